// gfx950 kernels for the folded product
//
//     C[k][j][i] = sum_{l < nq} B(j, l) * A[k][l][i],   B(j,l) = b[j*bsj + l*bsl]
//
// A slab k is the dense nq x w block at a + k*nq*w, C slab k the m x w block
// at c + k*m*w.
//
// MFMA conventions (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64, one block,
// wave64; lane = 16*lr + lc):
//   srcA (16 x 4):  lane holds A[lc][lr]
//   srcB (4 x 16):  lane holds B[lr][lc]
//   D    (16 x 16): accumulator element r of the lane holds D[row][lc] with
//                   row = 4*lr + r for f32 and row = 4*r + lr for f64
//                   (Traits<T>::acc_row; measured with a lane-layout probe)
// In every kernel here srcA is a fragment of the small matrix B (rows j) and
// srcB a fragment of the tensor A, so D rows are rows j of C.
#include "ttm_kernels.hpp"
#include "ttm_gemm.hpp"
#include "ttm_lowp.hpp"

#include <cstring>

namespace ttm_amd
{

namespace
{

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <class T>
struct Traits;
template <>
struct Traits<float>
{
    static constexpr int VEC = 4; // elements per 16-byte access
    using vec = f32x4;
    using acc = f32x4; // 16x16x4 accumulator fragment
    __device__ static constexpr int acc_row(int lr, int r) { return 4 * lr + r; }
};
template <>
struct Traits<double>
{
    static constexpr int VEC = 2;
    using vec = f64x2;
    using acc = f64x4;
    __device__ static constexpr int acc_row(int lr, int r) { return 4 * r + lr; }
};

template <class V>
__device__ __forceinline__ V splat_zero()
{
    V v;
#pragma unroll
    for (int e = 0; e < (int)(sizeof(V) / sizeof(v[0])); ++e) v[e] = 0;
    return v;
}

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f64x4 mfma4(double a, double b, f64x4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Four consecutive accumulator rows (r = 0..3) stored to 4 consecutive
// elements at p (16-byte aligned).
__device__ __forceinline__ void store_rows4(float *p, f32x4 v) { *reinterpret_cast<f32x4 *>(p) = v; }
__device__ __forceinline__ void store_rows4(double *p, f64x4 v)
{
    f64x2 lo, hi;
    lo[0] = v[0];
    lo[1] = v[1];
    hi[0] = v[2];
    hi[1] = v[3];
    reinterpret_cast<f64x2 *>(p)[0] = lo;
    reinterpret_cast<f64x2 *>(p)[1] = hi;
}

// ---------------------------------------------------------------------------
// generic: one thread per element of C, any shape and alignment.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_generic(const T *__restrict__ a, const T *__restrict__ b,
                                                 T *__restrict__ c, uint64_t outer, uint64_t nq,
                                                 uint64_t w, uint64_t m, uint64_t bsj, uint64_t bsl)
{
    uint64_t const total = outer * m * w;
    uint64_t const step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += step) {
        uint64_t const i = idx % w;
        uint64_t const kj = idx / w;
        uint64_t const j = kj % m;
        uint64_t const k = kj / m;
        T const *ap = a + k * nq * w + i;
        T const *bp = b + j * bsj;
        T s = 0;
        for (uint64_t l = 0; l < nq; ++l) s = __builtin_fma(bp[l * bsl], ap[l * w], s);
        c[idx] = s;
    }
}

// ---------------------------------------------------------------------------
// nn_stream: w % VEC == 0.  Fused column index n' = k*w + i over all slabs.
// Each wave owns a strip of 16*VEC columns; lane (lc, lr) loads the VEC
// consecutive columns n' = 16*VEC*strip + VEC*lc + v (one slab, since
// w % VEC == 0) of A row l0 + lr with one 16-byte load, and feeds element v to
// the MFMA of column sub-tile v.  The wave covers 16*JT rows j (grid.y blocks
// the rest).  A is read once per row block; B fragments come through L1/L2.
// ---------------------------------------------------------------------------
template <class T, int JT>
__global__ __launch_bounds__(256) void k_nn_stream(const T *__restrict__ a, const T *__restrict__ b,
                                                   T *__restrict__ c, uint64_t outer, uint64_t nq,
                                                   uint64_t w, uint64_t m, uint64_t bsj, uint64_t bsl)
{
    using Tr = Traits<T>;
    constexpr int VEC = Tr::VEC;
    using V = typename Tr::vec;
    using Acc = typename Tr::acc;
    constexpr int U = 4; // MFMA k-steps per unrolled iteration

    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    int const lc = lane & 15;
    int const lr = lane >> 4;

    uint64_t const ncols = outer * w;
    uint64_t const col = ((uint64_t)blockIdx.x * 4 + wave) * (16 * VEC) + (uint64_t)lc * VEC;
    bool const valid = col < ncols;
    uint64_t const k = valid ? col / w : 0;
    uint64_t const i = valid ? col - k * w : 0;
    T const *ap = a + k * nq * w + i;

    uint64_t const j0 = (uint64_t)blockIdx.y * (16 * JT);
    T const *bp[JT];
    bool jv[JT];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) {
        uint64_t const j = j0 + jt * 16 + lc;
        jv[jt] = j < m;
        bp[jt] = b + (jv[jt] ? j : 0) * bsj;
    }

    Acc acc[JT][VEC];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt)
#pragma unroll
        for (int v = 0; v < VEC; ++v) acc[jt][v] = splat_zero<Acc>();

    uint64_t l0 = 0;
    for (; l0 + 4 * U <= nq; l0 += 4 * U) {
        V av[U];
        T bv[U][JT];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            uint64_t const l = l0 + 4 * u + lr;
            av[u] = valid ? *reinterpret_cast<V const *>(ap + l * w) : splat_zero<V>();
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) bv[u][jt] = jv[jt] ? bp[jt][l * bsl] : T(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int jt = 0; jt < JT; ++jt)
#pragma unroll
                for (int v = 0; v < VEC; ++v) acc[jt][v] = mfma4(bv[u][jt], av[u][v], acc[jt][v]);
    }
    for (; l0 < nq; l0 += 4) {
        uint64_t const l = l0 + lr;
        bool const lv = l < nq;
        V av = (valid && lv) ? *reinterpret_cast<V const *>(ap + l * w) : splat_zero<V>();
        T bv[JT];
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) bv[jt] = (jv[jt] && lv) ? bp[jt][l * bsl] : T(0);
#pragma unroll
        for (int jt = 0; jt < JT; ++jt)
#pragma unroll
            for (int v = 0; v < VEC; ++v) acc[jt][v] = mfma4(bv[jt], av[v], acc[jt][v]);
    }

    if (!valid) return;
    T *cp = c + k * m * w + i;
#pragma unroll
    for (int jt = 0; jt < JT; ++jt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            uint64_t const j = j0 + jt * 16 + Tr::acc_row(lr, r);
            if (j < m) {
                V o;
#pragma unroll
                for (int v = 0; v < VEC; ++v) o[v] = acc[jt][v][r];
                *reinterpret_cast<V *>(cp + j * w) = o;
            }
        }
}

// ---------------------------------------------------------------------------
// tn_stream: w == 1, nq % VEC == 0.  C[k][j] = sum_l A[k*nq + l] * B(j, l).
// Each wave owns 16 rows k (srcB column lc <-> row k0 + lc) and 16*JT rows j.
// Lane (lc, lr) loads the VEC contraction indices l0 + VEC*lr + v of its row
// with one 16-byte load; MFMA step v pairs element v with the same l of B
// (the k-index of the MFMA is lr for both operands).  With BROW (row-major
// B) the B fragment is one 16-byte load as well.
// ---------------------------------------------------------------------------
template <class T, int JT, bool BROW>
__global__ __launch_bounds__(256) void k_tn_stream(const T *__restrict__ a, const T *__restrict__ b,
                                                   T *__restrict__ c, uint64_t outer, uint64_t nq,
                                                   uint64_t m, uint64_t bsj, uint64_t bsl)
{
    using Tr = Traits<T>;
    constexpr int VEC = Tr::VEC;
    using V = typename Tr::vec;
    using Acc = typename Tr::acc;
    constexpr int U = 2;
    constexpr int KSTEP = 4 * VEC; // contraction indices per wave-wide load

    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    int const lc = lane & 15;
    int const lr = lane >> 4;

    uint64_t const row = ((uint64_t)blockIdx.x * 4 + wave) * 16 + lc;
    bool const valid = row < outer;
    T const *ap = a + (valid ? row : 0) * nq + (uint64_t)lr * VEC;

    uint64_t const j0 = (uint64_t)blockIdx.y * (16 * JT);
    T const *bp[JT];
    bool jv[JT];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) {
        uint64_t const j = j0 + jt * 16 + lc;
        jv[jt] = j < m;
        bp[jt] = b + (jv[jt] ? j : 0) * bsj + (uint64_t)lr * VEC * bsl;
    }

    Acc acc[JT];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) acc[jt] = splat_zero<Acc>();

    auto load_b = [&](int jt, uint64_t l, bool ok) -> V {
        V r = splat_zero<V>();
        if (ok && jv[jt]) {
            if constexpr (BROW) {
                r = *reinterpret_cast<V const *>(bp[jt] + l);
            } else {
#pragma unroll
                for (int v = 0; v < VEC; ++v) r[v] = bp[jt][(l + v) * bsl];
            }
        }
        return r;
    };

    uint64_t l0 = 0;
    for (; l0 + KSTEP * U <= nq; l0 += KSTEP * U) {
        V av[U];
        V bv[U][JT];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            av[u] = valid ? *reinterpret_cast<V const *>(ap + l0 + KSTEP * u) : splat_zero<V>();
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) bv[u][jt] = load_b(jt, l0 + KSTEP * u, true);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int v = 0; v < VEC; ++v)
#pragma unroll
                for (int jt = 0; jt < JT; ++jt) acc[jt] = mfma4(bv[u][jt][v], av[u][v], acc[jt]);
    }
    for (; l0 < nq; l0 += KSTEP) {
        // nq % VEC == 0, so a vector is either wholly inside or wholly outside
        bool const lv = l0 + (uint64_t)lr * VEC < nq;
        V av = (valid && lv) ? *reinterpret_cast<V const *>(ap + l0) : splat_zero<V>();
        V bv[JT];
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) bv[jt] = load_b(jt, l0, lv);
#pragma unroll
        for (int v = 0; v < VEC; ++v)
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) acc[jt] = mfma4(bv[jt][v], av[v], acc[jt]);
    }

    if (!valid) return;
    T *cp = c + row * m;
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) {
        if constexpr (Tr::acc_row(1, 0) == 4) {
            // f32: the lane's four rows are consecutive in j
            uint64_t const jb = j0 + jt * 16 + lr * 4;
            if ((m & 3) == 0) {
                if (jb < m) store_rows4(cp + jb, acc[jt]);
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (jb + r < m) cp[jb + r] = acc[jt][r];
            }
        } else {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                uint64_t const j = j0 + jt * 16 + Tr::acc_row(lr, r);
                if (j < m) cp[j] = acc[jt][r];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

int pick_jt(uint64_t m) { return m <= 16 ? 1 : (m <= 32 ? 2 : 4); }

template <class T>
ttm_status_t launch_plain(const Fold &f, const T *a, const T *b, T *c, hipStream_t s, Kernel kern)
{
    uint64_t const bsj = f.b_stride_j(), bsl = f.b_stride_l();
    constexpr int VEC = Traits<T>::VEC;
    switch (kern) {
    case Kernel::generic: {
        uint64_t const total = f.numel_c();
        uint64_t blocks = (total + 255) / 256;
        if (blocks > (1u << 20)) blocks = 1u << 20;
        hipLaunchKernelGGL(k_generic<T>, dim3((unsigned)blocks), dim3(256), 0, s, a, b, c, f.outer, f.nq, f.w,
                           f.m, bsj, bsl);
        break;
    }
    case Kernel::nn_stream: {
        int const jt = pick_jt(f.m);
        uint64_t const ncols = f.outer * f.w;
        dim3 grid((unsigned)((ncols + 64 * VEC - 1) / (64 * VEC)), (unsigned)((f.m + 16 * jt - 1) / (16 * jt)));
        if (jt == 1)
            hipLaunchKernelGGL((k_nn_stream<T, 1>), grid, dim3(256), 0, s, a, b, c, f.outer, f.nq, f.w, f.m, bsj, bsl);
        else if (jt == 2)
            hipLaunchKernelGGL((k_nn_stream<T, 2>), grid, dim3(256), 0, s, a, b, c, f.outer, f.nq, f.w, f.m, bsj, bsl);
        else
            hipLaunchKernelGGL((k_nn_stream<T, 4>), grid, dim3(256), 0, s, a, b, c, f.outer, f.nq, f.w, f.m, bsj, bsl);
        break;
    }
    case Kernel::tn_stream: {
        int const jt = pick_jt(f.m);
        dim3 grid((unsigned)((f.outer + 63) / 64), (unsigned)((f.m + 16 * jt - 1) / (16 * jt)));
#define TTM_LAUNCH_TN(JT, BROW)                                                                          \
    hipLaunchKernelGGL((k_tn_stream<T, JT, BROW>), grid, dim3(256), 0, s, a, b, c, f.outer, f.nq, f.m, bsj, bsl)
        if (f.b_colmajor) {
            if (jt == 1) TTM_LAUNCH_TN(1, false);
            else if (jt == 2) TTM_LAUNCH_TN(2, false);
            else TTM_LAUNCH_TN(4, false);
        } else {
            if (jt == 1) TTM_LAUNCH_TN(1, true);
            else if (jt == 2) TTM_LAUNCH_TN(2, true);
            else TTM_LAUNCH_TN(4, true);
        }
#undef TTM_LAUNCH_TN
        break;
    }
    case Kernel::gemm: return launch_gemm<T>(f, a, b, c, s);
    default: set_error(TTM_ERR_ARGUMENT, "no kernel for this problem"); return TTM_ERR_ARGUMENT;
    }
    hipError_t const e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(TTM_ERR_HIP, hipGetErrorString(e));
        return TTM_ERR_HIP;
    }
    return TTM_SUCCESS;
}

const char *const k_kernel_names[] = {"none", "generic", "nn_stream", "tn_stream", "gemm", "lowp", "lowp_generic"};

} // namespace

const char *kernel_name(Kernel k)
{
    int const i = static_cast<int>(k);
    return (i >= 0 && i < static_cast<int>(Kernel::count)) ? k_kernel_names[i] : "none";
}

std::string kernel_symbol(Kernel k, const Fold &f, bool is_double, ttm_precision_t prec)
{
    std::string const t = is_double ? "double" : "float";
    std::string const jt = std::to_string(pick_jt(f.m));
    switch (k) {
    case Kernel::generic: return "k_generic<" + t + ">";
    case Kernel::nn_stream: return "k_nn_stream<" + t + ", " + jt + ">";
    case Kernel::tn_stream: return "k_tn_stream<" + t + ", " + jt + ", " + (f.b_colmajor ? "false" : "true") + ">";
    case Kernel::gemm: return gemm_symbol(f, is_double);
    case Kernel::lowp: return lowp_symbol(f, prec, false);
    case Kernel::lowp_generic: return lowp_symbol(f, prec, true);
    default: return "";
    }
}

Kernel kernel_from_name(const char *name)
{
    if (!name || !*name) return Kernel::none;
    for (int i = 1; i < static_cast<int>(Kernel::count); ++i)
        if (std::strcmp(name, k_kernel_names[i]) == 0) return static_cast<Kernel>(i);
    return Kernel::none;
}

Kernel choose_kernel(const Fold &f, bool is_double, ttm_precision_t prec, bool aligned, Kernel force)
{
    if (prec != TTM_PREC_DEFAULT) {
        if (is_double || prec > TTM_PREC_FP8_E4M3) return Kernel::none;
        bool const lp_ok = lowp_supports(f, aligned);
        if (force == Kernel::generic || force == Kernel::lowp_generic || !lp_ok) return Kernel::lowp_generic;
        return Kernel::lowp;
    }
    uint64_t const vec = is_double ? 2 : 4;
    // grid.y of the MFMA kernels covers m in blocks of >= 16 rows
    bool const grid_ok = (f.m + 15) / 16 <= 65535;
    bool const nn_ok = aligned && grid_ok && f.w % vec == 0;
    bool const tn_ok = aligned && grid_ok && f.w == 1 && f.nq % vec == 0;
    bool const gemm_ok = gemm_supports(f, is_double, aligned);
    if (force == Kernel::generic) return Kernel::generic;
    if (force == Kernel::nn_stream && nn_ok) return Kernel::nn_stream;
    if (force == Kernel::tn_stream && tn_ok) return Kernel::tn_stream;
    if (force == Kernel::gemm && gemm_ok) return Kernel::gemm;
    if (gemm_ok) return Kernel::gemm;
    if (tn_ok) return Kernel::tn_stream;
    if (nn_ok) return Kernel::nn_stream;
    return Kernel::generic;
}

ttm_status_t launch_f32(const Fold &f, const float *a, const float *b, float *c, ttm_precision_t prec,
                        hipStream_t stream, Kernel force, Kernel *used)
{
    bool const aligned = ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b) |
                           reinterpret_cast<uintptr_t>(c)) & 15u) == 0;
    Kernel const k = choose_kernel(f, false, prec, aligned, force);
    if (used) *used = k;
    if (k == Kernel::none) {
        set_error(TTM_ERR_PRECISION);
        return TTM_ERR_PRECISION;
    }
    if (k == Kernel::lowp || k == Kernel::lowp_generic)
        return launch_lowp(f, a, b, c, prec, stream, k == Kernel::lowp_generic);
    return launch_plain<float>(f, a, b, c, stream, k);
}

ttm_status_t launch_f64(const Fold &f, const double *a, const double *b, double *c, hipStream_t stream,
                        Kernel force, Kernel *used)
{
    bool const aligned = ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b) |
                           reinterpret_cast<uintptr_t>(c)) & 15u) == 0;
    Kernel const k = choose_kernel(f, true, TTM_PREC_DEFAULT, aligned, force);
    if (used) *used = k;
    return launch_plain<double>(f, a, b, c, stream, k);
}

} // namespace ttm_amd
