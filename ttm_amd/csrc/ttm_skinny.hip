// gfx950 kernels for skinny products in the TN view (w == 1, q fastest in
// the layout) with at most eight columns:
//
//     C[k][j] = sum_{l < nq} A[k*nq + l] * B(j, l),   k < outer <= 8,  j < m
//
// This is the reference's GEMV case (p = 1, include/tlib/detail/mtm.h:171,212)
// and its neighbours (case 6 with a handful of slabs).  B (m x nq) is the
// large operand, so the kernels stream it once with 16-byte loads and
// multiply on the vector ALUs (at most 16 flops per element of B: 32
// TFLOP/s of FMA at the HBM rate, against the 16x16 matrix cores that would
// waste 2 to 16 of every 16 columns); the rows of A are re-read per row block
// from L2.  The contraction is cut into `splits` ranges of kc indices (grid
// dimension y) so that a product with few rows still fills the chip; each
// range writes its partial C at c + s * c_split and k_sum_splits adds them in
// split order (no atomics: the result does not depend on scheduling).
#include "ttm_kernels.hpp"
#include "ttm_gemm.hpp"
#include "ttm_workspace.hpp"

#include <algorithm>
#include <cstdio>

namespace ttm_amd
{
namespace
{

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
// element-aligned 16-byte vectors (rows that do not start 16-byte aligned)
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));
typedef double f64x2u __attribute__((ext_vector_type(2), aligned(8)));

template <class T>
struct SV;
template <>
struct SV<float>
{
    static constexpr int VEC = 4;
    using vec = f32x4;
    using vec_u = f32x4u;
};
template <>
struct SV<double>
{
    static constexpr int VEC = 2;
    using vec = f64x2;
    using vec_u = f64x2u;
};

template <class V>
__device__ __forceinline__ V gld(const void *p)
{
    return *(const __attribute__((address_space(1))) V *)p;
}
template <class V>
__device__ __forceinline__ V gld_nt(const void *p)
{
    return __builtin_nontemporal_load((const __attribute__((address_space(1))) V *)p);
}

// a * b + c in T's own precision (__builtin_fma is the double one)
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <class T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Row-major B: B(j, l) at b[j * nq + l].  Block (jb, s) computes rows
// j0 = jb*R .. j0+R-1 (rows past m re-read row m-1 and are not stored) over
// the range [s*kc, s*kc + kc).  The 256 threads stride along l with one
// VEC-vector of each of the R rows of B and of the NC rows of A per step
// (U steps of loads in flight), accumulate R x NC sums, and reduce them over
// the block at the end.  The vectors are read at element alignment (rows of
// any length; one 16-byte load either way).
template <class T, int NC, int R, int U>
__global__ __launch_bounds__(256) void k_skinny_rows(const T *__restrict__ a, const T *__restrict__ b,
                                                     T *__restrict__ c, uint64_t outer, uint64_t nq, uint64_t m,
                                                     uint64_t kc, uint64_t c_split)
{
    constexpr int VEC = SV<T>::VEC;
    using V = typename SV<T>::vec;
    using VU = typename SV<T>::vec_u;
    __shared__ T red[4][R * NC];

    int const tid = threadIdx.x;
    uint64_t const j0 = (uint64_t)blockIdx.x * R;
    uint64_t const lb = (uint64_t)blockIdx.y * kc;
    uint64_t const le = nq - lb < kc ? nq : lb + kc;

    T const *brow[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        uint64_t const j = j0 + r < m ? j0 + r : m - 1;
        brow[r] = b + j * nq;
    }
    T const *arow[NC];
#pragma unroll
    for (int n = 0; n < NC; ++n) arow[n] = a + (uint64_t)(n < (int)outer ? n : (int)outer - 1) * nq;

    T acc[R][NC];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int n = 0; n < NC; ++n) acc[r][n] = 0;

    auto fma_step = [&](const V (&bv)[R], const V (&av)[NC]) {
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int n = 0; n < NC; ++n)
#pragma unroll
                for (int e = 0; e < VEC; ++e) acc[r][n] = fmadd(bv[r][e], av[n][e], acc[r][n]);
    };

    constexpr uint64_t STEP = 256 * VEC;
    uint64_t const nvec_end = lb + (le - lb) / VEC * VEC; // whole vectors end here
    uint64_t l = lb + (uint64_t)tid * VEC;
    // main loop: U steps whose vectors all lie before nvec_end
    for (; l + (U - 1) * STEP + VEC <= nvec_end; l += U * STEP) {
        V bv[U][R], av[U][NC];
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                bv[u][r] = gld_nt<VU>(brow[r] + l + u * STEP);
#pragma unroll
            for (int n = 0; n < NC; ++n) av[u][n] = gld<VU>(arow[n] + l + u * STEP);
        }
        // every load of the batch is issued before the first product (left
        // alone, the scheduler interleaves them and waits after two)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < U; ++u) fma_step(bv[u], av[u]);
    }
    // remaining whole vectors, one step at a time
    for (; l + VEC <= nvec_end; l += STEP) {
        V bv[R], av[NC];
#pragma unroll
        for (int r = 0; r < R; ++r) bv[r] = gld_nt<VU>(brow[r] + l);
#pragma unroll
        for (int n = 0; n < NC; ++n) av[n] = gld<VU>(arow[n] + l);
        __builtin_amdgcn_sched_barrier(0);
        fma_step(bv, av);
    }
    // the last nq % VEC indices of the product (last range only)
    if (tid < (int)(le - nvec_end)) {
        uint64_t const lt = nvec_end + tid;
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int n = 0; n < NC; ++n) acc[r][n] = fmadd(brow[r][lt], arow[n][lt], acc[r][n]);
    }

    int const lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int n = 0; n < NC; ++n) {
            T const v = wave_sum(acc[r][n]);
            if (lane == 0) red[wave][r * NC + n] = v;
        }
    __syncthreads();
    if (tid < R * NC) {
        int const r = tid / NC, n = tid % NC;
        uint64_t const j = j0 + r;
        if (j < m && (uint64_t)n < outer) {
            T const v = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
            c[(uint64_t)blockIdx.y * c_split + (uint64_t)n * m + j] = v;
        }
    }
}

// Column-major B: B(j, l) at b[j + l * m].  Thread t of a block owns the
// VEC rows j0 + VEC*(t % JV) + v (JV * VEC rows per block, JV a power of two
// <= 64, so that a wave covers whole columns of the block) and the indices
// l = lb + t / JV + (256 / JV) * step: one wave instruction reads 64 / JV
// consecutive columns of B's rows, one contiguous 1 KB piece when the block
// spans all m rows.  A(k, l) is the same for the JV lanes of a column (one
// address per column in the wave).  The NC x VEC sums of a thread are reduced
// over the threads that share its rows.
template <class T, int NC, int JV, int U>
__global__ __launch_bounds__(256) void k_skinny_cols(const T *__restrict__ a, const T *__restrict__ b,
                                                     T *__restrict__ c, uint64_t outer, uint64_t nq, uint64_t m,
                                                     uint64_t kc, uint64_t c_split)
{
    constexpr int VEC = SV<T>::VEC;
    using V = typename SV<T>::vec;
    using VU = typename SV<T>::vec_u;
    constexpr int LT = 256 / JV; // columns per step
    __shared__ T red[4][JV][VEC * NC];

    int const tid = threadIdx.x;
    int const tj = tid % JV, tl = tid / JV;
    uint64_t const jv = (uint64_t)blockIdx.x * (JV * VEC) + (uint64_t)tj * VEC; // first row of the thread
    int const nv = jv >= m ? 0 : (int)(m - jv < (uint64_t)VEC ? m - jv : VEC);  // rows of it inside B
    uint64_t const jr = nv == 0 ? 0 : jv;                                          // rows read (in bounds)
    uint64_t const lb = (uint64_t)blockIdx.y * kc;
    uint64_t const le = nq - lb < kc ? nq : lb + kc;

    T const *arow[NC];
#pragma unroll
    for (int n = 0; n < NC; ++n) arow[n] = a + (uint64_t)(n < (int)outer ? n : (int)outer - 1) * nq;

    T acc[VEC][NC];
#pragma unroll
    for (int v = 0; v < VEC; ++v)
#pragma unroll
        for (int n = 0; n < NC; ++n) acc[v][n] = 0;

    // the thread's VEC rows of column l: one element-aligned vector when
    // whole, element by element at the end of a column
    auto load_b = [&](uint64_t l) -> V {
        T const *p = b + l * m + jr;
        if (nv == VEC) return gld_nt<VU>(p);
        V x;
#pragma unroll
        for (int v = 0; v < VEC; ++v) x[v] = v < nv ? p[v] : T(0);
        return x;
    };
    auto fma_step = [&](V bv, const T (&av)[NC]) {
#pragma unroll
        for (int v = 0; v < VEC; ++v)
#pragma unroll
            for (int n = 0; n < NC; ++n) acc[v][n] = fmadd(bv[v], av[n], acc[v][n]);
    };

    uint64_t l = lb + tl;
    for (; l + (uint64_t)(U - 1) * LT < le; l += (uint64_t)U * LT) {
        V bv[U];
        T av[U][NC];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            bv[u] = load_b(l + (uint64_t)u * LT);
#pragma unroll
            for (int n = 0; n < NC; ++n) av[u][n] = arow[n][l + (uint64_t)u * LT];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < U; ++u) fma_step(bv[u], av[u]);
    }
    for (; l < le; l += LT) {
        V const bv = load_b(l);
        T av[NC];
#pragma unroll
        for (int n = 0; n < NC; ++n) av[n] = arow[n][l];
        fma_step(bv, av);
    }

    // reduce over the LW columns of a wave (lanes tj + JV*x), then the 4 waves
    int const wave = tid >> 6;
#pragma unroll
    for (int v = 0; v < VEC; ++v)
#pragma unroll
        for (int n = 0; n < NC; ++n) {
            T x = acc[v][n];
#pragma unroll
            for (int o = 32; o >= JV; o >>= 1) x += __shfl_xor(x, o, 64);
            if ((tid & 63) < JV) red[wave][tj][v * NC + n] = x;
        }
    __syncthreads();
    for (int idx = tid; idx < JV * VEC * NC; idx += 256) {
        int const t = idx / (VEC * NC), vn = idx % (VEC * NC);
        int const v = vn / NC, n = vn % NC;
        uint64_t const j = (uint64_t)blockIdx.x * (JV * VEC) + (uint64_t)t * VEC + v;
        if (j < m && (uint64_t)n < outer) {
            T const s = red[0][t][vn] + red[1][t][vn] + red[2][t][vn] + red[3][t][vn];
            c[(uint64_t)blockIdx.y * c_split + (uint64_t)n * m + j] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

int pick_nc(uint64_t outer) { return outer <= 1 ? 1 : (outer <= 4 ? 4 : 8); }

// Rows per block of the row-major kernel: R x NC accumulators per thread.
int pick_rows(uint64_t m, int nc)
{
    int const rmax = nc == 1 ? 16 : (nc == 4 ? 8 : 4);
    if (m <= 1) return 1;
    if (m <= 4) return 4;
    return rmax;
}

// Threads along j of the column-major kernel: enough VEC-vectors for m rows,
// at most 64 (a wave covers whole columns).
int pick_jv(uint64_t m, int vec)
{
    uint64_t const need = (m + vec - 1) / vec;
    int jv = 1;
    while ((uint64_t)jv < need && jv < 64) jv <<= 1;
    return jv;
}

bool rowmajor_b(const Fold &f) { return !f.b_colmajor || f.m == 1; }

struct SkinnyGeom
{
    int nc = 1, r = 1, jv = 1;
    bool rows = true;
    uint64_t nrb = 1; // row blocks
};

SkinnyGeom skinny_geom(const Fold &f, bool is_double)
{
    SkinnyGeom g;
    int const vec = is_double ? 2 : 4;
    g.nc = pick_nc(f.outer);
    g.rows = rowmajor_b(f);
    if (g.rows) {
        g.r = pick_rows(f.m, g.nc);
        g.nrb = (f.m + g.r - 1) / g.r;
    } else {
        g.jv = pick_jv(f.m, vec);
        g.nrb = (f.m + (uint64_t)g.jv * vec - 1) / ((uint64_t)g.jv * vec);
    }
    return g;
}

// Indices a block of the row-major kernel covers per step of its loop (all
// 256 threads, U vectors each), times the steps a range should at least take.
uint64_t rows_min_range(const SkinnyGeom &g, int vec)
{
    int const u = g.r >= 8 ? 1 : (g.r >= 4 ? 2 : 8);
    return std::max<uint64_t>(2048, (uint64_t)256 * vec * u * 2);
}

// Splits: about four blocks per compute unit, ranges long enough for a few
// rounds of the load pipeline (a multiple of 64), partials within 256 MiB
// (or the bound scratch).
Split skinny_split(const Fold &f, bool is_double, const SkinnyGeom &g, unsigned want)
{
    Split sp;
    sp.kc = f.nq;
    uint64_t s = want;
    if (s == 0) {
        uint64_t const target = (uint64_t)device_cus() * 4;
        s = g.nrb >= target ? 1 : (target + g.nrb - 1) / g.nrb;
    }
    uint64_t const by_len = f.nq / (g.rows ? rows_min_range(g, is_double ? 2 : 4) : 2048);
    uint64_t const room = std::min<uint64_t>(256ull << 20, scratch_available());
    uint64_t const by_bytes = room / (f.numel_c() * (is_double ? 8 : 4) + 1);
    s = std::min<uint64_t>({s, by_len, by_bytes, 65535});
    if (s <= 1) return sp;
    uint64_t kc = (f.nq + s - 1) / s;
    kc = (kc + 63) / 64 * 64;
    sp.kc = kc;
    sp.splits = (unsigned)((f.nq + kc - 1) / kc);
    if (sp.splits <= 1) sp = Split{1, f.nq};
    return sp;
}

template <class T, int NC, int R>
void launch_rows_r(const Fold &f, const T *a, const T *b, T *c, const Split &sp, uint64_t nrb, hipStream_t s)
{
    constexpr int U = R >= 8 ? 1 : (R >= 4 ? 2 : 8);
    hipLaunchKernelGGL((k_skinny_rows<T, NC, R, U>), dim3((unsigned)nrb, sp.splits), dim3(256), 0, s, a, b, c,
                       f.outer, f.nq, f.m, sp.kc, sp.splits > 1 ? f.numel_c() : 0);
}

template <class T, int NC>
void launch_rows(const Fold &f, const SkinnyGeom &g, const T *a, const T *b, T *c, const Split &sp, hipStream_t s)
{
    if (g.r == 1) launch_rows_r<T, NC, 1>(f, a, b, c, sp, g.nrb, s);
    else if (g.r == 4) launch_rows_r<T, NC, 4>(f, a, b, c, sp, g.nrb, s);
    else if constexpr (NC == 1) launch_rows_r<T, NC, 16>(f, a, b, c, sp, g.nrb, s);
    else if constexpr (NC == 4) launch_rows_r<T, NC, 8>(f, a, b, c, sp, g.nrb, s);
}

template <class T, int NC, int JV>
void launch_cols_jv(const Fold &f, const T *a, const T *b, T *c, const Split &sp, uint64_t nrb, hipStream_t s)
{
    hipLaunchKernelGGL((k_skinny_cols<T, NC, JV, 8>), dim3((unsigned)nrb, sp.splits), dim3(256), 0, s, a, b, c,
                       f.outer, f.nq, f.m, sp.kc, sp.splits > 1 ? f.numel_c() : 0);
}

template <class T, int NC>
void launch_cols(const Fold &f, const SkinnyGeom &g, const T *a, const T *b, T *c, const Split &sp, hipStream_t s)
{
    switch (g.jv) {
    case 1: launch_cols_jv<T, NC, 1>(f, a, b, c, sp, g.nrb, s); break;
    case 2: launch_cols_jv<T, NC, 2>(f, a, b, c, sp, g.nrb, s); break;
    case 4: launch_cols_jv<T, NC, 4>(f, a, b, c, sp, g.nrb, s); break;
    case 8: launch_cols_jv<T, NC, 8>(f, a, b, c, sp, g.nrb, s); break;
    case 16: launch_cols_jv<T, NC, 16>(f, a, b, c, sp, g.nrb, s); break;
    case 32: launch_cols_jv<T, NC, 32>(f, a, b, c, sp, g.nrb, s); break;
    default: launch_cols_jv<T, NC, 64>(f, a, b, c, sp, g.nrb, s); break;
    }
}

} // namespace

bool skinny_supports(const Fold &f)
{
    // TN view, at most eight columns, a contraction long enough to stream
    return f.w == 1 && f.outer <= 8 && f.nq >= 512;
}

std::string skinny_symbol(const Fold &f, bool is_double)
{
    SkinnyGeom const g = skinny_geom(f, is_double);
    char buf[96];
    char const *t = is_double ? "double" : "float";
    if (g.rows) {
        int const u = g.r >= 8 ? 1 : (g.r >= 4 ? 2 : 8);
        std::snprintf(buf, sizeof buf, "k_skinny_rows<%s, %d, %d, %d>", t, g.nc, g.r, u);
    } else {
        std::snprintf(buf, sizeof buf, "k_skinny_cols<%s, %d, %d, 8>", t, g.nc, g.jv);
    }
    return buf;
}

unsigned skinny_splits(const Fold &f, bool is_double, unsigned want)
{
    return skinny_split(f, is_double, skinny_geom(f, is_double), want).splits;
}

template <class T>
ttm_status_t launch_skinny(const Fold &f, const T *a, const T *b, T *c, hipStream_t s, unsigned want_splits)
{
    SkinnyGeom const g = skinny_geom(f, sizeof(T) == 8);
    Split const sp = skinny_split(f, sizeof(T) == 8, g, want_splits);
    T *out = c;
    void *ws = nullptr;
    if (sp.splits > 1) {
        size_t const bytes = (size_t)sp.splits * f.numel_c() * sizeof(T);
        if (ttm_status_t const st = ws_acquire(bytes, s, &ws, "workspace for the split-contraction partials");
            st != TTM_SUCCESS)
            return st;
        out = static_cast<T *>(ws);
    }
    if (g.rows) {
        if (g.nc == 1) launch_rows<T, 1>(f, g, a, b, out, sp, s);
        else if (g.nc == 4) launch_rows<T, 4>(f, g, a, b, out, sp, s);
        else launch_rows<T, 8>(f, g, a, b, out, sp, s);
    } else {
        if (g.nc == 1) launch_cols<T, 1>(f, g, a, b, out, sp, s);
        else if (g.nc == 4) launch_cols<T, 4>(f, g, a, b, out, sp, s);
        else launch_cols<T, 8>(f, g, a, b, out, sp, s);
    }
    if (sp.splits > 1) {
        launch_sum_splits<T>(out, f.numel_c(), sp.splits, c, s);
        ws_release(ws, s);
    }
    hipError_t const e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(TTM_ERR_HIP, hipGetErrorString(e));
        return TTM_ERR_HIP;
    }
    return TTM_SUCCESS;
}

template ttm_status_t launch_skinny<float>(const Fold &, const float *, const float *, float *, hipStream_t,
                                           unsigned);
template ttm_status_t launch_skinny<double>(const Fold &, const double *, const double *, double *, hipStream_t,
                                            unsigned);

} // namespace ttm_amd
