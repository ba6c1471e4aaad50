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
#include "ttm_tune.hpp"
#include "ttm_workspace.hpp"

#include <atomic>
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
        for (uint64_t l = 0; l < nq; ++l) {
            if constexpr (sizeof(T) == 4) s = __builtin_fmaf(bp[l * bsl], ap[l * w], s);
            else s = __builtin_fma(bp[l * bsl], ap[l * w], s);
        }
        c[idx] = s;
    }
}

// ---------------------------------------------------------------------------
// Streaming kernels for small m: A goes from HBM straight into MFMA operand
// registers (no LDS), B fragments come through L1/L2, and a software pipeline
// keeps D groups of A loads in flight per wave.  The wave covers every row j
// of C (16*JT >= m), so A is read exactly once: non-temporal loads and stores.
//
// Contraction indices are taken in groups; MFMA step s of a group pairs
// k-slot lr of both operands with one index l(s, lr), chosen so that each
// lane's loads are 16-byte vectors.
// ---------------------------------------------------------------------------

template <class V>
__device__ __forceinline__ V ld_nt(const V *p)
{
    return __builtin_nontemporal_load(p);
}

// VEC consecutive B values B(j, l .. l+VEC-1) of one row j.
template <class T, bool BROW>
__device__ __forceinline__ typename Traits<T>::vec load_bvec(const T *bj, uint64_t l, uint64_t bsl)
{
    using V = typename Traits<T>::vec;
    if constexpr (BROW) {
        return *reinterpret_cast<const V *>(bj + l);
    } else {
        V r;
#pragma unroll
        for (int e = 0; e < Traits<T>::VEC; ++e) r[e] = bj[(l + e) * bsl];
        return r;
    }
}

// Column-major B (B(j, l) at b[j + l*m]) copied row-major (bt[j*nq + l]) for
// the streams, which read 4 consecutive l of a row per lane.  One thread per
// element, writes coalesced (B is small: m <= 64 rows).
template <class T>
__global__ __launch_bounds__(256) void k_b_rows(const T *__restrict__ b, T *__restrict__ bt, uint64_t m,
                                                uint64_t nq)
{
    uint64_t const idx = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= m * nq) return;
    uint64_t const j = idx / nq, l = idx - j * nq;
    bt[idx] = b[j + l * m];
}

// nn_stream: w % VEC == 0.  Fused column index n = k*w + i over all slabs.
// A wave owns a strip of 16*VEC columns; lane (lc, lr) owns the VEC columns
// n = strip + VEC*lc + v (one slab, since w % VEC == 0), which feed column
// sub-tile v.  A group is 16 indices l0 .. l0+15; step u (< 4) uses
// l = l0 + 4*lr + u, so a lane loads A rows l0 + 4*lr + u (one 16-byte load
// each; one wave instruction covers 4 rows x 16*VEC columns) and the 4
// consecutive B values B(j, l0 + 4*lr .. l0 + 4*lr + 3).
template <class T, int JT, int D, bool BROW>
__global__ __launch_bounds__(256) void k_nn_stream(const T *__restrict__ a, const T *__restrict__ b,
                                                   T *__restrict__ c, uint64_t outer, uint64_t nq,
                                                   uint64_t w, uint64_t m, uint64_t bsj, uint64_t bsl)
{
    using Tr = Traits<T>;
    constexpr int VEC = Tr::VEC;
    using V = typename Tr::vec;
    using Acc = typename Tr::acc;
    constexpr int BV = 4 / VEC; // B vectors per group and row

    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    int const lc = lane & 15;
    int const lr = lane >> 4;

    uint64_t const ncols = outer * w;
    uint64_t const col = ((uint64_t)blockIdx.x * 4 + wave) * (16 * VEC) + (uint64_t)lc * VEC;
    bool const valid = col < ncols;
    // lanes past the end read column 0 (in bounds) and store nothing
    uint64_t const k = valid ? col / w : 0;
    uint64_t const i = valid ? col - k * w : 0;
    T const *ap = a + k * nq * w + i + (uint64_t)(4 * lr) * w;

    T const *bp[JT];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) {
        uint64_t const j = (uint64_t)jt * 16 + lc;
        bp[jt] = b + (j < m ? j : 0) * bsj; // rows past m are computed and dropped
    }

    Acc acc[JT][VEC];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt)
#pragma unroll
        for (int v = 0; v < VEC; ++v) acc[jt][v] = splat_zero<Acc>();

    V av[D][4];
    V bv[D][JT][BV];
    auto issue = [&](int slot, uint64_t g) {
        T const *p = ap + g * 16 * w;
#pragma unroll
        for (int u = 0; u < 4; ++u) av[slot][u] = ld_nt(reinterpret_cast<const V *>(p + u * w));
        uint64_t const l = g * 16 + 4 * lr;
#pragma unroll
        for (int jt = 0; jt < JT; ++jt)
#pragma unroll
            for (int h = 0; h < BV; ++h) bv[slot][jt][h] = load_bvec<T, BROW>(bp[jt], l + h * VEC, bsl);
    };
    auto compute = [&](int slot) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int jt = 0; jt < JT; ++jt)
#pragma unroll
                for (int v = 0; v < VEC; ++v)
                    acc[jt][v] = mfma4(bv[slot][jt][u / VEC][u % VEC], av[slot][u][v], acc[jt][v]);
    };

    uint64_t const ngroups = nq / 16;
#pragma unroll
    for (int d = 0; d < D - 1; ++d)
        if ((uint64_t)d < ngroups) issue(d, d);
    for (uint64_t g0 = 0; g0 < ngroups; g0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            uint64_t const g = g0 + d;
            if (g < ngroups) {
                if (g + D - 1 < ngroups) issue((d + D - 1) % D, g + D - 1);
                compute(d);
            }
        }
    }
    // tail: l0 = 16 * ngroups, indices past nq are zero in both operands
    uint64_t const l0 = ngroups * 16;
    if (l0 < nq) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            uint64_t const l = l0 + 4 * lr + u;
            bool const lv = l < nq;
            V const x = lv ? *reinterpret_cast<const V *>(ap + (l0 + u) * w) : splat_zero<V>();
            T bl[JT];
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) bl[jt] = lv ? bp[jt][l * bsl] : T(0);
#pragma unroll
            for (int jt = 0; jt < JT; ++jt)
#pragma unroll
                for (int v = 0; v < VEC; ++v) acc[jt][v] = mfma4(bl[jt], x[v], acc[jt][v]);
        }
    }

    if (!valid) return;
    T *cp = c + k * m * w + i;
#pragma unroll
    for (int jt = 0; jt < JT; ++jt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            uint64_t const j = (uint64_t)jt * 16 + Tr::acc_row(lr, r);
            if (j < m) {
                V o;
#pragma unroll
                for (int v = 0; v < VEC; ++v) o[v] = acc[jt][v][r];
                __builtin_nontemporal_store(o, reinterpret_cast<V *>(cp + j * w));
            }
        }
}

// tn_stream: w == 1, nq % VEC == 0.  C[k][j] = sum_l A[k*nq + l] * B(j, l).
// A wave owns 16 rows k (column lc of the MFMA <-> row k0 + lc).  A group is
// 16*VEC indices (256 bytes of each row): lane (lc, lr) loads, for v < 4,
// the VEC indices l0 + 4*VEC*v + VEC*lr + e (e < VEC) of its row with one
// 16-byte load, and step (v, e) uses them for k-slot lr; the matching B
// values are one 16-byte vector per v of a row-major B.
template <class T, int JT, int D, bool BROW>
__global__ __launch_bounds__(256) void k_tn_stream(const T *__restrict__ a, const T *__restrict__ b,
                                                   T *__restrict__ c, uint64_t outer, uint64_t nq,
                                                   uint64_t m, uint64_t bsj, uint64_t bsl)
{
    using Tr = Traits<T>;
    constexpr int VEC = Tr::VEC;
    using V = typename Tr::vec;
    using Acc = typename Tr::acc;
    constexpr int GL = 16 * VEC; // indices per group

    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    int const lc = lane & 15;
    int const lr = lane >> 4;

    uint64_t const row = ((uint64_t)blockIdx.x * 4 + wave) * 16 + lc;
    bool const valid = row < outer;
    T const *ap = a + (valid ? row : 0) * nq + (uint64_t)lr * VEC;

    T const *bp[JT];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) {
        uint64_t const j = (uint64_t)jt * 16 + lc;
        bp[jt] = b + (j < m ? j : 0) * bsj;
    }
    uint64_t const lane_l = (uint64_t)lr * VEC;

    Acc acc[JT];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) acc[jt] = splat_zero<Acc>();

    V av[D][4];
    V bv[D][JT][4];
    auto issue = [&](int slot, uint64_t g) {
#pragma unroll
        for (int v = 0; v < 4; ++v) av[slot][v] = ld_nt(reinterpret_cast<const V *>(ap + g * GL + 4 * VEC * v));
#pragma unroll
        for (int jt = 0; jt < JT; ++jt)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                bv[slot][jt][v] = load_bvec<T, BROW>(bp[jt], g * GL + 4 * VEC * v + lane_l, bsl);
    };
    auto compute = [&](int slot) {
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int e = 0; e < VEC; ++e)
#pragma unroll
                for (int jt = 0; jt < JT; ++jt) acc[jt] = mfma4(bv[slot][jt][v][e], av[slot][v][e], acc[jt]);
    };

    uint64_t const ngroups = nq / GL;
#pragma unroll
    for (int d = 0; d < D - 1; ++d)
        if ((uint64_t)d < ngroups) issue(d, d);
    for (uint64_t g0 = 0; g0 < ngroups; g0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            uint64_t const g = g0 + d;
            if (g < ngroups) {
                if (g + D - 1 < ngroups) issue((d + D - 1) % D, g + D - 1);
                compute(d);
            }
        }
    }
    // tail: nq % VEC == 0, so each lane's vector is wholly inside or outside
    for (uint64_t l0 = ngroups * GL; l0 < nq; l0 += 4 * VEC) {
        bool const lv = l0 + lane_l < nq;
        V const x = lv ? *reinterpret_cast<const V *>(ap + l0) : splat_zero<V>();
        V bl[JT];
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) bl[jt] = lv ? load_bvec<T, BROW>(bp[jt], l0 + lane_l, bsl) : splat_zero<V>();
#pragma unroll
        for (int e = 0; e < VEC; ++e)
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) acc[jt] = mfma4(bl[jt][e], x[e], acc[jt]);
    }

    if (!valid) return;
    T *cp = c + row * m;
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) {
        if constexpr (Tr::acc_row(1, 0) == 4) {
            // f32: the lane's four rows are consecutive in j
            uint64_t const jb = (uint64_t)jt * 16 + lr * 4;
            if ((m & 3) == 0) {
                if (jb < m) __builtin_nontemporal_store(acc[jt], reinterpret_cast<f32x4 *>(cp + jb));
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (jb + r < m) cp[jb + r] = acc[jt][r];
            }
        } else {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                uint64_t const j = (uint64_t)jt * 16 + Tr::acc_row(lr, r);
                if (j < m) cp[j] = acc[jt][r];
            }
        }
    }
}
// tn_lds: w == 1, m <= 16, nq % VEC == 0, all of B in LDS.
// C[k][j] = sum_l A[k*nq + l] * B(j, l).  Persistent: wave g of G takes the
// 16-row tiles g, g + G, ...; a tile's rows are walked in segments of 512
// bytes (32 pieces of 16 bytes).  A segment is read as eight wave loads of
// 2 rows x 512 B (lanes 0-31 row t, lanes 32-63 row t + 8, lane piece l & 31):
// whole 128-byte lines per load, the pattern that reads fastest here.  The
// wave writes the segment to its own 8 KB LDS buffer (row r's piece p at
// r*32 + ((p + r) & 31): each load's half is one contiguous 512 B store, and
// the 16 rows a read spans fall in 16 different bank groups) and reads it back
// by row: step s gives lane (lc, lr) piece 4s + lr of row lc, i.e. indices
// l = VEC*(32*seg + 4s + lr) + e, and MFMA (s, e) puts them in k-slot lr.  B is
// staged once per workgroup as rows of `spb` pieces (spb = 1 mod 16: the 16
// rows of a read fall in 16 different bank groups), zero past m and nq.  Only
// the wave that wrote a buffer reads it, and LDS operations of one wave
// complete in issue order, so the buffer needs no barrier.
template <class T, int WAVES, int D, int SPB_MAX>
__global__ __launch_bounds__(WAVES * 64) void k_tn_lds(const T *__restrict__ a, const T *__restrict__ b,
                                                       T *__restrict__ c, uint64_t outer, uint64_t nq, uint64_t m,
                                                       uint64_t bsj, uint64_t bsl, uint32_t spb)
{
    using Tr = Traits<T>;
    constexpr int VEC = Tr::VEC;
    using V = typename Tr::vec;
    using Acc = typename Tr::acc;
    constexpr int SEGP = 32; // pieces per segment

    __shared__ V bs[16 * SPB_MAX];
    __shared__ V as[WAVES][16 * SEGP];

    int const lane = threadIdx.x & 63;
    int const wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int const lc = lane & 15;
    int const lr = lane >> 4;

    // B rows 0..15, spb pieces each (zero past m and past nq).  A thread
    // issues up to 16 pieces' loads (16-byte vectors for a row-major B)
    // before its LDS stores, and zeroes at store time, so the loads overlap.
    // The row of piece idx is idx / spb, computed as a float reciprocal
    // multiply (exact for idx < 16 * spb, spb = 32k + 1 <= 513).
    {
        constexpr int NT = WAVES * 64;
        constexpr int KMAX = (16 * SPB_MAX + NT - 1) / NT;
        constexpr int KB = KMAX < 16 ? KMAX : 16;
        uint32_t const npieces = 16u * spb;
        float const inv = 1.0f / (float)spb;
        uint64_t const nvec_b = nq / VEC;
        auto fill = [&](auto rowvec_tag) {
            constexpr bool ROWVEC = decltype(rowvec_tag)::value;
            for (int k0 = 0; k0 < KMAX; k0 += KB) {
                V tmp[KB];
                bool ok[KB];
#pragma unroll
                for (int k = 0; k < KB; ++k) {
                    uint32_t const idx = threadIdx.x + (uint32_t)(k0 + k) * NT;
                    uint32_t const j = (uint32_t)(((float)idx + 0.5f) * inv);
                    uint32_t const p = idx - j * spb;
                    ok[k] = idx < npieces && j < m && p < nvec_b;
                    uint64_t const jj = ok[k] ? j : 0, pp = ok[k] ? p : 0; // in-bounds addresses for every lane
                    if constexpr (ROWVEC) {
                        tmp[k] = *reinterpret_cast<const V *>(b + jj * bsj + pp * VEC);
                    } else {
#pragma unroll
                        for (int e = 0; e < VEC; ++e) tmp[k][e] = b[jj * bsj + (pp * VEC + e) * bsl];
                    }
                }
#pragma unroll
                for (int k = 0; k < KB; ++k) {
                    uint32_t const idx = threadIdx.x + (uint32_t)(k0 + k) * NT;
                    if (idx < npieces) bs[idx] = ok[k] ? tmp[k] : splat_zero<V>();
                }
            }
        };
        if (bsl == 1 && bsj % VEC == 0 && (reinterpret_cast<uintptr_t>(b) & 15u) == 0) fill(std::true_type());
        else fill(std::false_type());
    }
    __syncthreads();

    uint64_t const nvec = nq / VEC; // pieces per row
    uint64_t const cpt = (nvec + SEGP - 1) / SEGP; // segments per row
    uint64_t const ntiles = (outer + 15) / 16;
    uint64_t const nw = (uint64_t)gridDim.x * WAVES;
    uint64_t const wid = (uint64_t)blockIdx.x * WAVES + wave;
    uint64_t const mytiles = wid < ntiles ? (ntiles - 1 - wid) / nw + 1 : 0;
    uint64_t const nch = mytiles * cpt;
    if (nch == 0) return;

    int const hrow = lane >> 5; // load: rows t (lanes 0-31) and t + 8 (lanes 32-63)
    int const piece = lane & 31;

    // Issue cursor.  Every step issues loads (the loop then has one shape and
    // its waits count only older segments); past the last segment of the wave
    // the cursor stays on it, so the extra loads repeat in-bounds addresses.
    V av[D][8];
    bool pv[D]; // the lane's piece is inside the row (zeroed when stored to LDS otherwise)
    bool rg[D]; // the slot holds a row's last, partial segment (uniform)
    bool const ragged = (nvec % SEGP) != 0;
    uint64_t i_tile = wid, i_seg = 0, i_count = 0;
    auto issue = [&](int slot) {
        uint64_t const p = i_seg * SEGP + piece;
        pv[slot] = p < nvec;
        rg[slot] = ragged && i_seg + 1 == cpt;
        uint64_t const pc = pv[slot] ? p : 0; // pieces past the row end load piece 0
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            uint64_t row = i_tile * 16 + t + 8 * hrow;
            row = row < outer ? row : outer - 1; // rows past the end read the last row; not stored
            av[slot][t] = ld_nt(reinterpret_cast<V const *>(a + row * nq) + pc);
        }
        if (++i_count < nch) {
            if (++i_seg == cpt) {
                i_seg = 0;
                i_tile += nw;
            }
        }
    };

    Acc acc[2] = {splat_zero<Acc>(), splat_zero<Acc>()}; // two chains: even and odd steps
    V *const buf = as[wave];
    uint64_t c_tile = wid, c_seg = 0; // compute cursor
    auto compute = [&](int slot) {
        if (rg[slot]) {
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                int const r = t + 8 * hrow;
                buf[r * SEGP + ((piece + r) & (SEGP - 1))] = pv[slot] ? av[slot][t] : splat_zero<V>();
            }
        } else {
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                int const r = t + 8 * hrow;
                buf[r * SEGP + ((piece + r) & (SEGP - 1))] = av[slot][t];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        V const *brow = bs + (uint32_t)lc * spb + (uint32_t)c_seg * SEGP;
#pragma unroll
        for (int s = 0; s < SEGP / 4; ++s) {
            int const q = 4 * s + lr;
            V const x = buf[lc * SEGP + ((q + lc) & (SEGP - 1))];
            V const y = brow[q];
#pragma unroll
            for (int e = 0; e < VEC; ++e) acc[s & 1] = mfma4(y[e], x[e], acc[s & 1]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (++c_seg == cpt) {
            uint64_t const row = c_tile * 16 + lc;
            Acc const sum = acc[0] + acc[1];
            if (row < outer) {
                T *cp = c + row * m;
                if constexpr (Tr::acc_row(1, 0) == 4) {
                    // f32: the lane's four values are rows 4*lr .. 4*lr + 3 of the product tile
                    uint64_t const jb = (uint64_t)lr * 4;
                    if ((m & 3) == 0) {
                        if (jb < m) __builtin_nontemporal_store(sum, reinterpret_cast<f32x4 *>(cp + jb));
                    } else {
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            if (jb + r < m) cp[jb + r] = sum[r];
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        uint64_t const j = (uint64_t)Tr::acc_row(lr, r);
                        if (j < m) cp[j] = sum[r];
                    }
                }
            }
            acc[0] = acc[1] = splat_zero<Acc>();
            c_seg = 0;
            c_tile += nw;
        }
    };

    // segment ch sits in slot ch % D
#pragma unroll
    for (int d = 0; d < D - 1; ++d) issue(d);
    uint64_t const nfull = nch / D * D;
    for (uint64_t ch = 0; ch < nfull; ch += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            issue((d + D - 1) % D);
            compute(d);
        }
    }
    // the last nch - nfull (< D) segments were issued by the loop or the prologue
#pragma unroll
    for (int d = 0; d < D - 1; ++d)
        if ((uint64_t)d < nch - nfull) compute(d);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

int pick_jt(uint64_t m) { return m <= 16 ? 1 : (m <= 32 ? 2 : 4); }

// Pipeline depth (groups of A loads in flight per wave) of the stream
// kernels; the register cost of a group grows with JT.  NN, JT = 1, measured
// on A 1024x1024x512 at m = 16 (fraction of 8 TB/s, float / double): depth 4
// 0.69 / 0.66 where q is the slowest mode (rows of A 2-4 MB apart), depth 6
// 0.76 / 0.72, depth 7 0.77 / 0.59 and depth 8 0.62 / 0.60 (occupancy falls).
constexpr int stream_depth(int jt, bool tn) { return jt == 1 ? (tn ? 3 : 6) : (jt == 2 ? (tn ? 2 : 3) : 2); }

// A column-major B of at most this many bytes is copied row-major for the
// NN stream (one small kernel; reading it in place takes VEC strided loads
// per lane and group: 0.60-0.73 of 8 TB/s at m = 16 against 0.79-0.86).
constexpr uint64_t kStreamBRowsMax = 16ull << 20;
bool stream_b_copy(const Fold &f, bool tn, int vec)
{
    return !tn && f.b_colmajor && f.m > 1 && f.nq % (uint64_t)vec == 0 &&
           f.m * f.nq * (vec == 4 ? 4 : 8) <= kStreamBRowsMax;
}
// 16-byte vectors of B rows: row-major B (or a single row, or the row-major
// copy) with nq % VEC == 0.
bool stream_bvec(const Fold &f, int vec, bool tn = false)
{
    return f.nq % (uint64_t)vec == 0 && (!f.b_colmajor || f.m == 1 || stream_b_copy(f, tn, vec));
}

// tn_lds instances: B rows of at most SPB_MAX pieces (nq <= VEC * (SPB_MAX - 1)),
// eight waves per workgroup while B and the eight 8 KB wave buffers fit the
// 160 KB of LDS, three waves for the 16 rows x 513 pieces of B (131 KB).
// m = 16 on A 1024x1024x512 (q fastest, n_q = 1024 / 512): float32 with 8
// waves 6.0 / 5.7 TB/s, with 4 waves 5.6 / 5.6 TB/s, pipeline depth 3 or 4 no
// better than 2 (the LDS-staged tile: 5.7 / 5.4 TB/s); float64 with 3 / 8
// waves 5.7 / 6.2 TB/s (the tile: 5.6 / 5.4 TB/s).
constexpr int kTnLdsDepth = 2;
uint32_t tn_lds_spb(const Fold &f, int vec) { return (uint32_t)(((f.nq / vec + 31) / 32) * 32 + 1); }
int tn_lds_spb_max(const Fold &f, bool is_double)
{
    uint32_t const spb = tn_lds_spb(f, is_double ? 2 : 4);
    if (spb <= 129) return 129;
    if (spb <= 257) return 257;
    return is_double && spb <= 513 ? 513 : 0;
}
constexpr int tn_lds_waves(int spb_max) { return spb_max <= 257 ? 8 : 3; }
bool tn_lds_ok(const Fold &f, bool is_double) { return f.m <= 16 && tn_lds_spb_max(f, is_double) > 0; }

template <class T, int SPB_MAX>
void launch_tn_lds(const Fold &f, const T *a, const T *b, T *c, hipStream_t s)
{
    constexpr int W = tn_lds_waves(SPB_MAX);
    auto const fn = k_tn_lds<T, W, kTnLdsDepth, SPB_MAX>;
    static int const occ = [&] {
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void *>(fn), W * 64, 0) !=
                hipSuccess ||
            blocks <= 0) {
            (void)hipGetLastError();
            blocks = 1;
        }
        return blocks;
    }();
    uint64_t const ntiles = (f.outer + 15) / 16;
    uint64_t const want = (ntiles + W - 1) / W;
    uint64_t const full = (uint64_t)device_cus() * occ;
    unsigned const grid = (unsigned)(want < full ? want : full);
    hipLaunchKernelGGL(fn, dim3(grid), dim3(W * 64), 0, s, a, b, c, f.outer, f.nq, f.m, f.b_stride_j(),
                       f.b_stride_l(), tn_lds_spb(f, Traits<T>::VEC));
}

template <class T, int JT, bool BROW>
void launch_stream(const Fold &f, const T *a, const T *b, T *c, hipStream_t s, bool tn)
{
    constexpr int VEC = Traits<T>::VEC;
    uint64_t const bsj = f.b_stride_j(), bsl = f.b_stride_l();
    constexpr bool is_double = std::is_same_v<T, double>;
    if (tn && JT == 1 && tn_lds_ok(f, is_double)) {
        int const smax = tn_lds_spb_max(f, is_double);
        if (smax == 129) launch_tn_lds<T, 129>(f, a, b, c, s);
        else if (smax == 257) launch_tn_lds<T, 257>(f, a, b, c, s);
        else if constexpr (is_double) launch_tn_lds<T, 513>(f, a, b, c, s);
    } else if (tn) {
        constexpr int D = stream_depth(JT, true);
        dim3 grid((unsigned)((f.outer + 63) / 64));
        hipLaunchKernelGGL((k_tn_stream<T, JT, D, BROW>), grid, dim3(256), 0, s, a, b, c, f.outer, f.nq, f.m, bsj,
                           bsl);
    } else {
        constexpr int D = stream_depth(JT, false);
        uint64_t const ncols = f.outer * f.w;
        dim3 grid((unsigned)((ncols + 64 * VEC - 1) / (64 * VEC)));
        hipLaunchKernelGGL((k_nn_stream<T, JT, D, BROW>), grid, dim3(256), 0, s, a, b, c, f.outer, f.nq, f.w, f.m,
                           bsj, bsl);
    }
}

template <class T>
ttm_status_t launch_plain(const Fold &f, const T *a, const T *b, T *c, hipStream_t s, Kernel kern, unsigned splits)
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
    case Kernel::nn_stream:
    case Kernel::tn_stream: {
        // the planner only picks these for m <= 64 (one row block)
        bool const tn = kern == Kernel::tn_stream;
        int const jt = pick_jt(f.m);
        bool bv = stream_bvec(f, VEC, tn);
        Fold g = f;
        void *ws = nullptr;
        if (stream_b_copy(f, tn, VEC)) {
            if (ws_acquire(f.m * f.nq * sizeof(T), s, &ws, "row-major copy of B") == TTM_SUCCESS) {
                uint64_t const n = f.m * f.nq;
                hipLaunchKernelGGL(k_b_rows<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, b,
                                   static_cast<T *>(ws), f.m, f.nq);
                b = static_cast<const T *>(ws);
                g.b_colmajor = false;
            } else {
                bv = false; // no room (the allocation-free entry): read B in place
            }
        } else if (f.b_colmajor && f.m == 1) {
            g.b_colmajor = false; // one row: B(0, l) at b[l] either way
        }
        if (jt == 1) bv ? launch_stream<T, 1, true>(g, a, b, c, s, tn) : launch_stream<T, 1, false>(g, a, b, c, s, tn);
        else if (jt == 2) bv ? launch_stream<T, 2, true>(g, a, b, c, s, tn) : launch_stream<T, 2, false>(g, a, b, c, s, tn);
        else bv ? launch_stream<T, 4, true>(g, a, b, c, s, tn) : launch_stream<T, 4, false>(g, a, b, c, s, tn);
        ws_release(ws, s);
        break;
    }
    case Kernel::gemm: return launch_gemm<T>(f, a, b, c, s, splits);
    case Kernel::skinny: return launch_skinny<T>(f, a, b, c, s, splits);
    default: set_error(TTM_ERR_ARGUMENT, "no kernel for this problem"); return TTM_ERR_ARGUMENT;
    }
    hipError_t const e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(TTM_ERR_HIP, hipGetErrorString(e));
        return TTM_ERR_HIP;
    }
    return TTM_SUCCESS;
}

const char *const k_kernel_names[] = {"none",  "generic", "nn_stream",    "tn_stream",
                                      "gemm",  "lowp",    "lowp_generic", "skinny"};
static_assert(sizeof(k_kernel_names) / sizeof(k_kernel_names[0]) == static_cast<int>(Kernel::count), "names");

} // namespace

const char *kernel_name(Kernel k)
{
    int const i = static_cast<int>(k);
    return (i >= 0 && i < static_cast<int>(Kernel::count)) ? k_kernel_names[i] : "none";
}

std::string kernel_symbol(Kernel k, const Fold &f, bool is_double, ttm_precision_t prec, bool aligned)
{
    std::string const t = is_double ? "double" : "float";
    std::string const jt = std::to_string(pick_jt(f.m));
    switch (k) {
    case Kernel::generic: return "k_generic<" + t + ">";
    case Kernel::nn_stream:
    case Kernel::tn_stream: {
        bool const tn = k == Kernel::tn_stream;
        int const j = pick_jt(f.m);
        if (tn && tn_lds_ok(f, is_double)) {
            int const smax = tn_lds_spb_max(f, is_double);
            return "k_tn_lds<" + t + ", " + std::to_string(tn_lds_waves(smax)) + ", " + std::to_string(kTnLdsDepth) +
                   ", " + std::to_string(smax) + ">";
        }
        return std::string(tn ? "k_tn_stream<" : "k_nn_stream<") + t + ", " + jt + ", " +
               std::to_string(stream_depth(j, tn)) + ", " +
               (stream_bvec(f, is_double ? 2 : 4, tn) ? "true" : "false") + ">";
    }
    case Kernel::gemm: return gemm_symbol(f, is_double, aligned);
    case Kernel::skinny: return skinny_symbol(f, is_double);
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
        if (is_double || prec > TTM_PREC_FP16X2) return Kernel::none;
        bool const lp_ok = lowp_supports(f, aligned);
        if (force == Kernel::generic || force == Kernel::lowp_generic || !lp_ok) return Kernel::lowp_generic;
        return Kernel::lowp;
    }
    uint64_t const vec = is_double ? 2 : 4;
    // the stream kernels hold every row j of C in one wave: m <= 64
    bool const small_m = f.m <= 64;
    bool const nn_ok = aligned && small_m && f.w % vec == 0;
    bool const tn_ok = aligned && small_m && f.w == 1 && f.nq % vec == 0;
    bool const skinny_ok = skinny_supports(f);
    bool const gemm_ok = gemm_supports(f, is_double);
    if (force == Kernel::generic) return Kernel::generic;
    if (force == Kernel::nn_stream && nn_ok) return Kernel::nn_stream;
    if (force == Kernel::tn_stream && tn_ok) return Kernel::tn_stream;
    if (force == Kernel::gemm && gemm_ok) return Kernel::gemm;
    if (force == Kernel::skinny && skinny_ok) return Kernel::skinny;
    // q fastest with at most eight slabs (the GEMV-like cases, p = 1
    // included): B is the large operand, streamed on the vector ALUs with
    // the contraction split over the chip
    if (skinny_ok) return Kernel::skinny;
    // The streams give a wave (nn) or a 16-row tile (tn) a whole strip of
    // the product: with fewer workgroups than compute units the blocks of
    // the gemm family, split along the contraction, fill the chip instead.
    int const cus = device_cus();
    bool const nn_full = (f.outer * f.w + 64 * vec - 1) / (64 * vec) >= (uint64_t)cus;
    bool const tn_full = (f.outer + 16 * 8 - 1) / (16 * 8) >= (uint64_t)cus;
    // m <= 16 in the NN view: the register-streaming kernel reads A at
    // 6.4-6.8 TB/s on the bench shapes, the LDS-staged tile at 5.3-5.8
    if (nn_ok && f.m <= 16 && nn_full) return Kernel::nn_stream;
    // m <= 16 in the TN view with all of B in LDS: tn_lds reads A at
    // 5.7-6.1 TB/s (float32) and 5.7-6.2 TB/s (float64) on the bench shapes,
    // the LDS-staged tile at 5.3-5.7 and 5.4-5.6
    if (tn_ok && tn_lds_ok(f, is_double) && tn_full) return Kernel::tn_stream;
    if (gemm_ok) return Kernel::gemm;
    if (tn_ok) return Kernel::tn_stream;
    if (nn_ok) return Kernel::nn_stream;
    return Kernel::generic;
}

unsigned launch_splits(Kernel k, const Fold &f, bool is_double, unsigned want)
{
    if (k != Kernel::gemm && k != Kernel::skinny) return 1;
    if (want == 0) want = tuned_splits(k, f, is_double);
    return k == Kernel::gemm ? gemm_split(f, is_double, want).splits : skinny_splits(f, is_double, want);
}

ttm_status_t launch_f32(const Fold &f, const float *a, const float *b, float *c, ttm_precision_t prec,
                        hipStream_t stream, Kernel force, LaunchInfo *info, unsigned want_splits)
{
    uintptr_t const bits = reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b) |
                           reinterpret_cast<uintptr_t>(c);
    bool const aligned = (bits & 15u) == 0;
    // pointers that are not even element-aligned only run element by element
    Kernel const k = (bits & 3u) ? (prec == TTM_PREC_DEFAULT ? Kernel::generic : Kernel::lowp_generic)
                                 : choose_kernel(f, false, prec, aligned, force);
    if (k == Kernel::none) {
        set_error(TTM_ERR_PRECISION);
        return TTM_ERR_PRECISION;
    }
    bool const ac_aligned = ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(c)) & 15u) == 0;
    if (k == Kernel::lowp || k == Kernel::lowp_generic) {
        if (info) *info = LaunchInfo{k, 1, ac_aligned};
        return launch_lowp(f, a, b, c, prec, stream, k == Kernel::lowp_generic);
    }
    if (want_splits == 0) want_splits = tuned_splits(k, f, false);
    if (info) *info = LaunchInfo{k, launch_splits(k, f, false, want_splits), ac_aligned};
    return launch_plain<float>(f, a, b, c, stream, k, want_splits);
}

ttm_status_t launch_f64(const Fold &f, const double *a, const double *b, double *c, hipStream_t stream,
                        Kernel force, LaunchInfo *info, unsigned want_splits)
{
    uintptr_t const bits = reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b) |
                           reinterpret_cast<uintptr_t>(c);
    bool const aligned = (bits & 15u) == 0;
    Kernel const k = (bits & 7u) ? Kernel::generic : choose_kernel(f, true, TTM_PREC_DEFAULT, aligned, force);
    if (want_splits == 0) want_splits = tuned_splits(k, f, true);
    bool const ac_aligned = ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(c)) & 15u) == 0;
    if (info) *info = LaunchInfo{k, launch_splits(k, f, true, want_splits), ac_aligned};
    return launch_plain<double>(f, a, b, c, stream, k, want_splits);
}

} // namespace ttm_amd
