// Grouped TT products with short contractions (float32 plain, BF16X3 and
// FP16X2, and float64), for the stages of batched multi-TTM chains:
//   C[j * ncols + n] = sum_{l < nq} B[j * nq + l] * A[n * nq + l],
// m <= 64 rows of B, 4 <= nq <= 256.
//
// A unit is 16 columns (16 rows of A, 16 * nq contiguous floats) and all m
// rows of C.  One wave computes a unit on the 16x16x4 f32 MFMA with the
// contraction taken in groups of 16: lane (c, q) = (lane % 16, lane / 16)
// reads the 16-byte piece A[n0 + c][16 g + 4 q .. 16 g + 4 q + 3] of every
// group g straight from global memory into registers (element-aligned when
// nq % 4 != 0: one global_load_dwordx4 either way), and MFMA x of the group
// takes element x of the piece, so the contraction index of MFMA x at lane
// quarter q is 16 g + 4 q + x in both operands.  Nothing of A passes through
// LDS: a wave issues the loads of its whole unit at once (at most 16 KB) and
// sixteen waves per compute unit (m <= 16; eight for taller B, whose
// accumulators and fragments need more registers) keep the HBM stream busy
// between the matrix work.  B (m x nq, rows padded with zeros to the row
// height and to a multiple of 16 columns) is staged in LDS once per problem
// with a row stride of 16 k + 4 floats, so the 16-byte fragment reads of the
// 16 rows of a fragment fall on distinct bank groups.  The last group of a
// row reads past the row end into the next row (masked to zero), and at the
// very end of A is read as A's last four elements and shifted into place.
//
// Workgroups are persistent over contiguous ranges of the launch's unit list
// (ranges of nearly equal cost, computed on the host), so a workgroup stages
// B only when its range crosses into the next problem; its waves take the
// units of each problem's part of the range round-robin.  The staging is
// bracketed by barriers that every wave of the workgroup reaches once per
// problem of the range.
#include "ttm_rows.hpp"
#include "ttm_plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace ttm_amd
{
namespace
{

typedef float f32x4 __attribute__((ext_vector_type(4)));
// 16-byte vectors at element alignment (a dword-aligned global_load_dwordx4)
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));

// Global accesses through descriptor pointers: the explicit address space
// keeps them global (not FLAT) instructions, which do not count against the
// LDS wait counter.
template <class V>
__device__ __forceinline__ V ldg(const void *p)
{
    return *(const __attribute__((address_space(1))) V *)p;
}
template <class V>
__device__ __forceinline__ void stg(void *p, V v)
{
    *(__attribute__((address_space(1))) V *)p = v;
}

constexpr int kGroups = (int)(kRowsMaxNq / 16); // contraction groups of 16 per row, at most
constexpr int kStrideMax = (int)kRowsMaxNq + 4; // LDS row stride of B for nq = 256

// LDS row stride of B for a contraction of nq: 16 * ceil(nq / 16) + 4 (a
// multiple of 4 whose quarter is odd).
__host__ __device__ constexpr unsigned rows_stride(unsigned nq) { return (nq + 15) / 16 * 16 + 4; }

// Element x of the piece that starts at element e of A (clamped to A's last
// four elements, shift d): elements past the row end (x >= lim) are zero.
__device__ __forceinline__ f32x4 piece_last(f32x4 rl, int d, int lim)
{
    f32x4 af;
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        float t = rl[x];
        if (d == 1) t = x + 1 < 4 ? rl[(x + 1) & 3] : 0.f;
        if (d == 2) t = x + 2 < 4 ? rl[(x + 2) & 3] : 0.f;
        if (d == 3) t = x + 3 < 4 ? rl[(x + 3) & 3] : 0.f;
        af[x] = x < lim ? t : 0.f;
    }
    return af;
}

// The launch's schedule: workgroup b runs the units [wg_units[b],
// wg_units[b + 1]) (contiguous, of nearly equal estimated cost, starting in
// problem wg_prob[b]) as blocks cut at problem ends; B is staged at the start
// of each block, and the waves take the block's units round-robin.
template <int W, class Stage, class Unit>
__device__ __forceinline__ void rows_run(const RowsProblem *__restrict__ probs, const RowsSched &sc, Stage &&stage,
                                         Unit &&unit)
{
    unsigned const wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    unsigned sb = sc.wg_units[blockIdx.x];           // cursor (uniform)
    unsigned const u1 = sc.wg_units[blockIdx.x + 1];
    unsigned q = sc.wg_prob[blockIdx.x];             // problem of the cursor
    while (sb < u1) {
        while (sb >= probs[q].unit_begin + probs[q].nunits) ++q;
        RowsProblem const P = probs[q];
        unsigned const pend = P.unit_begin + P.nunits;
        unsigned const be = pend < u1 ? pend : u1;
        __syncthreads(); // every wave is done with the previous problem's B
        stage(P);
        __syncthreads();
        for (unsigned v = sb + wave; v < be; v += W) unit(P, v);
        sb = be;
    }
}

template <int MB, int W, int WPE>
__global__ __launch_bounds__(64 * W, WPE) void k_rows_grouped(const RowsProblem *__restrict__ probs, RowsSched sc)
{
    __shared__ __attribute__((aligned(16))) float sB[MB * kStrideMax];
    constexpr int JT = MB / 16;
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    int const lc = lane & 15, lq = lane >> 4;
    // B of a problem into LDS: rows past m and columns past nq are zero.
    // Wave w stages rows w, w + W, ...; a lane's loads are all issued before
    // its stores (one round trip per problem).
    auto stage = [&](const RowsProblem &P) {
        static_assert(MB % W == 0, "whole rows per wave");
        constexpr int RPW = MB / W;          // rows per wave
        constexpr int CPL = kStrideMax / 64; // columns per lane, at most
        unsigned const nq = P.nq, S = rows_stride(nq), width = (nq + 15) / 16 * 16;
        float t[RPW][CPL];
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                unsigned const j = wave + W * r, l = lane + 64 * c;
                t[r][c] = (j < P.m && l < nq) ? ldg<float>(static_cast<const float *>(P.b) + (uint64_t)j * nq + l) : 0.f;
            }
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                unsigned const j = wave + W * r, l = lane + 64 * c;
                if (l < width) sB[j * S + l] = t[r][c];
            }
    };
    auto unit = [&](const RowsProblem &P, unsigned v) {
        unsigned const nq = P.nq, S = rows_stride(nq), ng = (nq + 15) / 16;
        uint64_t const total = P.ncols * nq;
        const float *brow = sB + lc * S + 4 * lq;
        int const lim = (int)nq - (int)(16 * (ng - 1) + 4 * lq); // valid elements of the last group's piece
        uint64_t const col0 = (uint64_t)(v - P.unit_begin) * 16u;
        // columns past the last one read the last one (not stored)
        uint64_t const n = col0 + lc < P.ncols ? col0 + lc : P.ncols - 1;
        const float *a = static_cast<const float *>(P.a);
        const float *arow = a + n * nq + 4 * lq;
        f32x4 ra[kGroups];
#pragma unroll
        for (int g = 0; g < kGroups; ++g)
            if (g + 1 < (int)ng) ra[g] = ldg<f32x4u>(arow + 16 * g);
        // the last group: clamped to A's last four elements (shift d)
        uint64_t const e = n * nq + 16 * (ng - 1) + 4 * lq;
        uint64_t const ec = e + 4 <= total ? e : total - 4;
        int const d = (int)(e - ec);
        f32x4 const rl = ldg<f32x4u>(a + ec);

        f32x4 acc[JT];
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) acc[jt] = f32x4{0.f, 0.f, 0.f, 0.f};
        auto group = [&](unsigned g, f32x4 af) {
            f32x4 bf[JT];
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) bf[jt] = *reinterpret_cast<const f32x4 *>(brow + 16 * jt * S + 16 * g);
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int jt = 0; jt < JT; ++jt)
                    acc[jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(bf[jt][x], af[x], acc[jt], 0, 0, 0);
        };
#pragma unroll
        for (int g = 0; g < kGroups; ++g)
            if (g + 1 < (int)ng) group(g, ra[g]);
        // element x of the piece is rl[x + d]; elements past the row end
        // (x >= lim, which covers everything past A's end) are zero
        group(ng - 1, piece_last(rl, d, lim));
        // lane (c, q) holds C rows 16 jt + 4 q + r of column col0 + c
        if (col0 + lc < P.ncols) {
#pragma unroll
            for (int jt = 0; jt < JT; ++jt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    unsigned const j = 16 * jt + 4 * lq + r;
                    if (j < P.m) stg<float>(static_cast<float *>(P.c) + (uint64_t)j * P.ncols + col0 + lc, acc[jt][r]);
                }
        }
    };
    rows_run<W>(probs, sc, stage, unit);
}

// ---------------------------------------------------------------------------
// BF16X3: the same units with A and B split into three bfloat16 terms each
// (x = x0 + x1 + x2 + O(2^-24 |x|), every difference exact in float32) and
// the six largest cross products on the 16x16x32 bf16 MFMA, smallest first:
// (1,1), (0,2), (2,0), (0,1), (1,0), (0,0) (A term, B term).  A group is 32
// contraction indices: lane (c, q) reads A[n0 + c][32 g + 8 q .. + 7] as two
// 16-byte pieces and splits them in registers; MFMA operand element e of
// lane quarter q is contraction index 32 g + 8 q + e for both operands.  B's
// three planes (bf16, row stride 32 k + 8) are split once per problem when
// they are staged.
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

constexpr int kGroups32 = (int)(kRowsMaxNq / 32);
constexpr int kStrideMaxH = (int)kRowsMaxNq + 8; // bf16 elements
__host__ __device__ constexpr unsigned rows_stride_h(unsigned nq) { return (nq + 31) / 32 * 32 + 8; }

__device__ __forceinline__ void split3(f32x2 x, bf16x2 &p0, bf16x2 &p1, bf16x2 &p2)
{
    p0 = __builtin_convertvector(x, bf16x2);
    f32x2 const r1 = x - __builtin_convertvector(p0, f32x2);
    p1 = __builtin_convertvector(r1, bf16x2);
    f32x2 const r2 = r1 - __builtin_convertvector(p1, f32x2);
    p2 = __builtin_convertvector(r2, bf16x2);
}

// The three terms of eight floats (two 16-byte pieces) as bf16x8 operands.
__device__ __forceinline__ void split3x8(f32x4 lo, f32x4 hi, bf16x8 (&t)[3])
{
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        f32x2 const x = h < 2 ? f32x2{lo[2 * h], lo[2 * h + 1]} : f32x2{hi[2 * h - 4], hi[2 * h - 3]};
        bf16x2 p0, p1, p2;
        split3(x, p0, p1, p2);
        t[0][2 * h] = p0[0];
        t[0][2 * h + 1] = p0[1];
        t[1][2 * h] = p1[0];
        t[1][2 * h + 1] = p1[1];
        t[2][2 * h] = p2[0];
        t[2][2 * h + 1] = p2[1];
    }
}

template <int MB, int W, int WPE>
__global__ __launch_bounds__(64 * W, WPE) void k_rows_bf16x3(const RowsProblem *__restrict__ probs, RowsSched sc)
{
    __shared__ __attribute__((aligned(16))) __bf16 sB[3][MB * kStrideMaxH];
    constexpr int JT = MB / 16;
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    int const lc = lane & 15, lq = lane >> 4;
    // B's three planes: rows w, w + W, ... of wave w; a lane takes pairs of
    // columns 2 (lane + 64 c), all loads before the stores
    auto stage = [&](const RowsProblem &P) {
        static_assert(MB % W == 0, "whole rows per wave");
        constexpr int RPW = MB / W;
        constexpr int CPL = (kRowsMaxNq / 2 + 63) / 64; // column pairs per lane, at most
        unsigned const nq = P.nq, S = rows_stride_h(nq), width = (nq + 31) / 32 * 32;
        f32x2 t[RPW][CPL];
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                unsigned const j = wave + W * r, l = 2 * (lane + 64 * c);
                const float *bj = static_cast<const float *>(P.b) + (uint64_t)j * nq;
                t[r][c][0] = (j < P.m && l < nq) ? ldg<float>(bj + l) : 0.f;
                t[r][c][1] = (j < P.m && l + 1 < nq) ? ldg<float>(bj + l + 1) : 0.f;
            }
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                unsigned const j = wave + W * r, l = 2 * (lane + 64 * c);
                if (l < width) {
                    bf16x2 p0, p1, p2;
                    split3(t[r][c], p0, p1, p2);
                    *reinterpret_cast<bf16x2 *>(&sB[0][j * S + l]) = p0;
                    *reinterpret_cast<bf16x2 *>(&sB[1][j * S + l]) = p1;
                    *reinterpret_cast<bf16x2 *>(&sB[2][j * S + l]) = p2;
                }
            }
    };
    auto unit = [&](const RowsProblem &P, unsigned v) {
        unsigned const nq = P.nq, S = rows_stride_h(nq), ng = (nq + 31) / 32;
        uint64_t const total = P.ncols * nq;
        unsigned const boff = lc * S + 8 * lq;
        int const lim0 = (int)nq - (int)(32 * (ng - 1) + 8 * lq); // valid elements of the last group's first piece
        uint64_t const col0 = (uint64_t)(v - P.unit_begin) * 16u;
        uint64_t const n = col0 + lc < P.ncols ? col0 + lc : P.ncols - 1;
        const float *a = static_cast<const float *>(P.a);
        const float *arow = a + n * nq + 8 * lq;
        f32x4 ra[kGroups32][2];
#pragma unroll
        for (int g = 0; g < kGroups32; ++g)
            if (g + 1 < (int)ng) {
                ra[g][0] = ldg<f32x4u>(arow + 32 * g);
                ra[g][1] = ldg<f32x4u>(arow + 32 * g + 4);
            }
        uint64_t const e0 = n * nq + 32 * (ng - 1) + 8 * lq, e1 = e0 + 4;
        uint64_t const ec0 = e0 + 4 <= total ? e0 : total - 4, ec1 = e1 + 4 <= total ? e1 : total - 4;
        int const d0 = (int)(e0 - ec0), d1 = (int)(e1 - ec1);
        f32x4 const rl0 = ldg<f32x4u>(a + ec0), rl1 = ldg<f32x4u>(a + ec1);

        f32x4 acc[JT];
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) acc[jt] = f32x4{0.f, 0.f, 0.f, 0.f};
        auto group = [&](unsigned g, f32x4 lo, f32x4 hi) {
            bf16x8 ta[3];
            split3x8(lo, hi, ta);
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) {
                unsigned const o = boff + 16 * jt * S + 32 * g;
                bf16x8 const b0 = *reinterpret_cast<const bf16x8 *>(&sB[0][o]);
                bf16x8 const b1 = *reinterpret_cast<const bf16x8 *>(&sB[1][o]);
                bf16x8 const b2 = *reinterpret_cast<const bf16x8 *>(&sB[2][o]);
                f32x4 a = acc[jt];
                a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b1, ta[1], a, 0, 0, 0);
                a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b2, ta[0], a, 0, 0, 0);
                a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b0, ta[2], a, 0, 0, 0);
                a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b1, ta[0], a, 0, 0, 0);
                a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b0, ta[1], a, 0, 0, 0);
                a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b0, ta[0], a, 0, 0, 0);
                acc[jt] = a;
            }
        };
#pragma unroll
        for (int g = 0; g < kGroups32; ++g)
            if (g + 1 < (int)ng) group(g, ra[g][0], ra[g][1]);
        group(ng - 1, piece_last(rl0, d0, lim0), piece_last(rl1, d1, lim0 - 4));
        if (col0 + lc < P.ncols) {
#pragma unroll
            for (int jt = 0; jt < JT; ++jt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    unsigned const j = 16 * jt + 4 * lq + r;
                    if (j < P.m) stg<float>(static_cast<float *>(P.c) + (uint64_t)j * P.ncols + col0 + lc, acc[jt][r]);
                }
        }
    };
    rows_run<W>(probs, sc, stage, unit);
}

// ---------------------------------------------------------------------------
// FP16X2: the BF16X3 units with A and B split into two binary16 terms after
// scaling by powers of two (x 2^e = x0 + x1 + O(2^-22 |x 2^e|), the
// difference exact in float32) and three products on the 16x16x32 f16 MFMA,
// smallest first: (1,0), (0,1), (0,0).  The exponent e = 15 - E of an
// operand's largest finite magnitude f 2^E (f in [0.5, 1)) keeps its scaled
// values below 2^15; A's exponent is taken per unit (the largest finite
// magnitude among the unit's 16 rows, a wave reduction), B's per problem (a
// workgroup reduction when it is staged), and the accumulators are scaled
// back by 2^-(ea + eb).  Infinities and NaNs set no scale and pass through
// the split.
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ int fp16x2_exp(float amax)
{
    if (!(amax > 0.f) || !__builtin_isfinite(amax)) return 0;
    int e = 0;
    (void)frexpf(amax, &e);
    return 15 - e;
}

__device__ __forceinline__ float finite_abs(float x) { return __builtin_isfinite(x) ? fabsf(x) : 0.f; }

__device__ __forceinline__ void split2h(f32x2 x, f16x2 &p0, f16x2 &p1)
{
    p0 = __builtin_convertvector(x, f16x2);
    p1 = __builtin_convertvector(x - __builtin_convertvector(p0, f32x2), f16x2);
}

// The two terms of eight floats scaled by 2^e, as f16x8 operands.
__device__ __forceinline__ void split2x8(f32x4 lo, f32x4 hi, int e, f16x8 (&t)[2])
{
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        f32x2 x = h < 2 ? f32x2{lo[2 * h], lo[2 * h + 1]} : f32x2{hi[2 * h - 4], hi[2 * h - 3]};
        x = f32x2{ldexpf(x[0], e), ldexpf(x[1], e)};
        f16x2 p0, p1;
        split2h(x, p0, p1);
        t[0][2 * h] = p0[0];
        t[0][2 * h + 1] = p0[1];
        t[1][2 * h] = p1[0];
        t[1][2 * h + 1] = p1[1];
    }
}

template <int MB, int W, int WPE>
__global__ __launch_bounds__(64 * W, WPE) void k_rows_fp16x2(const RowsProblem *__restrict__ probs, RowsSched sc)
{
    __shared__ __attribute__((aligned(16))) _Float16 sB[2][MB * kStrideMaxH];
    __shared__ float s_red[W];
    constexpr int JT = MB / 16;
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    int const lc = lane & 15, lq = lane >> 4;
    int eb = 0; // B's exponent of the problem in LDS
    auto wave_max = [](float v) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
        return v;
    };
    auto stage = [&](const RowsProblem &P) {
        static_assert(MB % W == 0, "whole rows per wave");
        constexpr int RPW = MB / W;
        constexpr int CPL = (kRowsMaxNq / 2 + 63) / 64; // column pairs per lane, at most
        unsigned const nq = P.nq, S = rows_stride_h(nq), width = (nq + 31) / 32 * 32;
        f32x2 t[RPW][CPL];
        float amax = 0.f;
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                unsigned const j = wave + W * r, l = 2 * (lane + 64 * c);
                const float *bj = static_cast<const float *>(P.b) + (uint64_t)j * nq;
                t[r][c][0] = (j < P.m && l < nq) ? ldg<float>(bj + l) : 0.f;
                t[r][c][1] = (j < P.m && l + 1 < nq) ? ldg<float>(bj + l + 1) : 0.f;
                amax = fmaxf(amax, fmaxf(finite_abs(t[r][c][0]), finite_abs(t[r][c][1])));
            }
        amax = wave_max(amax);
        if (lane == 0) s_red[wave] = amax;
        __syncthreads();
        amax = 0.f;
#pragma unroll
        for (int w = 0; w < W; ++w) amax = fmaxf(amax, s_red[w]);
        eb = fp16x2_exp(amax);
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                unsigned const j = wave + W * r, l = 2 * (lane + 64 * c);
                if (l < width) {
                    f16x2 p0, p1;
                    split2h(f32x2{ldexpf(t[r][c][0], eb), ldexpf(t[r][c][1], eb)}, p0, p1);
                    *reinterpret_cast<f16x2 *>(&sB[0][j * S + l]) = p0;
                    *reinterpret_cast<f16x2 *>(&sB[1][j * S + l]) = p1;
                }
            }
    };
    auto unit = [&](const RowsProblem &P, unsigned v) {
        unsigned const nq = P.nq, S = rows_stride_h(nq), ng = (nq + 31) / 32;
        uint64_t const total = P.ncols * nq;
        unsigned const boff = lc * S + 8 * lq;
        int const lim0 = (int)nq - (int)(32 * (ng - 1) + 8 * lq);
        uint64_t const col0 = (uint64_t)(v - P.unit_begin) * 16u;
        uint64_t const n = col0 + lc < P.ncols ? col0 + lc : P.ncols - 1;
        const float *a = static_cast<const float *>(P.a);
        const float *arow = a + n * nq + 8 * lq;
        f32x4 ra[kGroups32][2];
#pragma unroll
        for (int g = 0; g < kGroups32; ++g)
            if (g + 1 < (int)ng) {
                ra[g][0] = ldg<f32x4u>(arow + 32 * g);
                ra[g][1] = ldg<f32x4u>(arow + 32 * g + 4);
            }
        uint64_t const e0 = n * nq + 32 * (ng - 1) + 8 * lq, e1 = e0 + 4;
        uint64_t const ec0 = e0 + 4 <= total ? e0 : total - 4, ec1 = e1 + 4 <= total ? e1 : total - 4;
        int const d0 = (int)(e0 - ec0), d1 = (int)(e1 - ec1);
        f32x4 const last0 = piece_last(ldg<f32x4u>(a + ec0), d0, lim0);
        f32x4 const last1 = piece_last(ldg<f32x4u>(a + ec1), d1, lim0 - 4);
        // A's exponent: the largest finite magnitude of the unit
        float amax = 0.f;
#pragma unroll
        for (int g = 0; g < kGroups32; ++g)
            if (g + 1 < (int)ng)
#pragma unroll
                for (int x = 0; x < 4; ++x) amax = fmaxf(amax, fmaxf(finite_abs(ra[g][0][x]), finite_abs(ra[g][1][x])));
#pragma unroll
        for (int x = 0; x < 4; ++x) amax = fmaxf(amax, fmaxf(finite_abs(last0[x]), finite_abs(last1[x])));
        int const ea = fp16x2_exp(wave_max(amax));

        f32x4 acc[JT];
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) acc[jt] = f32x4{0.f, 0.f, 0.f, 0.f};
        auto group = [&](unsigned g, f32x4 lo, f32x4 hi) {
            f16x8 ta[2];
            split2x8(lo, hi, ea, ta);
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) {
                unsigned const o = boff + 16 * jt * S + 32 * g;
                f16x8 const b0 = *reinterpret_cast<const f16x8 *>(&sB[0][o]);
                f16x8 const b1 = *reinterpret_cast<const f16x8 *>(&sB[1][o]);
                f32x4 t = acc[jt];
                t = __builtin_amdgcn_mfma_f32_16x16x32_f16(b0, ta[1], t, 0, 0, 0);
                t = __builtin_amdgcn_mfma_f32_16x16x32_f16(b1, ta[0], t, 0, 0, 0);
                t = __builtin_amdgcn_mfma_f32_16x16x32_f16(b0, ta[0], t, 0, 0, 0);
                acc[jt] = t;
            }
        };
#pragma unroll
        for (int g = 0; g < kGroups32; ++g)
            if (g + 1 < (int)ng) group(g, ra[g][0], ra[g][1]);
        group(ng - 1, last0, last1);
        int const back = -(ea + eb);
        if (col0 + lc < P.ncols) {
#pragma unroll
            for (int jt = 0; jt < JT; ++jt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    unsigned const j = 16 * jt + 4 * lq + r;
                    if (j < P.m)
                        stg<float>(static_cast<float *>(P.c) + (uint64_t)j * P.ncols + col0 + lc, ldexpf(acc[jt][r], back));
                }
        }
    };
    rows_run<W>(probs, sc, stage, unit);
}

// ---------------------------------------------------------------------------
// Float64: the same units on the 16x16x4 f64 MFMA.  A group is 8 contraction
// indices: lane (c, q) reads the 16-byte piece A[n0 + c][8 g + 2 q .. + 1]
// (element-aligned when nq is odd) and MFMA x (of two) takes element x, so
// the contraction index of MFMA x at lane quarter q is 8 g + 2 q + x.  A unit
// of 256 contraction indices is 32 groups (128 registers), so the groups are
// loaded and multiplied in chunks of 16.  B is staged with a row stride of
// 8 k + 2 doubles (16-byte fragment reads of 16 rows on distinct bank groups).
// The accumulator of a lane holds C rows 16 jt + 4 r + q of column n0 + c.
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef double f64x2u __attribute__((ext_vector_type(2), aligned(8)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kGroups8 = (int)(kRowsMaxNq / 8);
constexpr int kChunk8 = 16;
constexpr int kStrideMaxD = (int)kRowsMaxNq + 2;
__host__ __device__ constexpr unsigned rows_stride_d(unsigned nq) { return (nq + 7) / 8 * 8 + 2; }

// The last group's piece: A's last two elements when it would run past A's
// end (shift d); elements past the row end (x >= lim) are zero.
__device__ __forceinline__ f64x2 piece_last_d(f64x2 rl, int d, int lim)
{
    f64x2 af;
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        double t = rl[x];
        if (d == 1) t = x == 0 ? rl[1] : 0.0;
        af[x] = x < lim ? t : 0.0;
    }
    return af;
}

template <int MB, int W, int WPE>
__global__ __launch_bounds__(64 * W, WPE) void k_rows_f64(const RowsProblem *__restrict__ probs, RowsSched sc)
{
    __shared__ __attribute__((aligned(16))) double sB[MB * kStrideMaxD];
    constexpr int JT = MB / 16;
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    int const lc = lane & 15, lq = lane >> 4;
    auto stage = [&](const RowsProblem &P) {
        static_assert(MB % W == 0, "whole rows per wave");
        constexpr int RPW = MB / W;
        constexpr int CPL = kStrideMaxD / 64;
        unsigned const nq = P.nq, S = rows_stride_d(nq), width = (nq + 7) / 8 * 8;
        double t[RPW][CPL];
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                unsigned const j = wave + W * r, l = lane + 64 * c;
                t[r][c] = (j < P.m && l < nq) ? ldg<double>(static_cast<const double *>(P.b) + (uint64_t)j * nq + l) : 0.0;
            }
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                unsigned const j = wave + W * r, l = lane + 64 * c;
                if (l < width) sB[j * S + l] = t[r][c];
            }
    };
    auto unit = [&](const RowsProblem &P, unsigned v) {
        unsigned const nq = P.nq, S = rows_stride_d(nq), ng = (nq + 7) / 8;
        uint64_t const total = P.ncols * nq;
        const double *brow = sB + lc * S + 2 * lq;
        int const lim = (int)nq - (int)(8 * (ng - 1) + 2 * lq); // valid elements of the last group's piece
        uint64_t const col0 = (uint64_t)(v - P.unit_begin) * 16u;
        uint64_t const n = col0 + lc < P.ncols ? col0 + lc : P.ncols - 1;
        const double *a = static_cast<const double *>(P.a);
        const double *arow = a + n * nq + 2 * lq;
        uint64_t const e = n * nq + 8 * (ng - 1) + 2 * lq;
        uint64_t const ec = e + 2 <= total ? e : total - 2;
        int const d = (int)(e - ec);
        f64x2 const rl = ldg<f64x2u>(a + ec);

        f64x4 acc[JT];
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) acc[jt] = f64x4{0.0, 0.0, 0.0, 0.0};
        auto group = [&](unsigned g, f64x2 af) {
            f64x2 bf[JT];
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) bf[jt] = *reinterpret_cast<const f64x2 *>(brow + 16 * jt * S + 8 * g);
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int jt = 0; jt < JT; ++jt)
                    acc[jt] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[jt][x], af[x], acc[jt], 0, 0, 0);
        };
        for (int c0 = 0; c0 + 1 < (int)ng; c0 += kChunk8) {
            f64x2 ra[kChunk8];
#pragma unroll
            for (int g = 0; g < kChunk8; ++g)
                if (c0 + g + 1 < (int)ng) ra[g] = ldg<f64x2u>(arow + 8 * (c0 + g));
#pragma unroll
            for (int g = 0; g < kChunk8; ++g)
                if (c0 + g + 1 < (int)ng) group(c0 + g, ra[g]);
        }
        group(ng - 1, piece_last_d(rl, d, lim));
        if (col0 + lc < P.ncols) {
#pragma unroll
            for (int jt = 0; jt < JT; ++jt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    unsigned const j = 16 * jt + 4 * r + lq;
                    if (j < P.m) stg<double>(static_cast<double *>(P.c) + (uint64_t)j * P.ncols + col0 + lc, acc[jt][r]);
                }
        }
    };
    rows_run<W>(probs, sc, stage, unit);
}

// Waves per workgroup (one workgroup per compute unit) and waves per SIMD
// the registers are budgeted for: 16 and 4 for float32 and float64 with
// m <= 16, 8 and 2 for the taller rows and every split-precision kernel
// (whose accumulators, split terms and fragments need up to 256 registers).
struct RowsCfg
{
    int w, wpe;
};
constexpr RowsCfg rows_cfg(int height, RowsKind kind)
{
    return (kind == ROWS_F32 || kind == ROWS_F64) && height == 16 ? RowsCfg{16, 4} : RowsCfg{8, 2};
}

} // namespace

int rows_height(uint64_t m) { return m <= 16 ? 16 : m <= 32 ? 32 : m <= 48 ? 48 : m <= 64 ? 64 : 0; }

bool rows_eligible(uint64_t nq, uint64_t m, uint64_t ncols)
{
    return nq >= kRowsMinNq && nq <= kRowsMaxNq && m >= 1 && m <= 64 && ncols >= 1;
}

unsigned rows_wg_per_cu(int) { return 1; }

unsigned rows_waves(int height, RowsKind kind) { return (unsigned)rows_cfg(height, kind).w; }

std::string rows_symbol(int height, RowsKind kind)
{
    char buf[64];
    RowsCfg const c = rows_cfg(height, kind);
    char const *name = kind == ROWS_BF16X3   ? "k_rows_bf16x3"
                       : kind == ROWS_FP16X2 ? "k_rows_fp16x2"
                       : kind == ROWS_F64    ? "k_rows_f64"
                                             : "k_rows_grouped";
    std::snprintf(buf, sizeof buf, "%s<%d, %d, %d>", name, height, c.w, c.wpe);
    return buf;
}

void rows_layout(std::vector<RowsProblem> &probs, int height, unsigned grid, RowsLayout &out, RowsKind kind)
{
    // cost of a unit in bytes: HBM traffic or the padded matrix work at the
    // MFMA ridge (float32: 157.3 TFLOP/s / 8 TB/s = 19.7 flop per byte;
    // float64: 78.6 / 8 = 9.8), whichever is larger, plus a fixed 2 KB for its
    // instruction stream.  BF16X3 takes six 16x16x32 bf16 products (16 cycles
    // each) per 32 contraction indices where float32 takes eight 16x16x4
    // products (32 cycles each): 0.375 of the matrix time; FP16X2 three
    // products: 0.1875.
    double const ridge = kind == ROWS_F64      ? 9.83
                         : kind == ROWS_BF16X3 ? 19.66 / 0.375
                         : kind == ROWS_FP16X2 ? 19.66 / 0.1875
                                               : 19.66;
    double const esz = kind == ROWS_F64 ? 8.0 : 4.0;
    double const grp = kind == ROWS_F64 ? 8.0 : kind == ROWS_F32 ? 16.0 : 32.0; // contraction group
    // every unit of a problem but its last (fewer columns) costs the same, so
    // the walk below takes whole runs of units at once: O(problems + grid)
    auto cost = [&](const RowsProblem &P, double cols) {
        double const padded = grp * std::ceil(P.nq / grp);
        double const bytes = cols * esz * (P.nq + P.m);
        double const flops = 2.0 * 16.0 * padded * height;
        return std::max(bytes, flops / ridge) + 2048.0;
    };
    std::vector<double> full(probs.size()), last(probs.size());
    unsigned begin = 0;
    double total = 0;
    for (size_t pi = 0; pi < probs.size(); ++pi) {
        RowsProblem &P = probs[pi];
        P.unit_begin = begin;
        P.nunits = (unsigned)((P.ncols + 15) / 16);
        begin += P.nunits;
        full[pi] = cost(P, 16.0);
        last[pi] = cost(P, (double)(P.ncols - (uint64_t)(P.nunits - 1) * 16));
        total += full[pi] * (P.nunits - 1) + last[pi];
    }
    out.units = begin;
    out.wg_units.assign(grid + 1, begin);
    out.wg_prob.assign(grid, 0);
    // workgroup b takes units while the running cost up to the middle of the
    // next unit stays within b + 1 grid-ths of the total
    double acc = 0;
    unsigned u = 0;
    size_t pi = 0;
    for (unsigned b = 0; b < grid; ++b) {
        while (pi < probs.size() && u >= probs[pi].unit_begin + probs[pi].nunits) ++pi;
        out.wg_units[b] = u;
        out.wg_prob[b] = pi < probs.size() ? (unsigned)pi : 0u;
        double const target = total * (b + 1) / grid;
        while (pi < probs.size()) {
            RowsProblem const &P = probs[pi];
            unsigned const lastu = P.unit_begin + P.nunits - 1;
            if (u < lastu) {
                // whole units at the full cost: n with acc + (n - 1/2) full <= target
                double const room = (target - acc) / full[pi] + 0.5;
                if (room < 1.0) break;
                unsigned const n = (unsigned)std::min<double>(std::floor(room), (double)(lastu - u));
                acc += n * full[pi];
                u += n;
                if (u < lastu) break; // the budget ends inside this problem
            }
            if (acc + 0.5 * last[pi] > target) break;
            acc += last[pi];
            u = lastu + 1;
            ++pi;
        }
    }
    out.wg_units[grid] = begin;
}

namespace
{
template <RowsKind K, int H>
void launch_one(const RowsProblem *d_probs, const RowsSched &sc, unsigned grid, hipStream_t s)
{
    constexpr RowsCfg c = rows_cfg(H, K);
    if constexpr (K == ROWS_F32)
        hipLaunchKernelGGL((k_rows_grouped<H, c.w, c.wpe>), dim3(grid), dim3(64 * c.w), 0, s, d_probs, sc);
    else if constexpr (K == ROWS_BF16X3)
        hipLaunchKernelGGL((k_rows_bf16x3<H, c.w, c.wpe>), dim3(grid), dim3(64 * c.w), 0, s, d_probs, sc);
    else if constexpr (K == ROWS_FP16X2)
        hipLaunchKernelGGL((k_rows_fp16x2<H, c.w, c.wpe>), dim3(grid), dim3(64 * c.w), 0, s, d_probs, sc);
    else
        hipLaunchKernelGGL((k_rows_f64<H, c.w, c.wpe>), dim3(grid), dim3(64 * c.w), 0, s, d_probs, sc);
}
template <RowsKind K>
bool launch_kind(int height, const RowsProblem *d_probs, const RowsSched &sc, unsigned grid, hipStream_t s)
{
    switch (height) {
    case 16: launch_one<K, 16>(d_probs, sc, grid, s); return true;
    case 32: launch_one<K, 32>(d_probs, sc, grid, s); return true;
    case 48: launch_one<K, 48>(d_probs, sc, grid, s); return true;
    case 64: launch_one<K, 64>(d_probs, sc, grid, s); return true;
    default: return false;
    }
}
} // namespace

ttm_status_t launch_rows_grouped(int height, const RowsProblem *d_probs, const RowsSched &sc, unsigned grid,
                                 hipStream_t s, RowsKind kind)
{
    if (grid == 0) return TTM_SUCCESS;
    bool const ok = kind == ROWS_F32      ? launch_kind<ROWS_F32>(height, d_probs, sc, grid, s)
                    : kind == ROWS_BF16X3 ? launch_kind<ROWS_BF16X3>(height, d_probs, sc, grid, s)
                    : kind == ROWS_FP16X2 ? launch_kind<ROWS_FP16X2>(height, d_probs, sc, grid, s)
                                          : launch_kind<ROWS_F64>(height, d_probs, sc, grid, s);
    if (!ok) {
        set_error(TTM_ERR_ARGUMENT, "short-row kernel: row height must be 16, 32 or 64");
        return TTM_ERR_ARGUMENT;
    }
    hipError_t const e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(TTM_ERR_HIP, hipGetErrorString(e));
        return TTM_ERR_HIP;
    }
    return TTM_SUCCESS;
}

} // namespace ttm_amd
