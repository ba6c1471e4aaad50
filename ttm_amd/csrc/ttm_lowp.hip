// Reduced-precision arithmetic for float32 products (no counterpart in the
// reference, which is float/double only).  Inputs and output stay float32;
// the operands are converted on the fly while a slice is staged in LDS, and
// the matrix cores accumulate in FP32.
//
//   BF16    x -> bf16(x)                                   1 product
//   FP16    x -> binary16(x)                               1 product
//   BF16X3  x = x0 + x1 + x2, x0 = bf16(x), x1 = bf16(x - x0),
//           x2 = bf16(x - x0 - x1); products x0y0 + x0y1 + x1y0 + x0y2 + x1y1 + x2y0
//           (the six terms of order >= 2^-16 relative; Henry, Tang &
//           Heinecke, "Leveraging the bfloat16 artificial intelligence
//           datatype for higher-precision computations", ARITH 2019)
//   FP8     OCP e4m3fn operands with one power-of-two (E8M0) scale per slice
//           of 32 consecutive contraction indices of each fibre of A and each
//           row of B, applied by the matrix core (v_mfma_scale_*_f8f6f4).
//
// Same folded product and NN / TN views as ttm_gemm.hip:
//   NN (w % 4 == 0): n = k*w + i,  At[l][n] = A[k*nq*w + l*w + i]
//   TN (w == 1):     n = k,        At[l][n] = A[k*nq + l]
//
// MFMA layouts (measured with a lane-layout probe, lane = 16*lg + lc):
//   16x16x32 bf16/f16: srcA lane holds Afrag[lc][8*lg + e], e < 8;
//                      srcB lane holds Bfrag[8*lg + e][lc];
//   16x16x128 f8f6f4:  32 bytes per lane, byte e at k = 16*lg + e (e < 16) and
//                      k = 64 + 16*lg + (e - 16) (e >= 16); the lane's E8M0
//                      scale applies to its row (column) and k in [32*lg, 32*lg + 32);
//   accumulator element r holds D[4*lg + r][lc].
// LDS tiles are therefore stored K-contiguous: s[row][k].
#include "ttm_lowp.hpp"
#include "ttm_workspace.hpp"

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

namespace ttm_amd
{
namespace
{

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
// 16-byte vector at float alignment (rows whose length is not a multiple of 4;
// dword-aligned multi-dword global loads are a supported access)
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

enum LpMode : int
{
    LP_BF16 = 0,
    LP_FP16 = 1,
    LP_BF16X3 = 2,
    LP_FP16X2 = 3
};

template <int MODE>
struct Lp;
template <>
struct Lp<LP_BF16>
{
    using op = __bf16;
    using v2 = bf16x2;
    using v4 = bf16x4;
    using v8 = bf16x8;
    static constexpr int PARTS = 1;
    __device__ static f32x4 mfma(v8 a, v8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
    __device__ static void split(float x, op (&p)[1]) { p[0] = (__bf16)x; }
    __device__ static void split2(f32x2 x, v2 (&p)[1]) { p[0] = __builtin_convertvector(x, bf16x2); }
};
template <>
struct Lp<LP_FP16>
{
    using op = _Float16;
    using v2 = f16x2;
    using v4 = f16x4;
    using v8 = f16x8;
    static constexpr int PARTS = 1;
    __device__ static f32x4 mfma(v8 a, v8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
    __device__ static void split(float x, op (&p)[1]) { p[0] = (_Float16)x; }
    __device__ static void split2(f32x2 x, v2 (&p)[1]) { p[0] = __builtin_convertvector(x, f16x2); }
};
template <>
struct Lp<LP_BF16X3>
{
    using op = __bf16;
    using v2 = bf16x2;
    using v4 = bf16x4;
    using v8 = bf16x8;
    static constexpr int PARTS = 3;
    __device__ static f32x4 mfma(v8 a, v8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
    // x - float(x0) and the next difference are exact in float32.
    __device__ static void split(float x, op (&p)[3])
    {
        p[0] = (__bf16)x;
        float const r1 = x - (float)p[0];
        p[1] = (__bf16)r1;
        float const r2 = r1 - (float)p[1];
        p[2] = (__bf16)r2;
    }
    // The same for two values at once (paired conversions).
    __device__ static void split2(f32x2 x, v2 (&p)[3])
    {
        p[0] = __builtin_convertvector(x, bf16x2);
        f32x2 const r1 = x - __builtin_convertvector(p[0], f32x2);
        p[1] = __builtin_convertvector(r1, bf16x2);
        f32x2 const r2 = r1 - __builtin_convertvector(p[1], f32x2);
        p[2] = __builtin_convertvector(r2, bf16x2);
    }
};

template <>
struct Lp<LP_FP16X2>
{
    using op = _Float16;
    using v2 = f16x2;
    using v4 = f16x4;
    using v8 = f16x8;
    static constexpr int PARTS = 2;
    __device__ static f32x4 mfma(v8 a, v8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
    // x (already scaled into the binary16 range) = x0 + x1 + O(2^-22 |x|);
    // x - float(x0) is exact in float32.
    __device__ static void split(float x, op (&p)[2])
    {
        p[0] = (_Float16)x;
        p[1] = (_Float16)(x - (float)p[0]);
    }
    __device__ static void split2(f32x2 x, v2 (&p)[2])
    {
        p[0] = __builtin_convertvector(x, f16x2);
        p[1] = __builtin_convertvector(x - __builtin_convertvector(p[0], f32x2), f16x2);
    }
};

// FP16X2 scaling: every element of A is multiplied by 2^ea and every element
// of B by 2^eb before the split, with e = 15 - E for the operand's largest
// finite magnitude f * 2^E (f in [0.5, 1)), so the scaled operand lies below
// 2^15 (binary16 reaches 65504); the product is multiplied by 2^-(ea + eb).
// An operand with no finite non-zero value gives e = 0.
__device__ __forceinline__ int fp16x2_exponent(float amax)
{
    if (!(amax > 0.f) || !__builtin_isfinite(amax)) return 0;
    int e = 0;
    (void)frexpf(amax, &e);
    return 15 - e;
}
__device__ __forceinline__ f32x2 scale2(f32x2 x, int e)
{
    f32x2 r = {ldexpf(x[0], e), ldexpf(x[1], e)};
    return r;
}

// Global memory accesses of the tiles.  The grouped kernel takes A, B and C
// from problem descriptors in memory, so the compiler cannot tell that they
// point to global memory and would emit FLAT instructions, which also count
// against the LDS wait counter.  The explicit address space keeps them
// global loads and stores.
template <class V>
__device__ __forceinline__ V ldg(const void *p)
{
    return *(const __attribute__((address_space(1))) V *)p;
}
template <class V>
__device__ __forceinline__ V ldg_nt(const void *p)
{
    return __builtin_nontemporal_load((const __attribute__((address_space(1))) V *)p);
}
template <class V>
__device__ __forceinline__ void stg(void *p, V v)
{
    *(__attribute__((address_space(1))) V *)p = v;
}

__device__ __forceinline__ f32x4 zero4()
{
    f32x4 z = {0.f, 0.f, 0.f, 0.f};
    return z;
}

// Workgroup id -> (row block, column block), row blocks fastest, grouped per XCD.
__device__ __forceinline__ void lp_block_coords(unsigned nj, unsigned total, unsigned &jb, unsigned &nb)
{
    unsigned b = blockIdx.x;
    unsigned const full = total & ~7u;
    if (b < full) b = (b & 7u) * (full >> 3) + (b >> 3);
    jb = b % nj;
    nb = b / nj;
}

// ---------------------------------------------------------------------------
// BF16 / FP16 / BF16X3: 128 x 128 blocks, K-slices of 32, 4 waves of 64 x 64.
// ---------------------------------------------------------------------------
// Views (as in ttm_gemm.hip): NN, TN (C[n][j]) and TT (TN loads, C[j * ncols + n]).
enum LpView : int
{
    LPV_NN = 0,
    LPV_TN = 1,
    LPV_TT = 2
};
constexpr int LP_BM = 128, LP_BN = 128, LP_BK = 32;
constexpr int LP_LDK = LP_BK; // 64-byte rows; bank spreading by the row order and chunk swizzle below

// LDS layout of a 128-row x 32-k operand tile (one part).  MFMA tile t of a
// wave reads the logical rows {64h + 4x + t : x < 16} (interleaved
// fragments), so logical row r = 64h + 4y + x is kept at physical row
// lds_row(r) = 64h + 16x + y: a fragment read covers 16 consecutive physical
// rows.  Each physical row P holds four 16-byte chunks (8 k each); chunk c
// sits at position c ^ lds_swz(P).  With this order every fragment read
// (16 rows, one chunk), every 4-row x 4-chunk store and every 16-row x
// 8-byte store touches distinct banks (checked for 32 banks / 8 lanes and
// 64 banks / 16 lanes per cycle).
__device__ __forceinline__ int lds_row(int r) { return (r & ~63) | ((r & 3) << 4) | ((r & 63) >> 2); }
__device__ __forceinline__ int lds_row_inv(int P) { return (P & ~63) | ((P & 15) << 2) | ((P & 63) >> 4); }
__device__ __forceinline__ int lds_swz(int P) { return ((P >> 1) & 3) ^ ((P >> 3) & 1); }
// element offset of (physical row P, chunk c) within a part
__device__ __forceinline__ int lds_at(int P, int c) { return P * LP_LDK + 8 * (c ^ lds_swz(P)); }

// One 128 x 128 block of D at (j0, n0).  flags: TILE_A_VEC / TILE_C_VEC /
// TILE_STREAM as in ttm_gemm.hpp.  With BSPLIT, B comes pre-converted: P
// planes of `op` at bs + p * plane, row j at j * ldb, zero padded to
// multiples of 32 columns and 128 rows (k_lowp_split_b); otherwise B is the
// row-major float32 m x nq matrix b, converted here.
// PIPE (with BSPLIT): A's global loads are issued two slices ahead (two
// register sets), and B's planes go straight from memory into two LDS
// buffers sB / sB1 by LDS DMA, one slice ahead, so no slice waits for the
// previous one's loads; the loads that miss the L2 (A's, the first time a
// column block is read) get two slices of matrix-core work to arrive.
template <int MODE, int VIEW, bool BSPLIT, int BM = LP_BM, bool PIPE = false>
__device__ __forceinline__ void lowp_tile(const float *__restrict__ a, const float *__restrict__ b,
                                          const typename Lp<MODE>::op *__restrict__ bs, uint64_t ldb, uint64_t plane,
                                          float *__restrict__ c, uint64_t ncols, uint64_t nq, uint64_t w, uint64_t m,
                                          uint64_t j0, uint64_t n0, unsigned flags,
                                          typename Lp<MODE>::op (*sA)[LP_BN * LP_LDK],
                                          typename Lp<MODE>::op (*sB)[BM * LP_LDK], int ea_in = 0, int ce = 0,
                                          typename Lp<MODE>::op (*sB1)[BM * LP_LDK] = nullptr)
{
    // BM = 128: wave wm owns the 64-row group wm (MFMA tiles t = 0..3);
    // BM = 64 (products of at most 64 rows): both waves along m share the one
    // 64-row group, wave wm taking tiles t = 2 wm, 2 wm + 1.  The LDS row
    // order is the same (tile t of group h reads physical rows 64h + 16t + lc).
    static_assert(BM == 128 || BM == 64, "64- or 128-row blocks");
    constexpr int BH = BM / 64;  // 64-row groups of the block
    constexpr bool TN = VIEW != LPV_NN;
    constexpr bool TSTORE = VIEW == LPV_TN;
    // COLSPLIT (64-row blocks storing C[n][j]): the four waves split the
    // columns instead (32 each) and every wave takes all four row tiles, so a
    // lane holds four consecutive rows j of a column of C and stores them as
    // one 16-byte vector (with the waves split along m it held two).
    constexpr bool COLSPLIT = BM == 64 && TSTORE;
    constexpr int JTL = COLSPLIT ? 4 : BM / 32; // MFMA row tiles per wave
    constexpr int NTL = COLSPLIT ? 2 : 4;       // MFMA column tiles per wave
    using L = Lp<MODE>;
    using op = typename L::op;
    using op2 = typename L::v2;
    using v4 = typename L::v4;
    using v8 = typename L::v8;
    constexpr int P = L::PARTS;
    constexpr int BN = LP_BN, BK = LP_BK;
    bool const a_vec = (flags & TILE_A_VEC) != 0;
    bool const c_vec = (flags & TILE_C_VEC) != 0;
    bool const stream = (flags & TILE_STREAM) != 0;
    int const tid = threadIdx.x;
    int const lane = tid & 63, wave = tid >> 6;
    int const lc = lane & 15, lg = lane >> 4;
    int const wm = wave & 1, wn = wave >> 1;
    // FP16X2 scale of A.  128-row blocks of a single product take one
    // exponent for all of A (ea_in, from a pass over A before the product).
    // 64-row blocks (at most 64 rows of B: HBM-bound products, where that
    // pass would read A a second time) and the grouped blocks take one per
    // 128-column block and 32-wide slice: the running minimum over the
    // block's slices so far of 15 - E(largest finite |A| of the slice), so
    // every scaled value stays below 2^15; the accumulators are rescaled
    // exactly (by 2^(new - old) <= 1) when it drops.  The grouped blocks
    // convert B themselves and scale it the same way per block and slice.
    constexpr bool SLICE_SCALE = MODE == LP_FP16X2 && (BM == 64 || !BSPLIT);
    constexpr bool B_SLICE = MODE == LP_FP16X2 && !BSPLIT;
    constexpr int NO_EXP = 1 << 20; // no finite non-zero value seen yet
    __shared__ float smax[2][4];
    int ea = SLICE_SCALE ? NO_EXP : ea_in;
    int eb_s = NO_EXP;
    auto a_exp = [&]() { return ea == NO_EXP ? 0 : ea; };
    auto b_exp = [&]() { return eb_s == NO_EXP ? 0 : eb_s; };

    // ---- loaders --------------------------------------------------------
    // "MN" source (NN A): a thread loads a 4 (k) x 4 (row) micro-block with
    // four 16-byte loads along the rows (logical rows 64H + 4y + x, x < 4;
    // k = 8*wave + 4*half + e) and writes each row's 4 k-values as one 8-byte
    // store.  "K" sources (TN A, row-major B): a thread loads 8 consecutive k
    // (chunk kc) of the two physical rows R and R + 64 and writes one 16-byte
    // store per row.  A tile: rows n, B tile: rows j; both 128 rows x 32 k.
    int const mn_half = lane & 1, mn_y = (lane >> 1) & 15, mn_h = lane >> 5;
    int const mn_c4 = 64 * mn_h + 4 * mn_y;    // first of the thread's 4 logical rows
    int const mn_k4 = 8 * wave + 4 * mn_half; // first of its 4 k
    int const k_c = tid & 3;                  // K sources: chunk (k = 8*k_c .. 8*k_c + 7)
    int const k_p = tid >> 2;                 // K sources: physical rows k_p and k_p + 64
    int const k_row = lds_row_inv(k_p);       // logical row of k_p (k_row + 64 for the other)

    // A: per-thread base addresses (rows past the end read row 0 and are dropped)
    uint64_t a_base[2] = {0, 0};
    bool a_ok[2] = {false, false};
    if constexpr (!TN) {
        uint64_t const n = n0 + mn_c4;
        a_ok[0] = n < ncols;
        uint64_t const k = a_ok[0] ? n / w : 0;
        a_base[0] = k * nq * w + (a_ok[0] ? n - k * w : 0) + (uint64_t)mn_k4 * w;
    } else {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            uint64_t const n = n0 + k_row + 64 * s;
            a_ok[s] = n < ncols;
            a_base[s] = (a_ok[s] ? n * nq : 0) + 8 * k_c;
        }
    }
    // every column of the tile in range and whole slices: unguarded 16-byte loads
    bool const full = n0 + BN <= ncols && nq % BK == 0 && (!TN || a_vec);

    // Registers of the next slice's global loads (one set; the index keeps
    // the loaders generic over register sets).
    using I0 = std::integral_constant<int, 0>;
    static_assert(!PIPE || BSPLIT, "LDS DMA copies B's planes");
    f32x4 ra[PIPE ? 2 : 1][4];
    f32x4 rb[1][4];
    i32x4 rbs[1][BSPLIT ? P : 1][2];
    auto load_a_vec = [&](const float *p) -> f32x4 {
        return stream ? ldg_nt<f32x4>(p) : ldg<f32x4>(p);
    };
    auto load_a = [&](uint64_t l0, auto full_tag, auto set_tag) {
        constexpr int S = decltype(set_tag)::value;
        if constexpr (decltype(full_tag)::value) {
            if constexpr (!TN) {
                const float *pa = a + a_base[0] + l0 * w;
#pragma unroll
                for (int e = 0; e < 4; ++e) ra[S][e] = load_a_vec(pa + e * w);
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) ra[S][e] = load_a_vec(a + a_base[e >> 1] + l0 + 4 * (e & 1));
            }
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) ra[S][e] = zero4();
            if constexpr (!TN) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    uint64_t const l = l0 + mn_k4 + e;
                    if (a_ok[0] && l < nq) ra[S][e] = load_a_vec(a + a_base[0] + (l0 + e) * w);
                }
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    int const s = e >> 1;
                    uint64_t const l = l0 + 8 * k_c + 4 * (e & 1);
                    if (a_ok[s] && l < nq) {
                        if (a_vec) {
                            ra[S][e] = load_a_vec(a + a_base[s] + l0 + 4 * (e & 1));
                        } else if (l + 4 <= nq) {
                            f32x4 const v = ldg<f32x4u>(a + a_base[s] + l0 + 4 * (e & 1));
                            ra[S][e] = v;
                        } else {
#pragma unroll
                            for (int x = 0; x < 4; ++x)
                                if (l + x < nq) ra[S][e][x] = ldg<float>(a + a_base[s] + l0 + 4 * (e & 1) + x);
                        }
                    }
                }
            }
        }
    };
    auto load_b = [&](uint64_t l0, auto set_tag) {
        constexpr int S = decltype(set_tag)::value;
        if constexpr (BSPLIT) {
            // padded planes: every load in range
#pragma unroll
            for (int p = 0; p < P; ++p)
#pragma unroll
                for (int s = 0; s < BH; ++s)
                    rbs[S][p][s] = ldg<i32x4>(bs + p * plane + (j0 + k_row + 64 * s) * ldb + l0 +
                                                                    8 * k_c);
        } else {
#pragma unroll
            for (int e = 0; e < 2 * BH; ++e) {
                rb[S][e] = zero4();
                uint64_t const j = j0 + k_row + 64 * (e >> 1);
                uint64_t const l = l0 + 8 * k_c + 4 * (e & 1);
                if (j >= m) continue;
                if (l + 4 <= nq) {
                    f32x4 const v = ldg<f32x4u>(b + j * nq + l); // 16-byte aligned when nq % 4 == 0
                    rb[S][e] = v;
                } else {
#pragma unroll
                    for (int x = 0; x < 4; ++x)
                        if (l + x < nq) rb[S][e][x] = ldg<float>(b + j * nq + l + x);
                }
            }
        }
    };
    auto load = [&](uint64_t l0, auto full_tag, auto set_tag) {
        load_a(l0, full_tag, set_tag);
        load_b(l0, set_tag);
    };

    // Convert and store, in four pieces: piece t of the A tile is row x = t
    // of the MN micro-block (r[e][x] = logical row mn_c4 + x, k mn_k4 + e),
    // or, for K rows, half h = t of physical rows k_p + 64h (r[2h], r[2h+1] =
    // its 8 k; pieces 2 and 3 empty); the B tile goes with piece 3.
    // (A from register set SA, B from set S)
    auto store_piece = [&](int t, auto set_tag, op (*sa)[BN * LP_LDK], op (*sb)[BM * LP_LDK], auto a_set_tag) {
        constexpr int S = decltype(set_tag)::value;
        constexpr int SA = decltype(a_set_tag)::value;
        if constexpr (!TN) {
            int const x = t;
            op2 lo[P], hi[P];
            f32x2 x01 = {ra[SA][0][x], ra[SA][1][x]};
            f32x2 x23 = {ra[SA][2][x], ra[SA][3][x]};
            if constexpr (MODE == LP_FP16X2) {
                x01 = scale2(x01, a_exp());
                x23 = scale2(x23, a_exp());
            }
            L::split2(x01, lo);
            L::split2(x23, hi);
            int const off = lds_at(lds_row(mn_c4 + x), wave) + 4 * mn_half;
#pragma unroll
            for (int p = 0; p < P; ++p)
                *reinterpret_cast<v4 *>(&sa[p][off]) = __builtin_shufflevector(lo[p], hi[p], 0, 1, 2, 3);
        } else if (t < 2) {
            int const h = t;
            op2 q[4][P];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                f32x2 x = {ra[SA][2 * h + u / 2][2 * (u % 2)], ra[SA][2 * h + u / 2][2 * (u % 2) + 1]};
                if constexpr (MODE == LP_FP16X2) x = scale2(x, a_exp());
                L::split2(x, q[u]);
            }
            int const off = lds_at(k_p + 64 * h, k_c);
#pragma unroll
            for (int p = 0; p < P; ++p) {
                v4 const lo = __builtin_shufflevector(q[0][p], q[1][p], 0, 1, 2, 3);
                v4 const hi = __builtin_shufflevector(q[2][p], q[3][p], 0, 1, 2, 3);
                *reinterpret_cast<v8 *>(&sa[p][off]) = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
            }
        }
        if (t == 3 && !PIPE) {
            if constexpr (BSPLIT) {
#pragma unroll
                for (int p = 0; p < P; ++p)
#pragma unroll
                    for (int s = 0; s < BH; ++s)
                        *reinterpret_cast<i32x4 *>(&sb[p][lds_at(k_p + 64 * s, k_c)]) = rbs[S][p][s];
            } else {
#pragma unroll
                for (int h = 0; h < BH; ++h) {
                    op2 q[4][P];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        f32x2 x = {rb[S][2 * h + u / 2][2 * (u % 2)], rb[S][2 * h + u / 2][2 * (u % 2) + 1]};
                        if constexpr (MODE == LP_FP16X2) x = scale2(x, b_exp());
                        L::split2(x, q[u]);
                    }
                    int const off = lds_at(k_p + 64 * h, k_c);
#pragma unroll
                    for (int p = 0; p < P; ++p) {
                        v4 const lo = __builtin_shufflevector(q[0][p], q[1][p], 0, 1, 2, 3);
                        v4 const hi = __builtin_shufflevector(q[2][p], q[3][p], 0, 1, 2, 3);
                        *reinterpret_cast<v8 *>(&sb[p][off]) = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
                    }
                }
            }
        }
    };
    auto store_all = [&](auto set_tag, op (*sa)[BN * LP_LDK], op (*sb)[BM * LP_LDK], auto a_set_tag) {
#pragma unroll
        for (int t = 0; t < 4; ++t) store_piece(t, set_tag, sa, sb, a_set_tag);
    };

    f32x4 acc[JTL][NTL];
#pragma unroll
    for (int jt = 0; jt < JTL; ++jt)
#pragma unroll
        for (int nt = 0; nt < NTL; ++nt) acc[jt][nt] = zero4();

    // Interleaved fragments: tile jt holds rows 4*x + jt (x < 16) of the
    // wave's 64 rows, tile nt columns 4*y + nt, so a lane's 4 results of a
    // row (column) are adjacent and leave with one 16-byte store.  Physical
    // rows: tile t, lane lc -> 64*half + 16*t + lc.
    int const arow = COLSPLIT ? wave * 32 + lc : wn * 64 + lc; // + 16 * nt
    int const brow = COLSPLIT ? lc : wm * (BM / 2) + lc;       // + 16 * jt

    // The products of one slice from the LDS tiles (sa, sb).
    auto multiply = [&](op (*sa)[BN * LP_LDK], op (*sb)[BM * LP_LDK]) {
        v8 fa[P][NTL];
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
            for (int nt = 0; nt < NTL; ++nt) fa[p][nt] = *reinterpret_cast<const v8 *>(&sa[p][lds_at(arow + 16 * nt, lg)]);
#pragma unroll
        for (int jt = 0; jt < JTL; ++jt) {
            v8 fb[P];
#pragma unroll
            for (int p = 0; p < P; ++p) fb[p] = *reinterpret_cast<const v8 *>(&sb[p][lds_at(brow + 16 * jt, lg)]);
            // per accumulator the products go smallest first; the four
            // accumulators of a row tile take turns (independent MFMAs
            // back to back)
            auto step = [&](int pb, int pa) {
#pragma unroll
                for (int nt = 0; nt < NTL; ++nt) acc[jt][nt] = L::mfma(fb[pb], fa[pa][nt], acc[jt][nt]);
            };
            if constexpr (P == 3) {
                step(1, 1);
                step(0, 2);
                step(2, 0);
                step(0, 1);
                step(1, 0);
            } else if constexpr (P == 2) {
                step(0, 1);
                step(1, 0);
            }
            step(0, 0);
        }
    };

    uint64_t const nslices = (nq + BK - 1) / BK;

    // SLICE_SCALE: the block-wide maximum of a slice, reduced at the barrier
    // the loop already has (smax is read after that barrier and written again
    // only after the next one).
    auto wave_max = [&](float mx) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
        return mx;
    };
    auto slice_max = [&]() {
        float mx = 0.f;
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                float const v = fabsf(ra[0][e][x]);
                mx = fmaxf(mx, v < __builtin_inff() ? v : 0.f); // infinities and NaNs set no scale
            }
        mx = wave_max(mx);
        if (lane == 0) smax[0][wave] = mx;
        if constexpr (B_SLICE) {
            float mb = 0.f;
#pragma unroll
            for (int e = 0; e < 2 * BH; ++e)
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    float const v = fabsf(rb[0][e][x]);
                    mb = fmaxf(mb, v < __builtin_inff() ? v : 0.f);
                }
            mb = wave_max(mb);
            if (lane == 0) smax[1][wave] = mb;
        }
    };
    // new exponent of one operand from its slice maximum; returns the drop
    // (<= 0) to apply to the accumulators
    auto lower = [&](float m, int &e_cur) {
        if (!(m > 0.f)) return 0;
        int E = 0;
        (void)frexpf(m, &E);
        int const e = 15 - E;
        if (e >= e_cur) return 0;
        int const d = e_cur == NO_EXP ? 0 : e - e_cur; // nothing accumulated yet in the old units
        e_cur = e;
        return d;
    };
    auto update_scale = [&]() {
        int d = lower(fmaxf(fmaxf(smax[0][0], smax[0][1]), fmaxf(smax[0][2], smax[0][3])), ea);
        if constexpr (B_SLICE) d += lower(fmaxf(fmaxf(smax[1][0], smax[1][1]), fmaxf(smax[1][2], smax[1][3])), eb_s);
        if (d != 0) {
#pragma unroll
            for (int jt = 0; jt < JTL; ++jt)
#pragma unroll
                for (int nt = 0; nt < NTL; ++nt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc[jt][nt][r] = ldexpf(acc[jt][nt][r], d);
        }
    };

    // one LDS buffer: store, barrier, prefetch the next slice, multiply,
    // barrier (two workgroups per CU overlap each other's phases)
    auto kloop = [&](auto full_tag) {
        load(0, full_tag, I0());
        if constexpr (SLICE_SCALE) {
            slice_max();
            __syncthreads();
            update_scale();
        }
        for (uint64_t s = 0; s < nslices; ++s) {
            store_all(I0(), sA, sB, I0());
            __syncthreads();
            if (s + 1 < nslices) load((s + 1) * BK, full_tag, I0());
            multiply(sA, sB);
            if constexpr (SLICE_SCALE)
                if (s + 1 < nslices) slice_max();
            __syncthreads();
            if constexpr (SLICE_SCALE)
                if (s + 1 < nslices) update_scale();
        }
    };
    // PIPE.  B's planes by LDS DMA: each lane moves 16 bytes and a wave's
    // 64 lanes fill 1 KB of LDS in lane order, i.e. 16 physical rows x 4
    // chunks; lane L of the piece at physical row P0 takes physical row
    // P = P0 + L / 4, position L % 4, so it loads logical row lds_row_inv(P),
    // chunk (L % 4) ^ lds_swz(P) (the layout the register path stores).
    // Wave w moves rows 64h + 16w .. + 15 of each part p: P * BH pieces.
    // The planes are padded to whole blocks and slices: no load is guarded.
    auto dma_b = [&](uint64_t l0, op (*sb)[BM * LP_LDK]) {
        int const wv = __builtin_amdgcn_readfirstlane(wave);
        int const cp = lane & 3;
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
            for (int h = 0; h < BH; ++h) {
                int const p0 = 64 * h + 16 * wv;
                int const pr = p0 + (lane >> 2);
                uint32_t const off = (uint32_t)(lds_row_inv(pr) & 63) * (uint32_t)ldb + 8u * (uint32_t)(cp ^ lds_swz(pr));
                const op *src = bs + p * plane + (j0 + 64 * h) * ldb + l0 + off;
                __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *)src,
                                                 (void __attribute__((address_space(3))) *)&sb[p][p0 * LP_LDK], 16,
                                                 0, 0);
            }
    };
    // Barrier that waits for this wave's LDS operations and for all but the
    // last `N` of its vector-memory operations (the LDS DMA of the slice
    // about to be read was issued before the A loads of the next slice).
    // No fence: the LDS DMA of the next slice stays in flight across it.
    // The waits go through the builtin (s_waitcnt immediate: vmcnt in bits
    // 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8), so the compiler counts them
    // and adds none of its own for the DMA they cover; the barrier is a
    // compiler memory barrier as well.
    auto barrier = [&]() { asm volatile("s_barrier" ::: "memory"); };
    auto sync_vm4 = [&]() {
        __builtin_amdgcn_s_waitcnt(0x0074); // vmcnt(4) lgkmcnt(0)
        barrier();
    };
    auto sync_vm0 = [&]() {
        __builtin_amdgcn_s_waitcnt(0x0070); // vmcnt(0) lgkmcnt(0)
        barrier();
    };
    auto sync_lds = [&]() {
        __builtin_amdgcn_s_waitcnt(0xC07F); // lgkmcnt(0)
        barrier();
    };
    using I1 = std::integral_constant<int, 1>;
    static_assert(!PIPE || !SLICE_SCALE, "per-slice scales read register set 0");
    // Slice s: A in register set s % 2 (loaded during slice s - 2), B in
    // LDS buffer s % 2 (DMA issued during slice s - 1).  Every slice issues
    // the same loads -- past the end they repeat the last slice, whose data
    // is never used -- so the compiler's waits count exactly: before A(s) is
    // stored, DMA(s) and A(s + 1) may stay in flight (10 instructions), and
    // the wait for this wave's DMA(s) before the barrier leaves A(s + 1) (4
    // instructions).  The guarded loop issues a varying number: it waits
    // for all.
    auto kloop_pipe = [&](auto full_tag) {
        constexpr bool FULL = decltype(full_tag)::value;
        uint64_t const last = (nslices - 1) * BK;
        auto at = [&](uint64_t t) { return t * BK < last ? t * BK : last; };
        auto wait_b = [&]() {
            if constexpr (FULL) sync_vm4();
            else sync_vm0();
        };
        load_a(0, full_tag, I0());
        __builtin_amdgcn_sched_barrier(0);
        dma_b(0, sB);
        __builtin_amdgcn_sched_barrier(0);
        load_a(at(1), full_tag, I1());
        // The compiler drains every load at the loop head (its count of
        // loads still in flight does not survive the back edge), so one
        // trip runs four slices: the drain waits only for the loads issued
        // during the trip's last slice.  The scheduling barriers keep each
        // slice's loads ahead of its multiply.
        auto half0 = [&](uint64_t s) {
            store_all(I0(), sA, sB, I0());
            wait_b();
            dma_b(at(s + 1), sB1);
            __builtin_amdgcn_sched_barrier(0);
            load_a(at(s + 2), full_tag, I0());
            __builtin_amdgcn_sched_barrier(0);
            multiply(sA, sB);
            sync_lds();
        };
        auto half1 = [&](uint64_t s) {
            store_all(I0(), sA, sB1, I1());
            wait_b();
            dma_b(at(s + 1), sB);
            __builtin_amdgcn_sched_barrier(0);
            load_a(at(s + 2), full_tag, I1());
            __builtin_amdgcn_sched_barrier(0);
            multiply(sA, sB1);
            sync_lds();
        };
        uint64_t s = 0;
        for (; s + 3 < nslices; s += 4) {
            half0(s);
            half1(s + 1);
            half0(s + 2);
            half1(s + 3);
        }
        if (s + 1 < nslices) {
            half0(s);
            half1(s + 1);
            s += 2;
        }
        if (s < nslices) half0(s);
        // no LDS DMA may still be writing when the workgroup ends
        __builtin_amdgcn_s_waitcnt(0x0F70); // vmcnt(0)
    };
    if constexpr (PIPE) {
        if (full) kloop_pipe(std::true_type());
        else kloop_pipe(std::false_type());
    } else {
        if (full) kloop(std::true_type());
        else kloop(std::false_type());
    }
    if constexpr (MODE == LP_FP16X2) {
        // ce = ea_in + eb (B's exponent from the pass over B); with slice
        // scales ea_in is 0 and A's exponent is ea, and the grouped blocks
        // hold B's exponent themselves
        int const cfin = B_SLICE ? a_exp() + b_exp() : (SLICE_SCALE ? ce + a_exp() : ce);
#pragma unroll
        for (int jt = 0; jt < JTL; ++jt)
#pragma unroll
            for (int nt = 0; nt < NTL; ++nt)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[jt][nt][r] = ldexpf(acc[jt][nt][r], -cfin);
    }

    // ---- epilogue ------------------------------------------------------------
    uint64_t const ncol0 = n0 + wn * 64 + 4 * lc; // this lane's 4 adjacent columns
    if constexpr (!TSTORE) {
        if (ncol0 >= ncols) return;
        uint64_t const k = ncol0 / w;
        float *cb = c + k * m * w + (ncol0 - k * w);
#pragma unroll
        for (int jt = 0; jt < JTL; ++jt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // logical row 64h + 4x + t of MFMA tile t, acc row x = 4 lg + r
                uint64_t const j = j0 + (BM == 128 ? 64 * wm + 4 * (4 * lg + r) + jt : 4 * (4 * lg + r) + 2 * wm + jt);
                if (j >= m) continue;
                if (c_vec) {
                    f32x4 o = {acc[jt][0][r], acc[jt][1][r], acc[jt][2][r], acc[jt][3][r]};
                    stg<f32x4>(cb + j * w, o);
                } else {
                    // TT with ncols % 4 != 0: the 4 columns may run past the end
#pragma unroll
                    for (int nt = 0; nt < 4; ++nt)
                        if (ncol0 + nt < ncols) stg<float>(cb + j * w + nt, acc[jt][nt][r]);
                }
            }
    } else if constexpr (COLSPLIT) {
        // C[n][j]: MFMA tile nt of wave w reads physical A rows 32 w + 16 nt + lc,
        // i.e. column 64 (w / 2) + 4 lc + 2 (w % 2) + nt; its row tile jt reads
        // physical B rows 16 jt + lc, so accumulator element r of tile jt is
        // row 4 (4 lg + r) + jt: the four tiles give four consecutive rows
#pragma unroll
        for (int nt = 0; nt < NTL; ++nt) {
            uint64_t const k = n0 + 64 * (wave >> 1) + 4 * lc + 2 * (wave & 1) + nt;
            if (k >= ncols) continue;
            float *crow = c + k * m;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                uint64_t const jr = j0 + 4 * (4 * lg + r);
                if ((m & 3) == 0 && jr + 4 <= m) {
                    f32x4 o = {acc[0][nt][r], acc[1][nt][r], acc[2][nt][r], acc[3][nt][r]};
                    stg<f32x4>(crow + jr, o);
                } else {
#pragma unroll
                    for (int jt = 0; jt < 4; ++jt)
                        if (jr + jt < m) stg<float>(crow + jr + jt, acc[jt][nt][r]);
                }
            }
        }
    } else {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            uint64_t const k = ncol0 + nt;
            if (k >= ncols) continue;
            float *crow = c + k * m;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                uint64_t const jr = j0 + wm * 64 + 4 * (4 * lg + r);
                if ((m & 3) == 0 && jr + 4 <= m) {
                    f32x4 o = {acc[0][nt][r], acc[1][nt][r], acc[2][nt][r], acc[3][nt][r]};
                    stg<f32x4>(crow + jr, o);
                } else {
#pragma unroll
                    for (int jt = 0; jt < 4; ++jt)
                        if (jr + jt < m) stg<float>(crow + jr + jt, acc[jt][nt][r]);
                }
            }
        }
    }
}

// B -> P planes of `op` (the operand conversion of the tile, done once per
// product): plane p at bs + p * plane, row j at j * ldb, zero outside m x nq.
// One thread per 8 consecutive l of a row.
template <int MODE>
__global__ __launch_bounds__(256) void k_lowp_split_b(const float *__restrict__ b, typename Lp<MODE>::op *__restrict__ bs,
                                                      uint64_t m, uint64_t nq, uint64_t bsj, uint64_t bsl, uint64_t ldb,
                                                      uint64_t rows, uint64_t plane, const int *__restrict__ exps)
{
    using L = Lp<MODE>;
    using v2 = typename L::v2;
    using v8 = typename L::v8;
    constexpr int P = L::PARTS;
    uint64_t const groups = ldb / 8;
    uint64_t const idx = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= rows * groups) return;
    uint64_t const j = idx / groups, l0 = (idx - j * groups) * 8;
    v2 q[4][P];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x2 x;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            uint64_t const l = l0 + 2 * t + e;
            x[e] = (j < m && l < nq) ? b[j * bsj + l * bsl] : 0.f;
        }
        if constexpr (MODE == LP_FP16X2) x = scale2(x, exps[1]);
        L::split2(x, q[t]);
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
        auto const lo = __builtin_shufflevector(q[0][p], q[1][p], 0, 1, 2, 3);
        auto const hi = __builtin_shufflevector(q[2][p], q[3][p], 0, 1, 2, 3);
        *reinterpret_cast<v8 *>(bs + p * plane + j * ldb + l0) = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
    }
}

// One product: a grid of BM x 128 blocks over the pre-converted B (BM = 64
// for products of at most 64 rows, else 128).  FP16X2 (two LDS planes, 32 KB
// per block) runs three blocks per CU: its register budget drops to 168 (a
// few values spill), which measured 3.85-4.54 ms against 4.22-4.85 ms with
// two blocks on the 1024-row float32 products of A 1024x1024x512.
// Blocks per CU the register budget is set for: three (168 registers) where
// the tile fits or nearly fits them, two (256) for the 128-row BF16X3 tile.
constexpr int lp_blocks_per_cu(int mode, int bm) { return mode == LP_BF16X3 && bm == 128 ? 2 : 3; }
// Single-product tiles with the pipelined slice loop (PIPE).
constexpr bool lp_pipe(int mode, int bm) { return mode == LP_BF16X3 && bm == 128; }

template <int MODE, bool TN, int BM>
__global__ __launch_bounds__(256, lp_blocks_per_cu(MODE, BM)) void k_lowp(const float *__restrict__ a, const typename Lp<MODE>::op *__restrict__ bs,
                                                 uint64_t ldb, uint64_t plane, float *__restrict__ c, uint64_t ncols,
                                                 uint64_t nq, uint64_t w, uint64_t m, unsigned nj, unsigned total,
                                                 const int *__restrict__ exps)
{
    using op = typename Lp<MODE>::op;
    __shared__ __attribute__((aligned(16))) op sA[Lp<MODE>::PARTS][LP_BN * LP_LDK];
    __shared__ __attribute__((aligned(16))) op sB[Lp<MODE>::PARTS][BM * LP_LDK];
    // second B buffer of the pipelined loop (unreferenced, and so not
    // allocated, otherwise)
    __shared__ __attribute__((aligned(16))) op sB1[Lp<MODE>::PARTS][BM * LP_LDK];
    unsigned jb, nb;
    lp_block_coords(nj, total, jb, nb);
    lowp_tile<MODE, TN ? LPV_TN : LPV_NN, true, BM, lp_pipe(MODE, BM)>(
        a, nullptr, bs, ldb, plane, c, ncols, nq, w, m, (uint64_t)jb * BM, (uint64_t)nb * LP_BN,
        TILE_A_VEC | TILE_C_VEC | (nj == 1 ? TILE_STREAM : 0u), sA, sB, MODE == LP_FP16X2 ? exps[0] : 0,
        MODE == LP_FP16X2 ? exps[0] + exps[1] : 0, sB1);
}

// FP16X2 scales: the largest magnitude of A (blocks [0, blocks_a)) and of B
// (the remaining blocks), one partial maximum per block, then one block
// turns both into exponents (no atomics: the result does not depend on the
// order of the blocks).  Infinities and NaNs set no scale (an infinity in
// one place must not flush every finite value); they stay what they are in
// the product through the conversion.
__device__ __forceinline__ float block_max(float v, float *red)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    v = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    __syncthreads();
    return v;
}

__global__ __launch_bounds__(256) void k_absmax_partial(const float *__restrict__ a, uint64_t na,
                                                        const float *__restrict__ b, uint64_t nb, unsigned blocks_a,
                                                        float *__restrict__ part)
{
    __shared__ float red[4];
    bool const is_a = blockIdx.x < blocks_a;
    const float *x = is_a ? a : b;
    uint64_t const n = is_a ? na : nb;
    uint64_t const blk = is_a ? blockIdx.x : blockIdx.x - blocks_a;
    uint64_t const nblk = is_a ? blocks_a : gridDim.x - blocks_a;
    uint64_t const stride = nblk * 256;
    float mx = 0.f;
    uint64_t vec_end = 0;
    if ((reinterpret_cast<uintptr_t>(x) & 15u) == 0) {
        uint64_t const nv = n / 4;
        for (uint64_t i = blk * 256 + threadIdx.x; i < nv; i += stride) {
            f32x4 const v = ldg_nt<f32x4>(x + 4 * i);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float const t = fabsf(v[e]);
                mx = fmaxf(mx, t < __builtin_inff() ? t : 0.f);
            }
        }
        vec_end = nv * 4;
    }
    for (uint64_t i = vec_end + blk * 256 + threadIdx.x; i < n; i += stride) {
        float const t = fabsf(x[i]);
        mx = fmaxf(mx, t < __builtin_inff() ? t : 0.f);
    }
    mx = block_max(mx, red);
    if (threadIdx.x == 0) part[blockIdx.x] = mx;
}

__global__ __launch_bounds__(256) void k_absmax_final(const float *__restrict__ part, unsigned blocks_a,
                                                      unsigned blocks, int *__restrict__ exps)
{
    __shared__ float red[4];
    float ma = 0.f, mb = 0.f;
    for (unsigned i = threadIdx.x; i < blocks; i += 256) {
        if (i < blocks_a) ma = fmaxf(ma, part[i]);
        else mb = fmaxf(mb, part[i]);
    }
    ma = block_max(ma, red);
    mb = block_max(mb, red);
    if (threadIdx.x == 0) {
        exps[0] = fp16x2_exponent(ma);
        exps[1] = fp16x2_exponent(mb);
    }
}

// Many TT products (stages of multi-TTM chains) in one persistent launch;
// BM rows (128, or 64 for products of at most 64 rows) by 128 columns.
// The grouped blocks also convert B and scale it per slice: 128-row blocks
// run two per CU.
constexpr int lp_group_blocks_per_cu(int bm) { return bm == 128 ? 2 : 3; }

template <int MODE, int BM>
__global__ __launch_bounds__(256, lp_group_blocks_per_cu(BM)) void k_lowp_grouped(const GroupProblem *__restrict__ probs,
                                                         const unsigned *__restrict__ tile_prob, unsigned total)
{
    using op = typename Lp<MODE>::op;
    __shared__ __attribute__((aligned(16))) op sA[Lp<MODE>::PARTS][LP_BN * LP_LDK];
    __shared__ __attribute__((aligned(16))) op sB[Lp<MODE>::PARTS][BM * LP_LDK];
    for (unsigned t = blockIdx.x; t < total; t += gridDim.x) {
        GroupProblem const &P = probs[tile_prob[t]];
        unsigned const local = t - P.tile_begin;
        unsigned const jb = local % P.nj, nb = local / P.nj;
        lowp_tile<MODE, LPV_TT, false, BM>(static_cast<const float *>(P.a), static_cast<const float *>(P.b), nullptr,
                                           0, 0, static_cast<float *>(P.c), P.ncols, P.nq, P.ncols, P.m,
                                           (uint64_t)jb * BM, (uint64_t)nb * LP_BN, P.flags, sA, sB);
    }
}

// ---------------------------------------------------------------------------
// FP8 e4m3fn with per-slice E8M0 scales: 64 x 64 blocks, K-slices of 128.
// ---------------------------------------------------------------------------
constexpr float FP8_MAX = 448.f;

// Non-finite values take no part in a slice's scale: |x| for finite x, 0 for
// infinities and NaNs (which the comparison sends to the 0 branch).
__device__ __forceinline__ float finite_abs(float x)
{
    float const a = fabsf(x);
    return a < __builtin_inff() ? a : 0.f;
}

// Power-of-two exponent e with amax * 2^-e <= 448 (0 for an all-zero slice;
// amax is a maximum of finite_abs values, so it is finite).
__device__ __forceinline__ int slice_exponent(float amax)
{
    if (!(amax > 0.f) || !(amax < __builtin_inff())) return 0;
    int e = (int)ceilf(log2f(amax / FP8_MAX));
    if (ldexpf(amax, -e) > FP8_MAX) ++e;
    return e < -126 ? -126 : (e > 127 ? 127 : e);
}

// x * 2^-e ready for the conversion: finite values clamped to +-448 (the
// conversion instruction returns NaN beyond +-448 instead of saturating);
// infinities and NaNs become NaN (e4m3fn has no infinity), so they reach
// every entry of C they take part in.
__device__ __forceinline__ float fp8_scaled(float x, int e)
{
    float const v = fminf(fmaxf(ldexpf(x, -e), -FP8_MAX), FP8_MAX);
    return fabsf(x) < __builtin_inff() ? v : __builtin_nanf("");
}

// 4 floats -> 4 e4m3fn bytes scaled by 2^-e.
__device__ __forceinline__ int quantize4(float x0, float x1, float x2, float x3, int e)
{
    int w = __builtin_amdgcn_cvt_pk_fp8_f32(fp8_scaled(x0, e), fp8_scaled(x1, e), 0, false);
    return __builtin_amdgcn_cvt_pk_fp8_f32(fp8_scaled(x2, e), fp8_scaled(x3, e), w, true);
}

// 32 floats -> 32 e4m3fn bytes scaled by 2^-e.
__device__ __forceinline__ void quantize32(const float (&x)[32], int e, i32x8 &out)
{
#pragma unroll
    for (int q = 0; q < 8; ++q) out[q] = quantize4(x[4 * q], x[4 * q + 1], x[4 * q + 2], x[4 * q + 3], e);
}

// A 64 x 128 tile of rows contiguous along l (A in the TN view; n_q % 4 == 0
// and 16-byte aligned rows): each wave instruction reads 2 rows x 512 bytes,
// 8 lanes per 32-wide scale slice, and the slice maximum is reduced over
// those 8 lanes.  Slice c = row * 4 + kb of the tile goes to lanes
// 8 * (c % 8) .. + 7 of wave c / 64 in step (c / 8) % 8; lane s of the 8
// holds l = l0 + 32 * kb + 4 * s .. + 3.  The bytes and scales equal those of
// one thread per row and slice: max is order-independent and each element is
// quantized alone.
struct Fp8Rows
{
    int q[8]; // 4 e4m3fn bytes per step
    int e[8]; // the slice's exponent per step
};

__device__ __forceinline__ void fp8_rows_load(const float *__restrict__ x, uint64_t row0, uint64_t nrows, uint64_t nq,
                                              uint64_t l0, int tid, Fp8Rows &out)
{
    int const lane = tid & 63, wave = tid >> 6;
    f32x4 v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        int const c = (wave * 8 + i) * 8 + (lane >> 3);
        uint64_t const r = row0 + (c >> 2);
        uint64_t const l = l0 + 32 * (c & 3) + 4 * (lane & 7);
        v[i] = (r < nrows && l < nq) ? *reinterpret_cast<const f32x4 *>(x + r * nq + l) : zero4();
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        float amax = fmaxf(fmaxf(finite_abs(v[i][0]), finite_abs(v[i][1])),
                           fmaxf(finite_abs(v[i][2]), finite_abs(v[i][3])));
        amax = fmaxf(amax, __shfl_xor(amax, 1));
        amax = fmaxf(amax, __shfl_xor(amax, 2));
        amax = fmaxf(amax, __shfl_xor(amax, 4));
        out.e[i] = slice_exponent(amax);
        out.q[i] = quantize4(v[i][0], v[i][1], v[i][2], v[i][3], out.e[i]);
    }
}

// B -> e4m3fn bytes and E8M0 scales, once per product (the tile copies them):
// row j at qb + j * ldq (ldq = n_q rounded up to 128), slice s of 32 bytes
// with its scale byte at sb[j * (ldq / 32) + s]; zero (scale 2^0) outside
// m x n_q.  One thread per row and slice.
__global__ __launch_bounds__(256) void k_fp8_split_b(const float *__restrict__ b, unsigned char *__restrict__ qb,
                                                     unsigned char *__restrict__ sb, uint64_t m, uint64_t nq,
                                                     uint64_t bsj, uint64_t bsl, uint64_t ldq, uint64_t rows)
{
    uint64_t const slices = ldq / 32;
    uint64_t const idx = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= rows * slices) return;
    uint64_t const j = idx / slices, sl = idx - j * slices;
    float x[32];
#pragma unroll
    for (int t = 0; t < 32; ++t) {
        uint64_t const l = 32 * sl + t;
        x[t] = (j < m && l < nq) ? b[j * bsj + l * bsl] : 0.f;
    }
    float amax = 0.f;
#pragma unroll
    for (int t = 0; t < 32; ++t) amax = fmaxf(amax, finite_abs(x[t]));
    int const e = slice_exponent(amax);
    i32x8 q;
    quantize32(x, e, q);
    *reinterpret_cast<i32x4 *>(qb + j * ldq + 32 * sl) = {q[0], q[1], q[2], q[3]};
    *reinterpret_cast<i32x4 *>(qb + j * ldq + 32 * sl + 16) = {q[4], q[5], q[6], q[7]};
    sb[j * slices + sl] = (unsigned char)(e + 127);
}

// A -> e4m3fn bytes and scales in the layout of the quantized B (column n
// of the fold as row n, padded to `rows` rows and ldq bytes), once per
// product.  TN view: rows of A are contiguous along l; one 64 x 128 tile per
// block, read 2 rows x 512 bytes per wave instruction (fp8_rows_load).
__global__ __launch_bounds__(256) void k_fp8_split_a_rows(const float *__restrict__ a, unsigned char *__restrict__ qa,
                                                          unsigned char *__restrict__ sa, uint64_t ncols, uint64_t nq,
                                                          uint64_t ldq, uint64_t lblocks)
{
    uint64_t const rb = blockIdx.x / lblocks, l0 = (blockIdx.x - rb * lblocks) * 128;
    uint64_t const row0 = rb * 64;
    int const tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    Fp8Rows r;
    fp8_rows_load(a, row0, ncols, nq, l0, tid, r);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        int const cix = (wave * 8 + i) * 8 + (lane >> 3);
        uint64_t const row = row0 + (cix >> 2);
        int const kb = cix & 3;
        *reinterpret_cast<int *>(qa + row * ldq + l0 + 32 * kb + 4 * (lane & 7)) = r.q[i];
        if ((lane & 7) == 0) sa[row * (ldq / 32) + l0 / 32 + kb] = (unsigned char)(r.e[i] + 127);
    }
}

// NN view: column n = k * w + i of A runs along l with stride w.  One block
// per 64 columns x 128 contraction indices (w % 4 == 0, so 4 consecutive
// columns share a slab and a 16-byte vector): the tile is read row by row in
// 16-byte vectors (16 lanes per 256-byte row), transposed through LDS, and
// quantized and written as in the TN view (eight lanes per 32-wide slice,
// 2 columns x 128 bytes per wave instruction).
__global__ __launch_bounds__(256) void k_fp8_split_a_cols(const float *__restrict__ a, unsigned char *__restrict__ qa,
                                                          unsigned char *__restrict__ sa, uint64_t ncols, uint64_t nq,
                                                          uint64_t w, uint64_t ldq, uint64_t lblocks)
{
    constexpr int LDT = 128 + 4; // floats per transposed column (16-byte aligned rows)
    __shared__ __attribute__((aligned(16))) float sT[64 * LDT];
    uint64_t const cb = blockIdx.x / lblocks, l0 = (blockIdx.x - cb * lblocks) * 128;
    uint64_t const n0 = cb * 64;
    int const tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // load: thread tid covers columns n0 + 4 (tid % 16) .. + 3 of rows l0 + tid / 16 + 16 i
    int const cg = tid & 15;
    uint64_t const n = n0 + 4 * cg;
    bool const ok = n < ncols;
    uint64_t const k = ok ? n / w : 0;
    uint64_t const base = ok ? k * nq * w + (n - k * w) : 0;
    f32x4 v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        uint64_t const l = l0 + (tid >> 4) + 16 * i;
        v[i] = (ok && l < nq) ? *reinterpret_cast<const f32x4 *>(a + base + l * w) : zero4();
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        int const lr = (tid >> 4) + 16 * i;
#pragma unroll
        for (int e = 0; e < 4; ++e) sT[(4 * cg + e) * LDT + lr] = v[i][e];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        int const c = (wave * 8 + i) * 8 + (lane >> 3);
        int const col = c >> 2, kb = c & 3;
        f32x4 const x = *reinterpret_cast<const f32x4 *>(&sT[col * LDT + 32 * kb + 4 * (lane & 7)]);
        float amax = fmaxf(fmaxf(finite_abs(x[0]), finite_abs(x[1])), fmaxf(finite_abs(x[2]), finite_abs(x[3])));
        amax = fmaxf(amax, __shfl_xor(amax, 1));
        amax = fmaxf(amax, __shfl_xor(amax, 2));
        amax = fmaxf(amax, __shfl_xor(amax, 4));
        int const ex = slice_exponent(amax);
        uint64_t const row = n0 + col;
        *reinterpret_cast<int *>(qa + row * ldq + l0 + 32 * kb + 4 * (lane & 7)) = quantize4(x[0], x[1], x[2], x[3], ex);
        if ((lane & 7) == 0) sa[row * (ldq / 32) + l0 / 32 + kb] = (unsigned char)(ex + 127);
    }
}

// The product over the quantized operands: BM rows of B x 64 columns of A
// per block, slices of 128; both tiles are byte copies into LDS.
template <bool TN, int BM>
__global__ __launch_bounds__(256) void k_fp8(const unsigned char *__restrict__ qa, const unsigned char *__restrict__ sa,
                                             const unsigned char *__restrict__ qb, const unsigned char *__restrict__ sb,
                                             uint64_t ldq, float *__restrict__ c, uint64_t ncols, uint64_t nq,
                                             uint64_t w, uint64_t m, unsigned nj, unsigned total)
{
    constexpr int BN = 64, BK = 128;
    constexpr int JT = BM / 32; // 16-row MFMA tiles per wave along j (2 x 2 waves)
    constexpr int LDK = BK + 16; // bytes per row
    __shared__ __attribute__((aligned(16))) unsigned char sA[BN * LDK];
    __shared__ __attribute__((aligned(16))) unsigned char sB[BM * LDK];
    __shared__ unsigned char sAs[BN * 4];
    __shared__ unsigned char sBs[BM * 4];

    unsigned jb, nb;
    lp_block_coords(nj, total, jb, nb);
    uint64_t const j0 = (uint64_t)jb * BM;
    uint64_t const n0 = (uint64_t)nb * BN;
    int const tid = threadIdx.x;
    int const lane = tid & 63, wave = tid >> 6;
    int const lc = lane & 15, lg = lane >> 4;
    int const wm = wave & 1, wn = wave >> 1;

    f32x4 acc[JT][2];
#pragma unroll
    for (int x = 0; x < JT; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = zero4();

    // A: the pre-quantized slice, 64 rows (columns n) x 128 bytes, and its
    // 64 x 4 scales, laid out as B's.  B: BM rows x 128 bytes (16-byte piece
    // p: row p / 8, bytes 16 * (p % 8) of the slice) and BM x 4 scales (scale
    // byte t: row t / 4, slice t % 4).  Slice l0 + BK is loaded into these
    // registers while slice l0 is multiplied.
    i32x4 pa[2];
    unsigned char ea;
    i32x4 pb[BM / 32];
    unsigned char eb[BM / 64];
    auto load = [&](uint64_t l0) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            int const p = tid + 256 * i;
            pa[i] = *reinterpret_cast<const i32x4 *>(qa + (n0 + (p >> 3)) * ldq + l0 + 16 * (p & 7));
        }
        ea = sa[(n0 + (tid >> 2)) * (ldq / 32) + l0 / 32 + (tid & 3)];
#pragma unroll
        for (int i = 0; i < BM / 32; ++i) {
            int const p = tid + 256 * i;
            pb[i] = *reinterpret_cast<const i32x4 *>(qb + (j0 + (p >> 3)) * ldq + l0 + 16 * (p & 7));
        }
#pragma unroll
        for (int i = 0; i < BM / 64; ++i) {
            int const t = tid + 256 * i;
            eb[i] = sb[(j0 + (t >> 2)) * (ldq / 32) + l0 / 32 + (t & 3)];
        }
    };
    load(0);

    for (uint64_t l0 = 0; l0 < nq; l0 += BK) {
        // (the barrier at the end of the previous slice: its fragments have been read)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            int const p = tid + 256 * i;
            *reinterpret_cast<i32x4 *>(&sA[(p >> 3) * LDK + 16 * (p & 7)]) = pa[i];
        }
        sAs[tid] = ea;
#pragma unroll
        for (int i = 0; i < BM / 32; ++i) {
            int const p = tid + 256 * i;
            *reinterpret_cast<i32x4 *>(&sB[(p >> 3) * LDK + 16 * (p & 7)]) = pb[i];
        }
#pragma unroll
        for (int i = 0; i < BM / 64; ++i) sBs[tid + 256 * i] = eb[i];
        __syncthreads();
        if (l0 + BK < nq) load(l0 + BK);
        // Fragments: the lane's 32 bytes are k = 16*lg + [0,16) and
        // k = 64 + 16*lg + [0,16); its scale covers k = 32*lg + [0,32).
        int sbv[JT], sav[2];
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) {
            int const rj = wm * (BM / 2) + jt * 16 + lc;
            i32x4 lo = *reinterpret_cast<const i32x4 *>(&sB[rj * LDK + 16 * lg]);
            i32x4 hi = *reinterpret_cast<const i32x4 *>(&sB[rj * LDK + 64 + 16 * lg]);
            i32x8 fb = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
            sbv[jt] = sBs[rj * 4 + lg];
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                int const rn = wn * 32 + nt * 16 + lc;
                i32x4 alo = *reinterpret_cast<const i32x4 *>(&sA[rn * LDK + 16 * lg]);
                i32x4 ahi = *reinterpret_cast<const i32x4 *>(&sA[rn * LDK + 64 + 16 * lg]);
                i32x8 fa = {alo[0], alo[1], alo[2], alo[3], ahi[0], ahi[1], ahi[2], ahi[3]};
                sav[nt] = sAs[rn * 4 + lg];
                acc[jt][nt] =
                    __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(fb, fa, acc[jt][nt], 0, 0, 0, sbv[jt], 0, sav[nt]);
            }
        }
        // All fragments of the slice have been read before the next slice is
        // stored.  Keep the scale registers allocated until then, so that no
        // later instruction reuses a register a matrix instruction may still
        // read its scale from.
        __syncthreads();
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) asm volatile("" ::"v"(sbv[jt]));
        asm volatile("" ::"v"(sav[0]), "v"(sav[1]));
    }

#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        uint64_t const col = n0 + wn * 32 + nt * 16 + lc;
        if (col >= ncols) continue;
        float *cbase;
        uint64_t stride;
        if constexpr (!TN) {
            uint64_t const k = col / w;
            cbase = c + k * m * w + (col - k * w);
            stride = w;
        } else {
            cbase = c + col * m;
            stride = 1;
        }
#pragma unroll
        for (int jt = 0; jt < JT; ++jt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                uint64_t const jj = j0 + wm * (BM / 2) + jt * 16 + 4 * lg + r;
                if (jj < m) cbase[jj * stride] = acc[jt][nt][r];
            }
    }
}

// ---------------------------------------------------------------------------
// Generic reduced-precision kernel (any shape / alignment): one thread per
// element of C, the same operand rounding, FP32 accumulation.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float fp8_value(float x, int e)
{
    int const wq = __builtin_amdgcn_cvt_pk_fp8_f32(fp8_scaled(x, e), 0.f, 0, false);
    return ldexpf(__builtin_amdgcn_cvt_f32_fp8(wq, 0), e);
}

__global__ __launch_bounds__(256) void k_lowp_generic(const float *__restrict__ a, const float *__restrict__ b,
                                                      float *__restrict__ c, uint64_t outer, uint64_t nq, uint64_t w,
                                                      uint64_t m, uint64_t bsj, uint64_t bsl, int prec,
                                                      const int *__restrict__ exps)
{
    uint64_t const total = outer * m * w;
    uint64_t const step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += step) {
        uint64_t const i = idx % w;
        uint64_t const kj = idx / w;
        uint64_t const j = kj % m;
        uint64_t const k = kj / m;
        float const *ap = a + k * nq * w + i;
        float const *bp = b + j * bsj;
        float s = 0.f;
        for (uint64_t l0 = 0; l0 < nq; l0 += 32) {
            uint64_t const l1 = l0 + 32 < nq ? l0 + 32 : nq;
            int ea = 0, eb = 0;
            if (prec == TTM_PREC_FP8_E4M3) {
                float amax = 0.f, bmax = 0.f;
                for (uint64_t l = l0; l < l1; ++l) {
                    amax = fmaxf(amax, finite_abs(ap[l * w]));
                    bmax = fmaxf(bmax, finite_abs(bp[l * bsl]));
                }
                ea = slice_exponent(amax);
                eb = slice_exponent(bmax);
            }
            for (uint64_t l = l0; l < l1; ++l) {
                float const x = ap[l * w], y = bp[l * bsl];
                switch (prec) {
                case TTM_PREC_BF16: s = fmaf((float)(__bf16)y, (float)(__bf16)x, s); break;
                case TTM_PREC_FP16: s = fmaf((float)(_Float16)y, (float)(_Float16)x, s); break;
                case TTM_PREC_FP8_E4M3: s = fmaf(fp8_value(y, eb), fp8_value(x, ea), s); break;
                case TTM_PREC_FP16X2: {
                    _Float16 px[2], py[2];
                    Lp<LP_FP16X2>::split(ldexpf(x, exps[0]), px);
                    Lp<LP_FP16X2>::split(ldexpf(y, exps[1]), py);
                    float t = (float)py[0] * (float)px[1] + (float)py[1] * (float)px[0];
                    s += t + (float)py[0] * (float)px[0];
                    break;
                }
                default: {
                    __bf16 px[3], py[3];
                    Lp<LP_BF16X3>::split(x, px);
                    Lp<LP_BF16X3>::split(y, py);
                    float t = (float)py[1] * (float)px[1] + (float)py[0] * (float)px[2] + (float)py[2] * (float)px[0];
                    t += (float)py[0] * (float)px[1] + (float)py[1] * (float)px[0];
                    s += t + (float)py[0] * (float)px[0];
                }
                }
            }
        }
        c[idx] = prec == TTM_PREC_FP16X2 ? ldexpf(s, -(exps[0] + exps[1])) : s;
    }
}

// Block height of the single-product tiles.
int lp_block_rows(const Fold &f) { return f.m <= 64 ? 64 : 128; }

// Convert B once into padded planes (ldb = nq rounded up to 32, rows = m
// rounded up to the block height), then run the blocks; the workspace is
// released in stream order after the product.
// FP16X2 scale pass: grid sizes and the workspace it needs (partial maxima,
// then the two exponents).
struct ScalePass
{
    unsigned blocks_a = 0, blocks = 0;
    size_t bytes() const { return (size_t)blocks * sizeof(float) + 2 * sizeof(int); }
};
ScalePass scale_pass(const Fold &f, bool with_a)
{
    auto nblk = [](uint64_t n, uint64_t cap) {
        uint64_t const b = (n / 4 + 255) / 256;
        return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
    };
    ScalePass sp;
    sp.blocks_a = with_a ? nblk(f.numel_a(), 2048) : 0;
    sp.blocks = sp.blocks_a + nblk(f.numel_b(), 64);
    return sp;
}
// Launch the scale pass into ws (partials first, exponents after them);
// returns the exponents' address.
int *launch_scale_pass(const Fold &f, const float *a, const float *b, const ScalePass &sp, void *ws, hipStream_t s)
{
    float *const part = static_cast<float *>(ws);
    int *const exps = reinterpret_cast<int *>(part + sp.blocks);
    hipLaunchKernelGGL(k_absmax_partial, dim3(sp.blocks), dim3(256), 0, s, a, f.numel_a(), b, f.numel_b(),
                       sp.blocks_a, part);
    hipLaunchKernelGGL(k_absmax_final, dim3(1), dim3(256), 0, s, part, sp.blocks_a, sp.blocks, exps);
    return exps;
}

// Convert B once into padded planes (ldb = nq rounded up to 32, rows = m
// rounded up to the block height), then run the blocks; the workspace is
// released in stream order after the product.  FP16X2 first finds the two
// scale exponents (kept after the planes in the same workspace).
template <int MODE, bool TN>
ttm_status_t launch_lp(const Fold &f, const float *a, const float *b, float *c, hipStream_t s)
{
    using op = typename Lp<MODE>::op;
    constexpr int P = Lp<MODE>::PARTS;
    constexpr bool SCALED = MODE == LP_FP16X2;
    uint64_t const bm = (uint64_t)lp_block_rows(f);
    uint64_t const ncols = TN ? f.outer : f.outer * f.w;
    unsigned const nj = (unsigned)((f.m + bm - 1) / bm);
    unsigned const total = nj * (unsigned)((ncols + 127) / 128);
    uint64_t const ldb = (f.nq + 31) / 32 * 32;
    uint64_t const rows = (uint64_t)nj * bm;
    uint64_t const plane = rows * ldb;
    uint64_t const split_blocks = (rows * (ldb / 8) + 255) / 256;
    if (split_blocks >= (1ull << 31)) {
        set_error(TTM_ERR_ARGUMENT, "B too large for the reduced-precision kernels");
        return TTM_ERR_ARGUMENT;
    }
    size_t const planes_bytes = (P * plane * sizeof(op) + 255) / 256 * 256;
    // 64-row blocks scale A per block and slice themselves: B's exponent only
    ScalePass const sp = SCALED ? scale_pass(f, bm != 64) : ScalePass();
    void *ws = nullptr;
    if (ttm_status_t const st = ws_acquire(planes_bytes + (SCALED ? sp.bytes() : 0), s, &ws,
                                           "workspace for the converted B operand");
        st != TTM_SUCCESS)
        return st;
    op *const bs = static_cast<op *>(ws);
    int const *exps = nullptr;
    if constexpr (SCALED) exps = launch_scale_pass(f, a, b, sp, static_cast<char *>(ws) + planes_bytes, s);
    hipLaunchKernelGGL((k_lowp_split_b<MODE>), dim3((unsigned)split_blocks), dim3(256), 0, s, b, bs, f.m, f.nq,
                       f.b_stride_j(), f.b_stride_l(), ldb, rows, plane, exps);
    if (bm == 64)
        hipLaunchKernelGGL((k_lowp<MODE, TN, 64>), dim3(total), dim3(256), 0, s, a, bs, ldb, plane, c, ncols, f.nq,
                           f.w, f.m, nj, total, exps);
    else
        hipLaunchKernelGGL((k_lowp<MODE, TN, 128>), dim3(total), dim3(256), 0, s, a, bs, ldb, plane, c, ncols, f.nq,
                           f.w, f.m, nj, total, exps);
    hipError_t const e = hipGetLastError();
    ws_release(ws, s);
    if (e != hipSuccess) {
        set_error(TTM_ERR_HIP, hipGetErrorString(e));
        return TTM_ERR_HIP;
    }
    return TTM_SUCCESS;
}

// Rows of B per FP8 block: 128 where B has more than 64 rows (each block's
// copy of A's slice then serves twice as many rows of C), else 64.
int fp8_block_rows(const Fold &f) { return f.m > 64 ? 128 : 64; }

// A and B quantized once per product (k_fp8_split_a_*, k_fp8_split_b) into
// pooled workspace: N_A + m n_q bytes (rounded up to whole blocks and
// slices) plus a scale byte per 32 of them; then the tiles.
template <bool TN>
ttm_status_t launch_fp8(const Fold &f, const float *a, const float *b, float *c, hipStream_t s)
{
    uint64_t const bm = (uint64_t)fp8_block_rows(f);
    uint64_t const ncols = TN ? f.outer : f.outer * f.w;
    unsigned const nj = (unsigned)((f.m + bm - 1) / bm);
    unsigned const total = nj * (unsigned)((ncols + 63) / 64);
    uint64_t const ldq = (f.nq + 127) / 128 * 128;
    uint64_t const rows = (uint64_t)nj * bm;
    uint64_t const rows_a = (ncols + 63) / 64 * 64;
    uint64_t const split_blocks = (rows * (ldq / 32) + 255) / 256;
    uint64_t const split_a_blocks = (rows_a / 64) * (ldq / 128);
    if (split_blocks >= (1ull << 31) || split_a_blocks >= (1ull << 31)) {
        set_error(TTM_ERR_ARGUMENT, "operands too large for the FP8 kernels");
        return TTM_ERR_ARGUMENT;
    }
    auto round256 = [](uint64_t x) { return (x + 255) / 256 * 256; };
    size_t const bytes_qb = round256(rows * ldq), bytes_sb = round256(rows * (ldq / 32));
    size_t const bytes_qa = round256(rows_a * ldq), bytes_sa = round256(rows_a * (ldq / 32));
    void *ws = nullptr;
    if (ttm_status_t const st = ws_acquire(bytes_qb + bytes_sb + bytes_qa + bytes_sa, s, &ws,
                                           "workspace for the quantized operands");
        st != TTM_SUCCESS)
        return st;
    unsigned char *const qb = static_cast<unsigned char *>(ws);
    unsigned char *const sb = qb + bytes_qb;
    unsigned char *const qa = sb + bytes_sb;
    unsigned char *const sa = qa + bytes_qa;
    hipLaunchKernelGGL(k_fp8_split_b, dim3((unsigned)split_blocks), dim3(256), 0, s, b, qb, sb, f.m, f.nq,
                       f.b_stride_j(), f.b_stride_l(), ldq, rows);
    if constexpr (TN)
        hipLaunchKernelGGL(k_fp8_split_a_rows, dim3((unsigned)split_a_blocks), dim3(256), 0, s, a, qa, sa, ncols,
                           f.nq, ldq, ldq / 128);
    else
        hipLaunchKernelGGL(k_fp8_split_a_cols, dim3((unsigned)split_a_blocks), dim3(256), 0, s, a, qa, sa, ncols,
                           f.nq, f.w, ldq, ldq / 128);
    if (bm == 128)
        hipLaunchKernelGGL((k_fp8<TN, 128>), dim3(total), dim3(256), 0, s, qa, sa, qb, sb, ldq, c, ncols, f.nq, f.w,
                           f.m, nj, total);
    else
        hipLaunchKernelGGL((k_fp8<TN, 64>), dim3(total), dim3(256), 0, s, qa, sa, qb, sb, ldq, c, ncols, f.nq, f.w,
                           f.m, nj, total);
    hipError_t const e = hipGetLastError();
    ws_release(ws, s);
    if (e != hipSuccess) {
        set_error(TTM_ERR_HIP, hipGetErrorString(e));
        return TTM_ERR_HIP;
    }
    return TTM_SUCCESS;
}

template <int MODE>
ttm_status_t launch_lp_mode(const Fold &f, const float *a, const float *b, float *c, hipStream_t s)
{
    return f.w == 1 ? launch_lp<MODE, true>(f, a, b, c, s) : launch_lp<MODE, false>(f, a, b, c, s);
}

} // namespace

bool lowp_supports(const Fold &f, bool aligned)
{
    if (!aligned) return false;
    bool const shape_ok = f.w == 1 ? f.nq % 4 == 0 : f.w % 4 == 0;
    if (!shape_ok) return false;
    uint64_t const ncols = f.w == 1 ? f.outer : f.outer * f.w;
    uint64_t const blocks = ((f.m + 63) / 64) * ((ncols + 63) / 64);
    return blocks < (1ull << 31);
}

std::string lowp_symbol(const Fold &f, ttm_precision_t prec, bool generic)
{
    char const *tn = f.w == 1 ? "true" : "false";
    char buf[96];
    if (generic) return "k_lowp_generic";
    if (prec == TTM_PREC_FP8_E4M3) {
        std::snprintf(buf, sizeof buf, "k_fp8<%s, %d>", tn, fp8_block_rows(f));
    } else {
        int const mode = prec == TTM_PREC_BF16 ? LP_BF16
                         : prec == TTM_PREC_FP16 ? LP_FP16
                         : prec == TTM_PREC_FP16X2 ? LP_FP16X2
                                                   : LP_BF16X3;
        std::snprintf(buf, sizeof buf, "k_lowp<%d, %s, %d>", mode, tn, lp_block_rows(f));
    }
    return buf;
}

ttm_status_t launch_lowp(const Fold &f, const float *a, const float *b, float *c, ttm_precision_t prec,
                         hipStream_t s, bool generic)
{
    if (generic) {
        uint64_t const total = f.numel_c();
        uint64_t blocks = (total + 255) / 256;
        if (blocks > (1u << 20)) blocks = 1u << 20;
        void *ws = nullptr;
        int const *exps = nullptr;
        if (prec == TTM_PREC_FP16X2) {
            ScalePass const sp = scale_pass(f, true);
            if (ttm_status_t const st = ws_acquire(sp.bytes(), s, &ws, "workspace for the FP16X2 scales");
                st != TTM_SUCCESS)
                return st;
            exps = launch_scale_pass(f, a, b, sp, ws, s);
        }
        hipLaunchKernelGGL(k_lowp_generic, dim3((unsigned)blocks), dim3(256), 0, s, a, b, c, f.outer, f.nq, f.w,
                           f.m, f.b_stride_j(), f.b_stride_l(), (int)prec, exps);
        ws_release(ws, s);
    } else if (prec == TTM_PREC_FP8_E4M3) {
        return f.w == 1 ? launch_fp8<true>(f, a, b, c, s) : launch_fp8<false>(f, a, b, c, s);
    } else if (prec == TTM_PREC_BF16) {
        return launch_lp_mode<LP_BF16>(f, a, b, c, s);
    } else if (prec == TTM_PREC_FP16) {
        return launch_lp_mode<LP_FP16>(f, a, b, c, s);
    } else if (prec == TTM_PREC_FP16X2) {
        return launch_lp_mode<LP_FP16X2>(f, a, b, c, s);
    } else {
        return launch_lp_mode<LP_BF16X3>(f, a, b, c, s);
    }
    hipError_t const e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(TTM_ERR_HIP, hipGetErrorString(e));
        return TTM_ERR_HIP;
    }
    return TTM_SUCCESS;
}

} // namespace ttm_amd

namespace ttm_amd
{

bool lowp_group_precision(ttm_precision_t prec) { return prec == TTM_PREC_BF16X3 || prec == TTM_PREC_FP16X2; }

namespace
{
const void *lowp_group_kernel(ttm_precision_t prec, int bm)
{
    if (prec == TTM_PREC_FP16X2)
        return bm == 64 ? reinterpret_cast<const void *>(&k_lowp_grouped<LP_FP16X2, 64>)
                        : reinterpret_cast<const void *>(&k_lowp_grouped<LP_FP16X2, 128>);
    return bm == 64 ? reinterpret_cast<const void *>(&k_lowp_grouped<LP_BF16X3, 64>)
                    : reinterpret_cast<const void *>(&k_lowp_grouped<LP_BF16X3, 128>);
}
} // namespace

int lowp_group_occupancy(ttm_precision_t prec, int bm)
{
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, lowp_group_kernel(prec, bm), 256, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return blocks;
}

std::string lowp_group_symbol(ttm_precision_t prec, int bm)
{
    return std::string("k_lowp_grouped<") + (prec == TTM_PREC_FP16X2 ? "3" : "2") + (bm == 64 ? ", 64>" : ", 128>");
}

ttm_status_t launch_lowp_grouped(ttm_precision_t prec, int bm, const GroupProblem *d_probs,
                                 const unsigned *d_tile_prob, unsigned total, unsigned grid, hipStream_t s)
{
    if (prec == TTM_PREC_FP16X2) {
        if (bm == 64)
            hipLaunchKernelGGL((k_lowp_grouped<LP_FP16X2, 64>), dim3(grid), dim3(256), 0, s, d_probs, d_tile_prob,
                               total);
        else
            hipLaunchKernelGGL((k_lowp_grouped<LP_FP16X2, 128>), dim3(grid), dim3(256), 0, s, d_probs, d_tile_prob,
                               total);
    } else {
        if (bm == 64)
            hipLaunchKernelGGL((k_lowp_grouped<LP_BF16X3, 64>), dim3(grid), dim3(256), 0, s, d_probs, d_tile_prob,
                               total);
        else
            hipLaunchKernelGGL((k_lowp_grouped<LP_BF16X3, 128>), dim3(grid), dim3(256), 0, s, d_probs, d_tile_prob,
                               total);
    }
    hipError_t const e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(TTM_ERR_HIP, hipGetErrorString(e));
        return TTM_ERR_HIP;
    }
    return TTM_SUCCESS;
}

} // namespace ttm_amd
