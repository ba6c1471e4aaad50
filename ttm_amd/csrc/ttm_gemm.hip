// LDS-staged MFMA kernels for the folded product (plain FP32 / FP64)
//
//     D[j][n] = sum_l B(j, l) * At[l][n]       j < m, n < ncols, l < nq
//
// with two views of the column index n:
//   NN (any w):   n = k*wp + i (i < wp, the columns i >= w of each slab are
//                 padding, wp = w rounded up to the vector width unless w is
//                 a multiple of it), At[l][n] = A[k*nq*w + l*w + i], C at
//                 k*m*w + j*w + i
//   TN (w == 1):  n = k,  At[l][n] = A[k*nq + l],  C at k*m + j
//
// A block may cover only the contraction range [l0, l0 + len) of a split
// product (split-K): A and B are then offset to l0 by the caller, `nq` is the
// block's length and `ld` the full extent that the strides are made of.
//
// A workgroup of 4 waves (WAVES_M x WAVES_N) computes a BM x BN block of D,
// stepping through the contraction in BK-slices that are staged in LDS with
// double buffering (one barrier per slice; the next slice's global loads are
// in flight while the current one is multiplied).
//
// Fragment interleaving.  A wave owns 16*JT rows and 16*NT columns.  MFMA
// tile jt holds the rows {JT*x + jt : x < 16} and tile nt the columns
// {NT*y + nt : y < 16}, so the JT (NT) fragment values a lane needs for one
// k are adjacent in LDS and arrive with one vector read; and a lane's NT
// results of one row are adjacent columns (NN: one 16-byte store) and its JT
// results of one column adjacent rows (TN: one 16-byte store).
//
// MFMA lane conventions (16x16x4, lane = 16*lr + lc): srcA holds
// Afrag[lc][lr], srcB holds Bfrag[lr][lc], accumulator element r holds
// D[acc_row(lr, r)][lc] with acc_row = 4*lr + r (f32) or 4*r + lr (f64).
#include "ttm_gemm.hpp"
#include "ttm_workspace.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>

namespace ttm_amd
{
namespace
{

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
// The same 16-byte vectors at element alignment: rows of A or B whose length
// is not a multiple of the vector width start at element-aligned addresses;
// dword-aligned multi-dword global loads are a supported access (one
// global_load_dwordx4).
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));
typedef double f64x2u __attribute__((ext_vector_type(2), aligned(8)));

template <class T>
struct G;
template <>
struct G<float>
{
    static constexpr int VEC = 4;
    using vec = f32x4;
    using vec_u = f32x4u;
    using acc = f32x4;
    __device__ static constexpr int acc_row(int lr, int r) { return 4 * lr + r; }
    __device__ static f32x4 mfma(float a, float b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};
template <>
struct G<double>
{
    static constexpr int VEC = 2;
    using vec = f64x2;
    using vec_u = f64x2u;
    using acc = f64x4;
    __device__ static constexpr int acc_row(int lr, int r) { return 4 * r + lr; }
    __device__ static f64x4 mfma(double a, double b, f64x4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
};

template <class V>
__device__ __forceinline__ V vzero()
{
    V v;
#pragma unroll
    for (int e = 0; e < (int)(sizeof(V) / sizeof(v[0])); ++e) v[e] = 0;
    return v;
}

// Read N consecutive T from LDS with 16-byte (or 8-byte) vector reads.
template <class T, int N>
__device__ __forceinline__ void lds_read(const T *p, T (&out)[N])
{
    constexpr int VEC = G<T>::VEC;
    if constexpr (N % VEC == 0) {
#pragma unroll
        for (int e = 0; e < N; e += VEC) {
            auto v = *reinterpret_cast<const typename G<T>::vec *>(p + e);
#pragma unroll
            for (int x = 0; x < VEC; ++x) out[e + x] = v[x];
        }
    } else {
#pragma unroll
        for (int e = 0; e < N; ++e) out[e] = p[e];
    }
}

// Global memory accesses of the tiles.  The grouped kernels take A, B and C
// from problem descriptors in memory, so the compiler cannot tell that they
// point to global memory and would emit FLAT instructions, which also count
// against the LDS counter: every LDS wait before a barrier then waits for
// all loads in flight.  The explicit address space keeps them global loads.
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
template <class V>
__device__ __forceinline__ void stg_nt(void *p, V v)
{
    __builtin_nontemporal_store(v, (__attribute__((address_space(1))) V *)p);
}

// Elements d .. d + VEC - 1 of v moved to 0 .. VEC - 1, the first nvalid
// kept and the rest zeroed (element x + d < VEC whenever x < nvalid).
// Selects only: no branch, no partial register write.
template <class V>
__device__ __forceinline__ V shift_mask(V v, int d, int nvalid)
{
    constexpr int N = (int)(sizeof(V) / sizeof(v[0]));
    V out;
#pragma unroll
    for (int x = 0; x < N; ++x) {
        auto r = v[x];
#pragma unroll
        for (int y = 1; x + y < N; ++y) r = d == y ? v[x + y] : r;
        out[x] = x < nvalid ? r : decltype(r)(0);
    }
    return out;
}

// Workgroup id -> (row block, column block).  Consecutive workgroup ids are
// dispatched round-robin over the 8 XCDs; regroup them so that the row
// blocks sharing one column block run on the same XCD (shared L2 for that
// slice of A).  Row blocks vary fastest.
__device__ __forceinline__ void block_coords(unsigned nj, unsigned total, unsigned &jb, unsigned &nb)
{
    unsigned b = blockIdx.x;
    unsigned const full = total & ~7u;
    if (b < full) b = (b & 7u) * (full >> 3) + (b >> 3);
    jb = b % nj;
    nb = b / nj;
}

// Views of the column index (see the file header).  TT loads like TN (rows of
// A contiguous along l) and stores like NN with a single slab of width ncols:
// C[j * ncols + n], i.e. the new mode becomes the slowest one.
enum View : int
{
    VIEW_NN = 0,
    VIEW_TN = 1,
    VIEW_TT = 2
};
template <class T, int JT, int NT, int WAVES_M, int BK>
struct TileShape
{
    static constexpr int VEC = G<T>::VEC;
    static constexpr int BM = WAVES_M * 16 * JT;
    static constexpr int BN = (4 / WAVES_M) * 16 * NT;
    static constexpr int LDA = BN + VEC; // At tile row stride (one 16-byte vector of padding)
    static constexpr int LDB = BM + VEC; // B tile row stride
};

// One BM x BN block of D starting at (j0, n0); sA / sB are the block's two
// LDS buffers per operand.  Begins and ends with all waves synchronised.
// PAD (NN view only): slabs of any width w and element-aligned A and C (a_len:
// elements of A from a to its end; A holds at least VEC); off, the NN view requires
// w % VEC == 0 and 16-byte aligned A and C (wp == w).
template <class T, int JT, int NT, int WAVES_M, int BK, int VIEW, bool BCOL, bool PAD = false>
__device__ __forceinline__ void gemm_tile(const T *__restrict__ a, const T *__restrict__ b, T *__restrict__ c,
                                          uint64_t ncols, uint64_t nq, uint64_t ld, uint64_t w, uint64_t wp,
                                          uint64_t m, uint64_t j0, uint64_t n0, unsigned flags, uint64_t a_len,
                                          T (*sA)[BK * TileShape<T, JT, NT, WAVES_M, BK>::LDA],
                                          T (*sB)[BK * TileShape<T, JT, NT, WAVES_M, BK>::LDB])
{
    using Gt = G<T>;
    using V = typename Gt::vec;
    using Acc = typename Gt::acc;
    using S = TileShape<T, JT, NT, WAVES_M, BK>;
    constexpr int VEC = Gt::VEC;
    constexpr int BM = S::BM;
    constexpr int BN = S::BN;
    constexpr int LDA = S::LDA;
    constexpr int LDB = S::LDB;
    constexpr bool TN = VIEW != VIEW_NN;     // rows of A contiguous along l
    constexpr bool TSTORE = VIEW == VIEW_TN; // C[n][j]
    static_assert(BK % 4 == 0 && BK % VEC == 0, "BK");
    bool const a_vec = (flags & TILE_A_VEC) != 0;
    bool const c_vec = (flags & TILE_C_VEC) != 0;
    bool const stream = (flags & TILE_STREAM) != 0;
    bool const nt_c = stream && (flags & TILE_KEEP_C) == 0;
    // 16-byte loads of A: non-temporal when A is streamed once (measured on
    // MI355X: ~7.1 TB/s for non-temporal against ~6.2 TB/s for ordinary
    // 16-byte read streams).  In the TN/TT views a slice reads BK * sizeof(T)
    // bytes of each row: only when that is a whole number of 128-byte lines
    // and rows start on a line does a slice consume its lines completely;
    // otherwise consecutive slices share lines and a non-temporal line is gone
    // before its second half is read (FETCH_SIZE 1.33x the bytes of A).
    constexpr bool TN_WHOLE_LINES = BK * sizeof(T) % 128 == 0;
    bool const nt_a = stream && (!TN || (TN_WHOLE_LINES && ld * sizeof(T) % 128 == 0));
    auto load_a = [&](const T *p) -> V {
        return nt_a ? ldg_nt<V>(p) : ldg<V>(p);
    };

    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const lc = lane & 15;
    int const lr = lane >> 4;
    int const wm = wave % WAVES_M;
    int const wn = wave / WAVES_M;

    // ---- global -> register loaders ------------------------------------
    // At tile: BK x BN.  NN: rows of BN contiguous columns (vectors along n).
    //                     TN: BN columns of BK contiguous l (vectors along l).
    constexpr int A_VECS = BK * BN / VEC;
    constexpr int A_PER = (A_VECS + 255) / 256;
    // B tile: BK x BM.  BCOL: B(j,l) at j + l*m (vectors along j);
    //                   else  at j*nq + l (vectors along l).
    constexpr int B_VECS = BK * BM / VEC;
    constexpr int B_PER = (B_VECS + 255) / 256;

    // per-thread fixed part of the A addresses; NN vectors are VEC padded
    // columns of one slab, of which a_nv are real (i < w)
    uint64_t a_base[A_PER];
    bool a_ok[A_PER];
    int a_row[A_PER], a_col[A_PER], a_nv[A_PER];
#pragma unroll
    for (int e = 0; e < A_PER; ++e) {
        int const v = tid + 256 * e;
        a_ok[e] = v < A_VECS;
        a_nv[e] = VEC;
        if constexpr (!TN) {
            a_row[e] = v / (BN / VEC);
            a_col[e] = (v % (BN / VEC)) * VEC;
            uint64_t const n = n0 + a_col[e];
            a_ok[e] = a_ok[e] && n < ncols;
            if constexpr (PAD) {
                uint64_t const k = a_ok[e] ? n / wp : 0;
                uint64_t const i = a_ok[e] ? n - k * wp : 0;
                a_ok[e] = a_ok[e] && i < w;
                a_nv[e] = a_ok[e] ? (int)(w - i < (uint64_t)VEC ? w - i : VEC) : 0;
                a_base[e] = a_ok[e] ? k * ld * w + i : 0;
            } else {
                uint64_t const k = a_ok[e] ? n / w : 0;
                a_base[e] = k * ld * w + (a_ok[e] ? n - k * w : 0);
            }
        } else {
            a_col[e] = v / (BK / VEC);          // column n within the tile
            a_row[e] = (v % (BK / VEC)) * VEC;  // first l within the slice
            uint64_t const n = n0 + a_col[e];
            a_ok[e] = a_ok[e] && n < ncols;
            a_base[e] = a_ok[e] ? n * ld : 0;
        }
    }
    int b_row[B_PER], b_col[B_PER];
    bool b_in[B_PER];
#pragma unroll
    for (int e = 0; e < B_PER; ++e) {
        int const v = tid + 256 * e;
        b_in[e] = v < B_VECS;
        if constexpr (BCOL) {
            b_row[e] = v / (BM / VEC);
            b_col[e] = (v % (BM / VEC)) * VEC;
        } else {
            b_col[e] = v / (BK / VEC);
            b_row[e] = (v % (BK / VEC)) * VEC;
        }
    }
    bool const b_vec_ok = (flags & TILE_B_VEC) != 0 && (BCOL ? (m % VEC == 0) : (ld % VEC == 0));

    // Register sets of slices in flight: two for tiles of at most 32 rows per
    // wave (little matrix work per slice to hide a load behind: the loads of
    // slice s + 2 are issued while slice s is multiplied) whose slice is at
    // most five vectors per thread, or the 16-row TN tile of one product; one
    // otherwise (larger sets would spill in the grouped kernels).
    constexpr int PF = (JT <= 2 && (A_PER + B_PER <= 5 || (VIEW == VIEW_TN && JT == 1))) ? 2 : 1;
    V ra[PF][A_PER], rb[PF][B_PER];
    // Guarded slices: shift d and count nvalid of each vector (d | nvalid << 8),
    // applied when the slice is stored so that no load is waited for early.
    int ra_dm[PF][A_PER], rb_dm[PF][B_PER];
    constexpr int DM_WHOLE = VEC << 8; // no shift, all VEC elements valid

    // Interior tile (every row and column in range, 16-byte vectors
    // throughout): the whole slices run with unguarded loads, all in flight
    // together, and only a ragged last slice (nq % BK != 0) is guarded.  With
    // PAD the A vectors are element-aligned; the tile is interior when every
    // one of them is whole (no slab row end inside a vector).
    // TN rows and B at element alignment take element-aligned 16-byte loads
    // (every vector of a whole slice lies inside its row or column); the TT
    // tiles of the grouped kernels keep the aligned-only rule.
    // Rows past m (the last row block of a ragged m) read row m - 1 of a
    // row-major B, or the column's last whole vector of a column-major B
    // (which needs m % VEC == 0, so that no vector is part in and part out
    // of range), and their products are never stored: such blocks keep the
    // unguarded loop.
    bool const rows_ok = j0 + BM <= m || !BCOL || m % VEC == 0;
    bool full = n0 + BN <= ncols && rows_ok &&
                (VIEW == VIEW_TT ? (a_vec && b_vec_ok) : (TN || a_vec || !PAD));
    if constexpr (PAD) {
        if (!a_vec) {
            bool whole = n0 + BN <= ncols && rows_ok;
#pragma unroll
            for (int e = 0; e < A_PER; ++e) whole = whole && a_nv[e] == VEC;
            full = __syncthreads_and(whole ? 1 : 0) != 0;
        }
    }
    uint64_t a_fix[A_PER]; // A offset of the thread's vector in slice 0
#pragma unroll
    for (int e = 0; e < A_PER; ++e) a_fix[e] = TN ? a_base[e] + a_row[e] : a_base[e] + (uint64_t)a_row[e] * w;
    uint64_t const a_step = TN ? 1 : w; // A offset per contraction index

    auto load_tiles = [&](uint64_t l0, auto set) {
        constexpr int S = decltype(set)::value;
#pragma unroll
        for (int e = 0; e < A_PER; ++e) {
            ra[S][e] = vzero<V>();
            ra_dm[S][e] = DM_WHOLE;
            if constexpr (!TN) {
                uint64_t const l = l0 + a_row[e];
                if (!PAD || a_vec) {
                    if (a_ok[e] && l < nq) ra[S][e] = load_a(a + a_base[e] + l * w);
                } else {
                    // Branch-free: the VEC elements from the vector's first
                    // (element-aligned) address; rows of a slab are
                    // contiguous, so elements past the slab row's end belong
                    // to the next row and are masked.  Only at the very end of
                    // A is the read moved back and shifted into place (d).
                    // Invalid vectors and rows read in-bounds addresses and
                    // keep no element (nvalid 0).
                    // (a_len may be below VEC for the last rows of a split:
                    // the read then starts before `a`, still inside A, which
                    // holds at least VEC elements)
                    uint64_t const lc = l < nq ? l : nq - 1;
                    int64_t const addr = (int64_t)(a_base[e] + lc * w);
                    int64_t const lim = (int64_t)a_len - VEC;
                    int64_t const ac = addr < lim ? addr : lim;
                    V const v = ldg<typename Gt::vec_u>(a + ac);
                    ra[S][e] = v;
                    ra_dm[S][e] = (int)(addr - ac) | (l < nq ? a_nv[e] : 0) << 8;
                }
            } else if (nq >= (uint64_t)VEC) {
                // One in-bounds vector per thread and no divergent branch (a
                // per-element fill of a vector register whose load is still in
                // flight makes the compiler wait for every load): a vector
                // that would run past the row end is read as the row's last
                // VEC elements and shifted into place; elements past the row
                // end, and rows past the last column (read as row 0), are zero.
                uint64_t const l = l0 + a_row[e];
                uint64_t const last = nq - VEC;
                uint64_t const lc = l < last ? l : last;
                V const v = ldg<typename Gt::vec_u>(a + a_base[e] + lc);
                int const nvalid = (!a_ok[e] || l >= nq) ? 0 : (int)(nq - l < (uint64_t)VEC ? nq - l : VEC);
                ra[S][e] = v;
                ra_dm[S][e] = (int)(l - lc) | nvalid << 8;
            } else {
                uint64_t const l = l0 + a_row[e];
                if (a_ok[e] && l < nq) {
                    if (a_vec) {
                        // nq % VEC == 0: a vector is wholly in or out of range
                        ra[S][e] = load_a(a + a_base[e] + l);
                    } else if (l + VEC <= nq) {
                        // element-aligned vector inside the row
                        V const v = ldg<typename Gt::vec_u>(a + a_base[e] + l);
                        ra[S][e] = v;
                    } else {
                        // the row's last, partial vector
#pragma unroll
                        for (int x = 0; x < VEC; ++x)
                            if (l + x < nq) ra[S][e][x] = ldg<T>(a + a_base[e] + l + x);
                    }
                }
            }
        }
#pragma unroll
        for (int e = 0; e < B_PER; ++e) {
            rb[S][e] = vzero<V>();
            rb_dm[S][e] = DM_WHOLE;
            if (!b_in[e]) continue;
            if constexpr (BCOL) {
                uint64_t const l = l0 + b_row[e];
                uint64_t const j = j0 + b_col[e];
                if (l < nq) {
                    if (j + VEC <= m) {
                        // element-aligned vector inside column l (16-byte aligned when m % VEC == 0)
                        V const v = ldg<typename Gt::vec_u>(b + j + l * m);
                        rb[S][e] = v;
                    } else {
#pragma unroll
                        for (int x = 0; x < VEC; ++x)
                            if (j + x < m) rb[S][e][x] = ldg<T>(b + j + x + l * m);
                    }
                }
            } else if (nq >= (uint64_t)VEC) {
                // as the TN rows of A; rows past m read row m - 1 (their
                // products are never stored)
                uint64_t const j = j0 + b_col[e];
                uint64_t const jc = j < m ? j : m - 1;
                uint64_t const l = l0 + b_row[e];
                uint64_t const last = nq - VEC;
                uint64_t const lc = l < last ? l : last;
                V const v = ldg<typename Gt::vec_u>(b + jc * ld + lc);
                int const nvalid = l >= nq ? 0 : (int)(nq - l < (uint64_t)VEC ? nq - l : VEC);
                rb[S][e] = v;
                rb_dm[S][e] = (int)(l - lc) | nvalid << 8;
            } else {
                uint64_t const j = j0 + b_col[e];
                uint64_t const l = l0 + b_row[e];
                if (j < m) {
                    if (l + VEC <= nq) {
                        // element-aligned vector inside row j (16-byte aligned when ld % VEC == 0)
                        V const v = ldg<typename Gt::vec_u>(b + j * ld + l);
                        rb[S][e] = v;
                    } else {
#pragma unroll
                        for (int x = 0; x < VEC; ++x)
                            if (l + x < nq) rb[S][e][x] = ldg<T>(b + j * ld + l + x);
                    }
                }
            }
        }
    };

    // Interior tiles: the same loads without guards or branches, so all of a
    // slice's loads stay in flight while the previous slice is multiplied.
    // Every thread owns whole A vectors (A_VECS is a multiple of 256); B
    // vectors past B_VECS (BM * BK < 256 * VEC) re-read the tile's first
    // vector, B(j0, l0), which is in range, and are not stored.
    static_assert(A_VECS % 256 == 0, "every thread loads whole A vectors");
    auto load_full = [&](uint64_t l0, auto set) {
        constexpr int S = decltype(set)::value;
#pragma unroll
        for (int e = 0; e < A_PER; ++e) {
            if constexpr (PAD) {
                if (!a_vec) {
                    // element-aligned whole vectors
                    using VU = typename Gt::vec_u;
                    ra[S][e] = nt_a ? (V)ldg_nt<VU>(a + a_fix[e] + l0 * a_step) : (V)ldg<VU>(a + a_fix[e] + l0 * a_step);
                    continue;
                }
            }
            if constexpr (TN && VIEW != VIEW_TT) {
                if (!a_vec) {
                    ra[S][e] = ldg<typename Gt::vec_u>(a + a_fix[e] + l0 * a_step);
                    continue;
                }
            }
            ra[S][e] = load_a(a + a_fix[e] + l0 * a_step);
        }
#pragma unroll
        for (int e = 0; e < B_PER; ++e) {
            uint64_t const jj = b_in[e] ? (uint64_t)b_col[e] : 0;
            uint64_t const ll = b_in[e] ? (uint64_t)b_row[e] : 0;
            uint64_t const j = j0 + jj;
            T const *bp = BCOL ? b + (j + VEC <= m ? j : m - VEC) + (l0 + ll) * m
                               : b + (j < m ? j : m - 1) * ld + l0 + ll;
            if (VIEW == VIEW_TT || b_vec_ok) rb[S][e] = ldg<V>(bp);
            else rb[S][e] = ldg<typename Gt::vec_u>(bp);
        }
    };

    // TN views store each 16-byte A vector (VEC consecutive l of one column n)
    // as VEC scalars into rows l of the At tile; column n of row l sits at
    // n ^ tn_swz(l).  The XOR moves whole groups of NT columns (the lanes'
    // fragment reads stay one vector read) and spreads the 8 vectors of a row
    // segment over distinct banks: the float stores of every compiled TN shape
    // are conflict-free (32- and 64-bank models; 4-way for BK = 32 without it),
    // the double stores 2-way as before, the fragment reads unchanged.
    auto tn_swz = [](int l) { return NT * (((l >> 3) * 3) & 3); };
    // full_tag false: the slice came from load_tiles (shift and mask the TN
    // rows of A and the rows of a row-major B here).
    auto store_tiles = [&](int buf, auto set, auto full_tag) {
        constexpr int S = decltype(set)::value;
        constexpr bool GUARDED = !decltype(full_tag)::value;
#pragma unroll
        for (int e = 0; e < A_PER; ++e) {
            if (tid + 256 * e >= A_VECS) continue;
            if constexpr (!TN) {
                V const v = (PAD && GUARDED && !a_vec)
                                ? shift_mask(ra[S][e], ra_dm[S][e] & 0xff, ra_dm[S][e] >> 8)
                                : ra[S][e];
                *reinterpret_cast<V *>(&sA[buf][a_row[e] * LDA + a_col[e]]) = v;
            } else {
                V const v = GUARDED ? shift_mask(ra[S][e], ra_dm[S][e] & 0xff, ra_dm[S][e] >> 8) : ra[S][e];
#pragma unroll
                for (int x = 0; x < VEC; ++x)
                    sA[buf][(a_row[e] + x) * LDA + (a_col[e] ^ tn_swz(a_row[e] + x))] = v[x];
            }
        }
#pragma unroll
        for (int e = 0; e < B_PER; ++e) {
            if (!b_in[e]) continue;
            if constexpr (BCOL) {
                *reinterpret_cast<V *>(&sB[buf][b_row[e] * LDB + b_col[e]]) = rb[S][e];
            } else {
                V const v = GUARDED ? shift_mask(rb[S][e], rb_dm[S][e] & 0xff, rb_dm[S][e] >> 8) : rb[S][e];
#pragma unroll
                for (int x = 0; x < VEC; ++x) sB[buf][(b_row[e] + x) * LDB + b_col[e]] = v[x];
            }
        }
    };

    Acc acc[JT][NT];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[jt][nt] = vzero<Acc>();

    int const a_off = wn * 16 * NT + NT * lc; // this lane's columns within the At tile
    int const b_off = wm * 16 * JT + JT * lc; // this lane's rows within the B tile

    uint64_t const nslices = (nq + BK - 1) / BK;
    // The products of one slice from LDS buffer buf.
    auto multiply = [&](int buf) {
#pragma unroll
        for (int kk = 0; kk < BK / 4; ++kk) {
            int const kr = kk * 4 + lr;
            T fa[JT], fb[NT];
            lds_read<T, JT>(&sB[buf][kr * LDB + b_off], fa);
            lds_read<T, NT>(&sA[buf][kr * LDA + (TN ? (a_off ^ tn_swz(kr)) : a_off)], fb);
#pragma unroll
            for (int jt = 0; jt < JT; ++jt)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) acc[jt][nt] = Gt::mfma(fa[jt], fb[nt], acc[jt][nt]);
        }
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, PF - 1>;
    // Slices [sb, se) of the contraction; starts and ends with all waves
    // synchronised and both LDS buffers free.
    auto kloop = [&](auto full_tag, uint64_t sb, uint64_t se) {
        constexpr bool FULL = decltype(full_tag)::value;
        auto load = [&](uint64_t s, auto set) {
            if constexpr (FULL) load_full(s * BK, set);
            else load_tiles(s * BK, set);
        };
        uint64_t const n = se - sb;
        if constexpr (PF == 1) {
            load(sb, I0());
            store_tiles(0, I0(), full_tag);
            __syncthreads();
            for (uint64_t s = 0; s < n; ++s) {
                int const buf = (int)(s & 1);
                bool const more = s + 1 < n;
                if (more) load(sb + s + 1, I0());
                multiply(buf);
                if (more) store_tiles(buf ^ 1, I0(), full_tag);
                __syncthreads();
            }
        } else {
            // Slice sb + s sits in LDS buffer s & 1 and arrives in register
            // set s & 1.  A set is reloaded only after its slice was stored, a
            // buffer only after the barrier that ended its slice's products.
            load(sb, I0());
            if (n > 1) load(sb + 1, I1());
            store_tiles(0, I0(), full_tag);
            __syncthreads();
            for (uint64_t s = 0; s < n; s += 2) {
                if (s + 2 < n) load(sb + s + 2, I0());
                multiply(0);
                if (s + 1 < n) store_tiles(1, I1(), full_tag);
                __syncthreads();
                if (s + 1 < n) {
                    if (s + 3 < n) load(sb + s + 3, I1());
                    multiply(1);
                    if (s + 2 < n) store_tiles(0, I0(), full_tag);
                    __syncthreads();
                }
            }
        }
    };
    if constexpr (VIEW == VIEW_TT) {
        // grouped chain stages (contractions of 16..256): one loop per tile,
        // unguarded only when every slice is whole
        if (full && nq % BK == 0) kloop(std::true_type(), 0, nslices);
        else kloop(std::false_type(), 0, nslices);
    } else {
        uint64_t const nwhole = full ? nq / BK : 0; // slices with unguarded loads
        if (nwhole > 0) kloop(std::true_type(), 0, nwhole);
        if (nwhole < nslices) kloop(std::false_type(), nwhole, nslices);
    }

    // ---- epilogue ---------------------------------------------------------
    uint64_t const ncol0 = n0 + a_off; // first of this lane's NT adjacent columns
    if constexpr (!TSTORE) {
        // one row j, NT adjacent columns; each VEC-group of them lies in one
        // slab (wp % VEC == 0), of which gnv are real columns, but the NT
        // columns may span two slabs.
        constexpr int NG = NT / VEC;
        static_assert(NT % VEC == 0, "NT");
        T *cgroup[NG];
        int gnv[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            uint64_t const n = ncol0 + g * VEC;
            bool const in = n < ncols;
            if constexpr (VIEW == VIEW_NN && !PAD) {
                // w % VEC == 0: every group is whole
                uint64_t const k = in ? n / w : 0;
                gnv[g] = in ? VEC : 0;
                cgroup[g] = c + k * m * w + (in ? n - k * w : 0);
            } else {
                uint64_t const k = in ? n / wp : 0;
                uint64_t const i = in ? n - k * wp : 0;
                gnv[g] = (in && i < w) ? (int)(w - i < (uint64_t)VEC ? w - i : VEC) : 0;
                cgroup[g] = c + k * m * w + i;
            }
        }
#pragma unroll
        for (int jt = 0; jt < JT; ++jt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                uint64_t const j = j0 + wm * 16 * JT + JT * Gt::acc_row(lr, r) + jt;
                if (j >= m) continue;
#pragma unroll
                for (int g = 0; g < NG; ++g) {
                    if (gnv[g] == 0) continue;
                    if (c_vec || (VIEW == VIEW_NN && !PAD)) {
                        V o;
#pragma unroll
                        for (int x = 0; x < VEC; ++x) o[x] = acc[jt][g * VEC + x][r];
                        if (nt_c) stg_nt<V>(cgroup[g] + j * w, o);
                        else stg<V>(cgroup[g] + j * w, o);
                    } else if (VIEW == VIEW_NN && gnv[g] == VEC) {
                        // whole group at an element-aligned address
                        V o;
#pragma unroll
                        for (int x = 0; x < VEC; ++x) o[x] = acc[jt][g * VEC + x][r];
                        stg<typename Gt::vec_u>(cgroup[g] + j * w, o);
                    } else {
                        // the group runs past the slab row's (or C's) end
#pragma unroll
                        for (int x = 0; x < VEC; ++x)
                            if (x < gnv[g]) stg<T>(cgroup[g] + j * w + x, acc[jt][g * VEC + x][r]);
                    }
                }
            }
    } else {
        // one column (row k of C), JT adjacent rows j
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            uint64_t const k = ncol0 + nt;
            if (k >= ncols) continue;
            T *crow = c + k * m;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                uint64_t const jr = j0 + wm * 16 * JT + JT * Gt::acc_row(lr, r);
                if constexpr (JT % VEC == 0) {
                    if (c_vec && jr + JT <= m) {
#pragma unroll
                        for (int jt = 0; jt < JT; jt += VEC) {
                            V o;
#pragma unroll
                            for (int x = 0; x < VEC; ++x) o[x] = acc[jt + x][nt][r];
                            if (nt_c) stg_nt<V>(crow + jr + jt, o);
                            else stg<V>(crow + jr + jt, o);
                        }
                        continue;
                    }
                }
#pragma unroll
                for (int jt = 0; jt < JT; ++jt)
                    if (jr + jt < m) stg<T>(crow + jr + jt, acc[jt][nt][r]);
            }
        }
    }
}

// One product: a grid of BM x BN blocks (row blocks fastest, XCD-grouped)
// times gridDim.y contraction splits.  Split s covers [s*kc, s*kc + kc) of
// the contraction and writes its partial product, in C's layout, at
// c + s * c_split (the partials are summed by k_sum_splits); with one split
// the block writes C itself.
template <class T, int JT, int NT, int WAVES_M, int BK, bool TN, bool BCOL, bool PAD = false>
__global__ __launch_bounds__(256, 2) void k_gemm(const T *__restrict__ a, const T *__restrict__ b,
                                              T *__restrict__ c, uint64_t ncols, uint64_t nq, uint64_t w,
                                              uint64_t wp, uint64_t m, unsigned nj, unsigned total, uint64_t kc,
                                              uint64_t c_split, unsigned flags)
{
    using S = TileShape<T, JT, NT, WAVES_M, BK>;
    __shared__ __attribute__((aligned(16))) T sA[2][BK * S::LDA];
    __shared__ __attribute__((aligned(16))) T sB[2][BK * S::LDB];
    unsigned jb, nb;
    block_coords(nj, total, jb, nb);
    uint64_t const l0 = (uint64_t)blockIdx.y * kc;
    uint64_t const len = nq - l0 < kc ? nq - l0 : kc;
    // elements of A from the split's first row to the end of A (PAD reads)
    uint64_t const a_len = TN ? 0 : (ncols / wp) * nq * w - l0 * w;
    gemm_tile<T, JT, NT, WAVES_M, BK, TN ? VIEW_TN : VIEW_NN, BCOL, PAD>(
        a + l0 * (TN ? 1 : w), b + l0 * (BCOL ? m : 1), c + (uint64_t)blockIdx.y * c_split, ncols, len, nq, w, wp,
        m, (uint64_t)jb * S::BM, (uint64_t)nb * S::BN, flags, a_len, sA, sB);
}

// C[i] = sum_{s < S} P[s * n + i] in the order s = 0, 1, ... (deterministic).
// VEC-wide when n % VEC == 0 and C is 16-byte aligned (the partials are).
// One thread per element (vector): for many elements and few splits.
template <class T>
__global__ __launch_bounds__(256) void k_sum_splits(const T *__restrict__ p, uint64_t n, unsigned nsplit,
                                                    T *__restrict__ c, int vec)
{
    constexpr int VEC = G<T>::VEC;
    using V = typename G<T>::vec;
    uint64_t const stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t const t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (vec) {
        uint64_t const nv = n / VEC;
        for (uint64_t v = t0; v < nv; v += stride) {
            V acc = ldg<V>(p + v * VEC);
            for (unsigned s = 1; s < nsplit; ++s) acc += ldg<V>(p + (uint64_t)s * n + v * VEC);
            stg<V>(c + v * VEC, acc);
        }
    } else {
        for (uint64_t i = t0; i < n; i += stride) {
            T acc = p[i];
            for (unsigned s = 1; s < nsplit; ++s) acc += p[(uint64_t)s * n + i];
            c[i] = acc;
        }
    }
}

// The same sum with a wave per element (vector): lane l adds the splits
// l, l + 64, ... in order, then the 64 lane sums are added by a fixed
// butterfly (the result does not depend on scheduling).
template <class T>
__global__ __launch_bounds__(256) void k_sum_splits_wave(const T *__restrict__ p, uint64_t n, unsigned nsplit,
                                                         T *__restrict__ c, int vec)
{
    constexpr int VEC = G<T>::VEC;
    using V = typename G<T>::vec;
    int const lane = threadIdx.x & 63;
    uint64_t const v = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (vec) {
        if (v >= n / VEC) return;
        V acc = vzero<V>();
        for (unsigned s = lane; s < nsplit; s += 64) acc += ldg<V>(p + (uint64_t)s * n + v * VEC);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
#pragma unroll
            for (int e = 0; e < VEC; ++e) acc[e] += __shfl_xor(acc[e], o, 64);
        if (lane == 0) stg<V>(c + v * VEC, acc);
    } else {
        if (v >= n) return;
        T acc = 0;
        for (unsigned s = lane; s < nsplit; s += 64) acc += p[(uint64_t)s * n + v];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (lane == 0) c[v] = acc;
    }
}

// Many products in one persistent launch: workgroup g processes tiles
// g, g + gridDim.x, ... of the concatenated tile lists of all problems.
// Every problem is a TT product with a row-major B (a stage of a multi-TTM
// chain); tile_prob[t] names the problem of tile t.
template <class T, int JT, int NT, int WAVES_M, int BK>
__global__ __launch_bounds__(256, 2) void k_gemm_grouped(const GroupProblem *__restrict__ probs,
                                                      const unsigned *__restrict__ tile_prob, unsigned total)
{
    using S = TileShape<T, JT, NT, WAVES_M, BK>;
    __shared__ __attribute__((aligned(16))) T sA[2][BK * S::LDA];
    __shared__ __attribute__((aligned(16))) T sB[2][BK * S::LDB];
    // Tiles are short (a stage contracts 16..256 indices), so the descriptor
    // of the next tile is read while this one runs and the problem id of the
    // tile after it one round earlier: no tile starts behind two dependent
    // descriptor loads.
    unsigned t = blockIdx.x;
    unsigned const step = gridDim.x;
    if (t >= total) return;
    GroupProblem cur = probs[tile_prob[t]];
    unsigned pid_next = t + step < total ? tile_prob[t + step] : 0u;
    for (;;) {
        unsigned const tn = t + step;
        bool const has_next = tn < total;
        GroupProblem nxt = cur;
        unsigned pid_after = 0;
        if (has_next) {
            nxt = probs[pid_next];
            if (tn + step < total) pid_after = tile_prob[tn + step];
        }
        unsigned const local = t - cur.tile_begin;
        unsigned const jb = local % cur.nj, nb = local / cur.nj;
        gemm_tile<T, JT, NT, WAVES_M, BK, VIEW_TT, false>(
            static_cast<const T *>(cur.a), static_cast<const T *>(cur.b), static_cast<T *>(cur.c), cur.ncols,
            cur.nq, cur.nq, cur.ncols, cur.ncols, cur.m, (uint64_t)jb * S::BM, (uint64_t)nb * S::BN, cur.flags, 0,
            sA, sB);
        if (!has_next) break;
        t = tn;
        cur = nxt;
        pid_next = pid_after;
    }
}

// ---------------------------------------------------------------------------
// host side: tile configurations
// ---------------------------------------------------------------------------
struct Config
{
    int jt, nt, waves_m, bk;
};

// Small m streams A once with wide column blocks; large m uses square blocks.
Config pick_config(const Fold &f, bool is_double)
{
    // float32 rows read in the TN view: slices of 32 (128 bytes, one whole
    // L2 line of each row per slice; with 16, the two halves of a line are
    // read one slice apart and a third of the lines is fetched twice)
    if (f.m <= 16 && f.w == 1 && !is_double) return {1, 4, 1, 32};
    // double rows in the TN view: 128-column blocks (half the LDS, twice the
    // workgroups per CU): 0.64-0.68 -> 0.67-0.69 of 8 TB/s at m = 16
    if (f.m <= 16 && f.w == 1) return {1, 2, 1, 16};
    // float64 blocks of 32 and 64 rows (both views): 128 columns.  With 256
    // columns the 64-row block needs 83 KB of LDS and 388 registers (one
    // wave per SIMD), the 32-row one 75 KB (two); with 128 columns 50 / 42 KB
    // and 167 / 146 registers, three blocks per CU
    if (is_double && f.m > 16 && f.m <= 32) return {2, 2, 1, 16};
    if (is_double && f.m > 16 && f.m <= 64) return {4, 2, 1, 16};
    if (f.m <= 16) return {1, 4, 1, 16};
    if (f.m <= 32) return {2, 4, 1, 16};
    if (f.m <= 64) return {4, 4, 1, 16};
    return {4, 4, 2, 16};
}

// Geometry of a product in the gemm family: tile shape, padded slab width
// (NN view), column count in that padded space and the block grid.
struct Geom
{
    Config cfg;
    bool tn = false;
    bool pad = false; // NN view that needs the PAD tile (set by the launcher from w and the pointers)
    uint64_t wp = 1, ncols = 0;
    unsigned nj = 0, nn = 0;
    int bm() const { return cfg.waves_m * 16 * cfg.jt; }
    int bn() const { return (4 / cfg.waves_m) * 16 * cfg.nt; }
};

Geom geom(const Fold &f, bool is_double)
{
    uint64_t const vec = is_double ? 2 : 4;
    Geom g;
    g.cfg = pick_config(f, is_double);
    g.tn = f.w == 1;
    g.wp = g.tn ? 1 : (f.w + vec - 1) / vec * vec;
    g.ncols = g.tn ? f.outer : f.outer * g.wp;
    g.nj = (unsigned)((f.m + g.bm() - 1) / g.bm());
    g.nn = (unsigned)((g.ncols + g.bn() - 1) / g.bn());
    return g;
}

// Contraction splits: when the blocks of a product cannot fill the chip
// (fewer than two per compute unit, the occupancy of the tiles), the
// contraction is cut into `splits` ranges of kc indices (a multiple of 64,
// so that every range starts 16-byte aligned and on a whole 128-byte line of
// a float32 row) whose partial products are summed by a second pass.  Ranges
// keep at least 512 indices, and the partials stay within 256 MiB (within the
// fixed scratch region when one is bound).
constexpr uint64_t kSplitMinChunk = 512;
constexpr uint64_t kSplitMaxBytes = 256ull << 20;

Split split_for(const Fold &f, bool is_double, const Geom &g, unsigned want)
{
    Split sp;
    sp.kc = f.nq;
    uint64_t const tiles = (uint64_t)g.nj * g.nn;
    uint64_t s = want;
    if (s == 0) {
        uint64_t const target = (uint64_t)device_cus() * 2;
        s = tiles >= target ? 1 : (target + tiles - 1) / tiles;
    }
    uint64_t const by_len = f.nq / kSplitMinChunk;
    uint64_t const room = std::min<uint64_t>(kSplitMaxBytes, scratch_available());
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

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

template <class T, int JT, int NT, int WM, int BK, bool TN, bool BCOL>
void launch_cfg(const Fold &f, const Geom &g, const Split &sp, unsigned flags, const T *a, const T *b, T *c,
                hipStream_t s)
{
    unsigned const total = g.nj * g.nn;
    uint64_t const c_split = sp.splits > 1 ? f.numel_c() : 0;
    if constexpr (!TN) {
        if (g.pad) {
            hipLaunchKernelGGL((k_gemm<T, JT, NT, WM, BK, TN, BCOL, true>), dim3(total, sp.splits), dim3(256), 0, s,
                               a, b, c, g.ncols, f.nq, f.w, g.wp, f.m, g.nj, total, sp.kc, c_split, flags);
            return;
        }
    }
    hipLaunchKernelGGL((k_gemm<T, JT, NT, WM, BK, TN, BCOL>), dim3(total, sp.splits), dim3(256), 0, s, a, b, c,
                       g.ncols, f.nq, f.w, g.wp, f.m, g.nj, total, sp.kc, c_split, flags);
}

template <class T, bool TN, bool BCOL>
void launch_layout(const Fold &f, const Geom &g, const Split &sp, unsigned flags, const T *a, const T *b, T *c,
                   hipStream_t s)
{
    Config const cfg = g.cfg;
    if constexpr (sizeof(T) == 4 && TN) {
        if (cfg.jt == 1 && cfg.bk == 32) {
            launch_cfg<T, 1, 4, 1, 32, TN, BCOL>(f, g, sp, flags, a, b, c, s);
            return;
        }
    }
    if constexpr (sizeof(T) == 8) {
        if (cfg.nt == 2) {
            if constexpr (TN) {
                if (cfg.jt == 1) {
                    launch_cfg<T, 1, 2, 1, 16, TN, BCOL>(f, g, sp, flags, a, b, c, s);
                    return;
                }
            }
            if (cfg.jt == 2) launch_cfg<T, 2, 2, 1, 16, TN, BCOL>(f, g, sp, flags, a, b, c, s);
            else launch_cfg<T, 4, 2, 1, 16, TN, BCOL>(f, g, sp, flags, a, b, c, s);
            return;
        }
    }
    if (cfg.jt == 1) launch_cfg<T, 1, 4, 1, 16, TN, BCOL>(f, g, sp, flags, a, b, c, s);
    else if (cfg.jt == 2) launch_cfg<T, 2, 4, 1, 16, TN, BCOL>(f, g, sp, flags, a, b, c, s);
    else if (cfg.waves_m == 1) launch_cfg<T, 4, 4, 1, 16, TN, BCOL>(f, g, sp, flags, a, b, c, s);
    else launch_cfg<T, 4, 4, 2, 16, TN, BCOL>(f, g, sp, flags, a, b, c, s);
}

} // namespace

std::string gemm_symbol(const Fold &f, bool is_double, bool aligned)
{
    Config const cfg = pick_config(f, is_double);
    uint64_t const vec = is_double ? 2 : 4;
    bool const pad = f.w != 1 && (!aligned || f.w % vec != 0);
    char buf[128];
    std::snprintf(buf, sizeof buf, "k_gemm<%s, %d, %d, %d, %d, %s, %s, %s>", is_double ? "double" : "float", cfg.jt,
                  cfg.nt, cfg.waves_m, cfg.bk, f.w == 1 ? "true" : "false", f.b_colmajor ? "true" : "false",
                  pad ? "true" : "false");
    return buf;
}

bool gemm_supports(const Fold &f, bool is_double)
{
    // the element-aligned reads take whole vectors from inside A
    if (f.numel_a() < (is_double ? 2u : 4u)) return false;
    Geom const g = geom(f, is_double);
    // one-dimensional grid of blocks (the splits go along y)
    return (uint64_t)g.nj * g.nn < (1ull << 31) && g.nj > 0 && g.nn > 0;
}

Split gemm_split(const Fold &f, bool is_double, unsigned want)
{
    return split_for(f, is_double, geom(f, is_double), want);
}

bool gemm_underfilled(const Fold &f, bool is_double)
{
    Geom const g = geom(f, is_double);
    return (uint64_t)g.nj * g.nn < (uint64_t)device_cus() * 2;
}

template <class T>
ttm_status_t launch_gemm(const Fold &f, const T *a, const T *b, T *c, hipStream_t s, unsigned want_splits)
{
    constexpr uint64_t VEC = 16 / sizeof(T);
    Geom g = geom(f, sizeof(T) == 8);
    Split const sp = split_for(f, sizeof(T) == 8, g, want_splits);
    T *out = c;
    void *ws = nullptr;
    if (sp.splits > 1) {
        size_t const bytes = (size_t)sp.splits * f.numel_c() * sizeof(T);
        if (ttm_status_t const st = ws_acquire(bytes, s, &ws, "workspace for the split-contraction partials");
            st != TTM_SUCCESS)
            return st;
        out = static_cast<T *>(ws);
    }
    bool const c_al = sp.splits > 1 ? f.numel_c() % VEC == 0 : aligned16(c);
    unsigned flags = 0;
    if (aligned16(a) && (g.tn ? f.nq % VEC == 0 : f.w % VEC == 0)) flags |= TILE_A_VEC;
    if (c_al && (g.tn ? f.m % VEC == 0 : f.w % VEC == 0)) flags |= TILE_C_VEC;
    if (aligned16(b)) flags |= TILE_B_VEC;
    if (g.nj == 1) flags |= TILE_STREAM | (sp.splits > 1 ? TILE_KEEP_C : 0u);
    g.pad = !g.tn && !((flags & TILE_A_VEC) && (flags & TILE_C_VEC));
    if (g.tn) {
        if (f.b_colmajor) launch_layout<T, true, true>(f, g, sp, flags, a, b, out, s);
        else launch_layout<T, true, false>(f, g, sp, flags, a, b, out, s);
    } else {
        if (f.b_colmajor) launch_layout<T, false, true>(f, g, sp, flags, a, b, out, s);
        else launch_layout<T, false, false>(f, g, sp, flags, a, b, out, s);
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

template <class T>
void launch_sum_splits(const T *partials, uint64_t n, unsigned splits, T *c, hipStream_t s)
{
    constexpr uint64_t VEC = 16 / sizeof(T);
    int const vec = (n % VEC == 0 && aligned16(c) && aligned16(partials)) ? 1 : 0;
    uint64_t const work = vec ? n / VEC : n;
    uint64_t const cus = (uint64_t)device_cus();
    if (splits >= 16 && work < cus * 1024) {
        // few elements, many splits: a wave per element (vector), lanes over
        // the splits (the per-thread loops would be long chains of loads)
        uint64_t const blocks = (work + 3) / 4;
        hipLaunchKernelGGL(k_sum_splits_wave<T>, dim3((unsigned)blocks), dim3(256), 0, s, partials, n, splits, c,
                           vec);
        return;
    }
    uint64_t blocks = (work + 255) / 256;
    uint64_t const cap = cus * 8;
    if (blocks > cap) blocks = cap;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_sum_splits<T>, dim3((unsigned)blocks), dim3(256), 0, s, partials, n, splits, c, vec);
}

template void launch_sum_splits<float>(const float *, uint64_t, unsigned, float *, hipStream_t);
template void launch_sum_splits<double>(const double *, uint64_t, unsigned, double *, hipStream_t);

template ttm_status_t launch_gemm<float>(const Fold &, const float *, const float *, float *, hipStream_t, unsigned);
template ttm_status_t launch_gemm<double>(const Fold &, const double *, const double *, double *, hipStream_t,
                                          unsigned);

// ---- grouped launches -------------------------------------------------------
// A chain stage over a mode without a factor: the layout rotation that the
// product with the identity would do, as a copy, C[l * ncols + n] =
// A[n * nq + l] (an ncols x nq matrix transposed).  Persistent like the
// grouped products: workgroup g takes 64 x 64 tiles g, g + gridDim.x, ...
// of the concatenated tile lists (nj = ceil(nq / 64) l-blocks per column
// block, l-blocks fastest).  A tile goes through LDS so that the reads (along
// l) and the writes (along n) are both contiguous across the lanes.
template <class T>
__global__ __launch_bounds__(256) void k_transpose_grouped(const GroupProblem *__restrict__ probs,
                                                        const unsigned *__restrict__ tile_prob, unsigned total)
{
    __shared__ T tile[64][65];
    int const tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (unsigned t = blockIdx.x; t < total; t += gridDim.x) {
        GroupProblem const P = probs[tile_prob[t]];
        unsigned const local = t - P.tile_begin;
        uint64_t const l0 = (uint64_t)(local % P.nj) * 64, n0 = (uint64_t)(local / P.nj) * 64;
        const T *a = static_cast<const T *>(P.a);
        T *c = static_cast<T *>(P.c);
        __syncthreads(); // every read of the previous tile is done
#pragma unroll 4
        for (int r = ty; r < 64; r += 4) {
            uint64_t const n = n0 + r, l = l0 + tx;
            if (n < P.ncols && l < P.nq) tile[r][tx] = a[n * P.nq + l];
        }
        __syncthreads();
#pragma unroll 4
        for (int r = ty; r < 64; r += 4) {
            uint64_t const l = l0 + r, n = n0 + tx;
            if (l < P.nq && n < P.ncols) c[l * P.ncols + n] = tile[tx][r];
        }
    }
}

template <class T>
ttm_status_t launch_transpose_grouped(const GroupProblem *d_probs, const unsigned *d_tile_prob, unsigned total,
                                      unsigned grid, hipStream_t s)
{
    if (total == 0) return TTM_SUCCESS;
    hipLaunchKernelGGL(k_transpose_grouped<T>, dim3(grid), dim3(256), 0, s, d_probs, d_tile_prob, total);
    hipError_t const e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(TTM_ERR_HIP, hipGetErrorString(e));
        return TTM_ERR_HIP;
    }
    return TTM_SUCCESS;
}
template ttm_status_t launch_transpose_grouped<float>(const GroupProblem *, const unsigned *, unsigned, unsigned,
                                                      hipStream_t);
template ttm_status_t launch_transpose_grouped<double>(const GroupProblem *, const unsigned *, unsigned, unsigned,
                                                       hipStream_t);

namespace
{
// (jt, nt, waves_m, bk): rows 16/32/64 with 256 columns, 128 x 128, and 64 or 32
// rows with 128 columns for short problems.
// Indices 6..10 (slices of 32) are float32 only: rows read in the TT view
// then cover whole 128-byte lines per slice, and short rows (16..256
// indices) take half as many slices, each with twice the matrix work.
constexpr GroupConfig k_group_configs[] = {
    {1, 4, 1, 16}, {2, 4, 1, 16}, {4, 4, 1, 16}, {4, 4, 2, 16}, {2, 4, 2, 16}, {1, 4, 2, 16},
    {1, 4, 1, 32}, {2, 4, 1, 32}, {4, 4, 1, 32}, {1, 4, 2, 32}, {2, 4, 2, 32},
    {1, 4, 4, 16}, // 64 x 64: four times the tiles of 64 x 256 for the small later stages
};
constexpr int k_num_group_configs = sizeof(k_group_configs) / sizeof(k_group_configs[0]);

template <class T, int JT, int NT, int WM, int BK>
const void *grouped_fn()
{
    return reinterpret_cast<const void *>(&k_gemm_grouped<T, JT, NT, WM, BK>);
}

template <class T>
const void *grouped_kernel(int index)
{
    switch (index) {
    case 0: return grouped_fn<T, 1, 4, 1, 16>();
    case 1: return grouped_fn<T, 2, 4, 1, 16>();
    case 2: return grouped_fn<T, 4, 4, 1, 16>();
    case 3: return grouped_fn<T, 4, 4, 2, 16>();
    case 4: return grouped_fn<T, 2, 4, 2, 16>();
    case 5: return grouped_fn<T, 1, 4, 2, 16>();
    case 11: return grouped_fn<T, 1, 4, 4, 16>();
    default: break;
    }
    if constexpr (sizeof(T) == 4) {
        switch (index) {
        case 6: return grouped_fn<T, 1, 4, 1, 32>();
        case 7: return grouped_fn<T, 2, 4, 1, 32>();
        case 8: return grouped_fn<T, 4, 4, 1, 32>();
        case 9: return grouped_fn<T, 1, 4, 2, 32>();
        case 10: return grouped_fn<T, 2, 4, 2, 32>();
        default: break;
        }
    }
    return nullptr;
}

int config_index(const GroupConfig &cfg)
{
    for (int i = 0; i < k_num_group_configs; ++i) {
        GroupConfig const &g = k_group_configs[i];
        if (g.jt == cfg.jt && g.nt == cfg.nt && g.waves_m == cfg.waves_m && g.bk == cfg.bk) return i;
    }
    return -1;
}
} // namespace

int group_config_count() { return k_num_group_configs; }

GroupConfig group_config(int index) { return k_group_configs[index]; }

std::string group_symbol(const GroupConfig &cfg, bool is_double)
{
    char buf[96];
    std::snprintf(buf, sizeof buf, "k_gemm_grouped<%s, %d, %d, %d, %d>", is_double ? "double" : "float", cfg.jt,
                  cfg.nt, cfg.waves_m, cfg.bk);
    return buf;
}

int group_occupancy(const GroupConfig &cfg, bool is_double)
{
    int const idx = config_index(cfg);
    if (idx < 0) return 0;
    const void *fn = is_double ? grouped_kernel<double>(idx) : grouped_kernel<float>(idx);
    if (!fn) return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, 256, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return blocks;
}

template <class T>
ttm_status_t launch_gemm_grouped(const GroupConfig &cfg, const GroupProblem *d_probs, const unsigned *d_tile_prob,
                                 unsigned total, unsigned grid, hipStream_t s)
{
    int const idx = config_index(cfg);
    if (idx < 0 || grid == 0) {
        set_error(TTM_ERR_ARGUMENT, "unknown grouped tile configuration");
        return TTM_ERR_ARGUMENT;
    }
    switch (idx) {
    case 0: hipLaunchKernelGGL((k_gemm_grouped<T, 1, 4, 1, 16>), dim3(grid), dim3(256), 0, s, d_probs, d_tile_prob, total); break;
    case 1: hipLaunchKernelGGL((k_gemm_grouped<T, 2, 4, 1, 16>), dim3(grid), dim3(256), 0, s, d_probs, d_tile_prob, total); break;
    case 2: hipLaunchKernelGGL((k_gemm_grouped<T, 4, 4, 1, 16>), dim3(grid), dim3(256), 0, s, d_probs, d_tile_prob, total); break;
    case 3: hipLaunchKernelGGL((k_gemm_grouped<T, 4, 4, 2, 16>), dim3(grid), dim3(256), 0, s, d_probs, d_tile_prob, total); break;
    case 4: hipLaunchKernelGGL((k_gemm_grouped<T, 2, 4, 2, 16>), dim3(grid), dim3(256), 0, s, d_probs, d_tile_prob, total); break;
    case 5: hipLaunchKernelGGL((k_gemm_grouped<T, 1, 4, 2, 16>), dim3(grid), dim3(256), 0, s, d_probs, d_tile_prob, total); break;
    case 11: hipLaunchKernelGGL((k_gemm_grouped<T, 1, 4, 4, 16>), dim3(grid), dim3(256), 0, s, d_probs, d_tile_prob, total); break;
    default:
        if constexpr (sizeof(T) == 4) {
            if (idx == 6) hipLaunchKernelGGL((k_gemm_grouped<T, 1, 4, 1, 32>), dim3(grid), dim3(256), 0, s, d_probs, d_tile_prob, total);
            else if (idx == 7) hipLaunchKernelGGL((k_gemm_grouped<T, 2, 4, 1, 32>), dim3(grid), dim3(256), 0, s, d_probs, d_tile_prob, total);
            else if (idx == 8) hipLaunchKernelGGL((k_gemm_grouped<T, 4, 4, 1, 32>), dim3(grid), dim3(256), 0, s, d_probs, d_tile_prob, total);
            else if (idx == 9) hipLaunchKernelGGL((k_gemm_grouped<T, 1, 4, 2, 32>), dim3(grid), dim3(256), 0, s, d_probs, d_tile_prob, total);
            else if (idx == 10) hipLaunchKernelGGL((k_gemm_grouped<T, 2, 4, 2, 32>), dim3(grid), dim3(256), 0, s, d_probs, d_tile_prob, total);
        } else {
            set_error(TTM_ERR_ARGUMENT, "grouped tile configuration not compiled for float64");
            return TTM_ERR_ARGUMENT;
        }
        break;
    }
    hipError_t const e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(TTM_ERR_HIP, hipGetErrorString(e));
        return TTM_ERR_HIP;
    }
    return TTM_SUCCESS;
}

template ttm_status_t launch_gemm_grouped<float>(const GroupConfig &, const GroupProblem *, const unsigned *,
                                                 unsigned, unsigned, hipStream_t);
template ttm_status_t launch_gemm_grouped<double>(const GroupConfig &, const GroupProblem *, const unsigned *,
                                                  unsigned, unsigned, hipStream_t);

} // namespace ttm_amd
