// LDS-staged MFMA kernels (plain FP32 / FP64) for the folded product.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "ttm_amd.h"
#include "ttm_plan.hpp"

namespace ttm_amd
{

// True when the gemm family can run the fold (any extents and element-aligned
// buffers; false only beyond its grid limits).
bool gemm_supports(const Fold &f, bool is_double);

// Template instance the gemm launcher runs for a fold (as profilers print it);
// `aligned`: A and C 16-byte aligned.
std::string gemm_symbol(const Fold &f, bool is_double, bool aligned = true);

// Contraction split of a product: `splits` ranges of kc indices.
struct Split
{
    unsigned splits = 1;
    uint64_t kc = 0;
};
// The split the launcher uses: `want` splits (0: the planner's choice, which
// is 1 unless the blocks cannot fill the chip), clamped to ranges of at least
// 512 indices and 256 MiB of partials.
Split gemm_split(const Fold &f, bool is_double, unsigned want = 0);
// Fewer blocks than two per compute unit.
bool gemm_underfilled(const Fold &f, bool is_double);

// C[i] = sum over the splits of partials[s * n + i], in split order
// (deterministic), for i < n; partials is 16-byte aligned.
template <class T>
void launch_sum_splits(const T *partials, uint64_t n, unsigned splits, T *c, hipStream_t s);

// want_splits as for gemm_split.  Partials come from the call's workspace.
template <class T>
ttm_status_t launch_gemm(const Fold &f, const T *a, const T *b, T *c, hipStream_t s, unsigned want_splits = 0);

// ---- grouped (persistent, many products per launch) ------------------------

enum TileFlags : unsigned
{
    // A may be read with aligned 16-byte loads: TN/TT views, rows of A start
    // 16-byte aligned (row extent % vec == 0); NN view, w % vec == 0 and A aligned
    TILE_A_VEC = 1u,
    // C may be written with aligned 16-byte stores (NN/TT: w % vec == 0, TN:
    // m % vec == 0, and C aligned)
    TILE_C_VEC = 2u,
    TILE_STREAM = 4u, // A is read exactly once (one row block): non-temporal loads of A and stores of C
    TILE_B_VEC = 8u,  // B is 16-byte aligned (its row or column stride is checked in the tile)
    TILE_KEEP_C = 16u // C stores stay temporal under TILE_STREAM (partials that are read back)
};

// One product of a grouped launch, in the TT view:
//   C[j * ncols + n] = sum_{l < nq} B[j * nq + l] * A[n * nq + l]
// (rows of A contiguous along the contracted mode, B row-major m x nq, the
// new mode of C slowest).  tile_begin is the problem's first tile in the
// launch's tile list, nj its number of row blocks.
struct GroupProblem
{
    const void *a;
    const void *b;
    void *c;
    uint64_t ncols;
    uint64_t nq;
    uint64_t m;
    unsigned tile_begin;
    unsigned nj;
    unsigned flags;
    unsigned reserved;
};

// Tile shape of a grouped launch: BM = waves_m * 16 * jt rows,
// BN = (4 / waves_m) * 16 * nt columns, contraction slices of bk.
struct GroupConfig
{
    int jt, nt, waves_m, bk;
    int bm() const { return waves_m * 16 * jt; }
    int bn() const { return (4 / waves_m) * 16 * nt; }
};

// The configurations the grouped launcher is compiled for.
// Chain stages over modes without a factor: the TT view's layout rotation
// as a copy, C[l * ncols + n] = A[n * nq + l], in 64 x 64 tiles (nj =
// ceil(nq / 64) per column block); `grid` persistent workgroups.
template <class T>
ttm_status_t launch_transpose_grouped(const GroupProblem *d_probs, const unsigned *d_tile_prob, unsigned total,
                                      unsigned grid, hipStream_t s);
int group_config_count();
GroupConfig group_config(int index);
std::string group_symbol(const GroupConfig &cfg, bool is_double);

// Resident workgroups per compute unit for a configuration (0 on error).
int group_occupancy(const GroupConfig &cfg, bool is_double);

// Launch `grid` persistent workgroups over `total` tiles.
template <class T>
ttm_status_t launch_gemm_grouped(const GroupConfig &cfg, const GroupProblem *d_probs, const unsigned *d_tile_prob,
                                 unsigned total, unsigned grid, hipStream_t s);

} // namespace ttm_amd
