// LDS-staged MFMA kernels (plain FP32 / FP64) for the folded product.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "ttm_amd.h"
#include "ttm_plan.hpp"

namespace ttm_amd
{

// True when the gemm family can run the fold: 16-byte aligned buffers and
// either w == 1 with nq % vec == 0 (TN view) or w % vec == 0 (NN view).
bool gemm_supports(const Fold &f, bool is_double, bool aligned);

// Template instance the gemm launcher runs for a fold (as profilers print it).
std::string gemm_symbol(const Fold &f, bool is_double);

template <class T>
ttm_status_t launch_gemm(const Fold &f, const T *a, const T *b, T *c, hipStream_t s);

// ---- grouped (persistent, many products per launch) ------------------------

enum TileFlags : unsigned
{
    TILE_A_VEC = 1u, // rows of A may be read with 16-byte loads (nq % vec == 0, aligned)
    TILE_C_VEC = 2u, // rows of C may be written with 16-byte stores (ncols % vec == 0, aligned)
    TILE_STREAM = 4u // A is read exactly once (one row block): non-temporal loads of A and stores of C
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
