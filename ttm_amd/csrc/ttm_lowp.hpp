// Reduced-precision arithmetic (BF16, FP16, BF16X3, FP16X2, FP8 e4m3fn) for float32 products.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "ttm_amd.h"
#include "ttm_gemm.hpp"
#include "ttm_plan.hpp"

namespace ttm_amd
{

// True when the MFMA kernels can run the fold: 16-byte aligned buffers and
// either w == 1 with nq % 4 == 0 (TN view) or w % 4 == 0 (NN view).
bool lowp_supports(const Fold &f, bool aligned);

// Template instance launched for a fold and precision (as profilers print it).
std::string lowp_symbol(const Fold &f, ttm_precision_t prec, bool generic);

// `generic` selects the one-thread-per-element kernel (any shape/alignment).
ttm_status_t launch_lowp(const Fold &f, const float *a, const float *b, float *c, ttm_precision_t prec,
                         hipStream_t s, bool generic);

// Grouped BF16X3 or FP16X2 launches of TT products (stages of multi-TTM
// chains) in bm x 128 tiles, bm = 128 or 64 (products of at most 64 rows);
// tile lists as for launch_gemm_grouped.  FP16X2 scales both operands per
// tile and slice (no pass before the launch).
bool lowp_group_precision(ttm_precision_t prec);
int lowp_group_occupancy(ttm_precision_t prec, int bm);
std::string lowp_group_symbol(ttm_precision_t prec, int bm);
ttm_status_t launch_lowp_grouped(ttm_precision_t prec, int bm, const GroupProblem *d_probs,
                                 const unsigned *d_tile_prob, unsigned total, unsigned grid, hipStream_t s);

} // namespace ttm_amd
