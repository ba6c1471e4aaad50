// Kernel families and the launcher interface shared by the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "ttm_amd.h"
#include "ttm_plan.hpp"

namespace ttm_amd
{

enum class Kernel : int
{
    none = 0,
    generic,    // one thread per C element, any shape / alignment
    nn_stream,  // MFMA 16x16x4, w % vec == 0: streams A slabs along the contiguous block
    tn_stream,  // MFMA 16x16x4, w == 1: rows of A contiguous along the contraction mode
    gemm,       // LDS-staged MFMA blocks (ttm_gemm.hip), NN (w % vec == 0) or TN (w == 1) view
    lowp,       // reduced-precision MFMA blocks (ttm_lowp.hip): bf16, fp16, bf16x3, fp8
    lowp_generic, // reduced precision, one thread per element (any shape / alignment)
    skinny,     // vector-ALU stream of B (ttm_skinny.hip): w == 1 with at most eight columns
    count
};

// skinny family (ttm_skinny.hip): TN view (w == 1), outer <= 8, nq >= 512.
bool skinny_supports(const Fold &f);
std::string skinny_symbol(const Fold &f, bool is_double);
// Contraction splits the skinny launcher uses (want: 0 = its own choice).
unsigned skinny_splits(const Fold &f, bool is_double, unsigned want = 0);
template <class T>
ttm_status_t launch_skinny(const Fold &f, const T *a, const T *b, T *c, hipStream_t s, unsigned want_splits = 0);

const char *kernel_name(Kernel k);
Kernel kernel_from_name(const char *name);

// Kernel the planner picks for a fold (aligned: all base pointers 16-byte
// aligned).  `force` (may be Kernel::none) overrides when applicable.
Kernel choose_kernel(const Fold &f, bool is_double, ttm_precision_t prec, bool aligned, Kernel force);

// Template instance of the family's kernel for a fold (as profilers print it);
// `aligned`: A, B and C 16-byte aligned (selects the NN tile without padding).
std::string kernel_symbol(Kernel k, const Fold &f, bool is_double, ttm_precision_t prec, bool aligned = true);

// Contraction splits a launch of family k runs with: `want` (0: the cached
// tuning of the fold's class, else the family's own choice), clamped by the
// family's limits; 1 for families that do not split.
unsigned launch_splits(Kernel k, const Fold &f, bool is_double, unsigned want);

// What a launch ran (for diagnostics).
struct LaunchInfo
{
    Kernel kernel = Kernel::none;
    unsigned splits = 1;
    bool aligned = true;
};

// Launch the product on `stream`; returns TTM_SUCCESS or TTM_ERR_HIP.
// want_splits as for launch_splits.
ttm_status_t launch_f32(const Fold &f, const float *a, const float *b, float *c, ttm_precision_t prec,
                        hipStream_t stream, Kernel force, LaunchInfo *info, unsigned want_splits = 0);
ttm_status_t launch_f64(const Fold &f, const double *a, const double *b, double *c, hipStream_t stream,
                        Kernel force, LaunchInfo *info, unsigned want_splits = 0);

} // namespace ttm_amd
