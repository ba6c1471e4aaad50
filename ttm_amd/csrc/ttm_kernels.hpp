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
    count
};

const char *kernel_name(Kernel k);
Kernel kernel_from_name(const char *name);

// Kernel the planner picks for a fold (aligned: all base pointers 16-byte
// aligned).  `force` (may be Kernel::none) overrides when applicable.
Kernel choose_kernel(const Fold &f, bool is_double, ttm_precision_t prec, bool aligned, Kernel force);

// Template instance of the family's kernel for a fold (as profilers print it).
std::string kernel_symbol(Kernel k, const Fold &f, bool is_double, ttm_precision_t prec);

// Launch the product on `stream`; returns TTM_SUCCESS or TTM_ERR_HIP.
ttm_status_t launch_f32(const Fold &f, const float *a, const float *b, float *c, ttm_precision_t prec,
                        hipStream_t stream, Kernel force, Kernel *used);
ttm_status_t launch_f64(const Fold &f, const double *a, const double *b, double *c, hipStream_t stream,
                        Kernel force, Kernel *used);

} // namespace ttm_amd
