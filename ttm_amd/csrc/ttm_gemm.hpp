// LDS-staged MFMA kernels (plain FP32 / FP64) for the folded product.
#pragma once

#include <hip/hip_runtime.h>

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

} // namespace ttm_amd
