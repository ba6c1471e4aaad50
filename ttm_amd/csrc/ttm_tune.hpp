// Contraction-split autotuner of single products: the split count that the
// timing in ttm_stune / ttm_dtune found fastest, cached per shape class.
//
// A class is the kernel family, element type, view (w == 1 or not), B's
// orientation and the power-of-two classes of m, n_q and the column count
// (outer * w).  q and the layout enter only through the fold they produce
// (which of the views, and the extents w and outer), so the class also
// stands for every (q, pi) that folds alike.  The cache is process-wide;
// with TTM_AMD_SPLIT_CACHE set it is read from and written to that file.
#pragma once

#include <string>

#include "ttm_kernels.hpp"
#include "ttm_plan.hpp"

namespace ttm_amd
{

std::string split_class_key(Kernel k, const Fold &f, bool is_double);
// Cached split count of the fold's class (0: none cached).
unsigned tuned_splits(Kernel k, const Fold &f, bool is_double);
void store_tuned_splits(Kernel k, const Fold &f, bool is_double, unsigned splits);
void clear_tuned_splits();

} // namespace ttm_amd
