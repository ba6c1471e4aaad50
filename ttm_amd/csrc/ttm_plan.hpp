// Host-side planner: validates a reference-style call and folds any runtime
// (q, pi) into the batched product
//
//     C[k][j][i] = sum_{l < n_q} B(j, l) * A[k][l][i]
//     k < outer,  j < m,  i < w,
//
// where w is the product of the extents of the modes stored below q in the
// layout (contiguous block) and outer the product of those above it.  For
// dense tensors slab k of A is the contiguous n_q x w block at k*n_q*w and
// slab k of C the m x w block at k*m*w.  This is the form the reference's
// subtensor executor walks one GEMM at a time
// (include/tlib/detail/ttm.h:1135-1194); its cases 1..7
// (include/tlib/detail/mtm.h:145-223) are the special cases w == 1 or
// outer == 1 of the same form.
#pragma once

#include <cstddef>
#include <cstdint>

#include "ttm_amd.h"

namespace ttm_amd
{

struct Fold
{
    uint64_t outer = 1; // slabs above mode q
    uint64_t nq = 1;    // contraction extent
    uint64_t w = 1;     // contiguous block below mode q
    uint64_t m = 1;     // rows of B = extent of C in mode q
    bool b_colmajor = true; // pib[0] == 1: B(j,l) at j + l*m, else at j*nq + l
    unsigned qhat = 1;  // one-based position of q in the layout
    int ref_case = 0;   // the reference's case number 1..8 (reporting only)
    uint64_t numel_a() const { return outer * nq * w; }
    uint64_t numel_c() const { return outer * m * w; }
    uint64_t numel_b() const { return m * nq; }
    uint64_t b_stride_j() const { return b_colmajor ? 1 : nq; }
    uint64_t b_stride_l() const { return b_colmajor ? m : 1; }
};

// Validation in the reference's order (include/tlib/ttm.h:65-90), then the
// checks the reference leaves out (pib, C's non-q extents, density).  On
// failure the thread's last-error message is set.
ttm_status_t plan(unsigned q, unsigned p, const void *a, const size_t *na, const size_t *wa,
                  const size_t *pia, const void *b, const size_t *nb, const size_t *pib,
                  const void *c, const size_t *nc, const size_t *wc, Fold &f);

// Thread-local last error.
void set_error(ttm_status_t st, const char *detail = nullptr);
void clear_error();
const char *last_error();
const char *status_text(ttm_status_t st);

} // namespace ttm_amd
