// Validation and (q, pi) folding for the C ABI.
#include "ttm_plan.hpp"

#include <string>

#include "tlib/detail/cases.h"
#include "tlib/detail/layout.h"
#include "tlib/detail/shape.h"
#include "tlib/detail/strides.h"

namespace ttm_amd
{

namespace
{
thread_local std::string g_last_error;

// Messages 1..18 are the reference's texts (include/tlib/ttm.h:65-90).
const char *const k_messages[] = {
    "",
    "Error in tlib::tensor_times_matrix: input tensor order should be greater zero.",
    "Error in tlib::tensor_times_matrix: contraction mode should be greater zero or less than or equal to p.",
    "Error in tlib::tensor_times_matrix: pointer to input tensor A should not be zero.",
    "Error in tlib::tensor_times_matrix: pointer to input vector B should not be zero.",
    "Error in tlib::tensor_times_matrix: pointer to output tensor C should not be zero.",
    "Error in tlib::tensor_times_matrix: pointer to input tensor shape vector na should not be zero.",
    "Error in tlib::tensor_times_matrix: pointer to input vector shape vector nb should not be zero.",
    "Error in tlib::tensor_times_matrix: pointer to output tensor shape vector nc should not be zero.",
    "Error in tlib::tensor_times_matrix: pointer to input tensor stride vector wa should not be zero.",
    "Error in tlib::tensor_times_matrix: pointer to output tensor stride vector wc should not be zero.",
    "Error in tlib::tensor_times_matrix: pointer to input tensor permutation vector pia should not be zero.",
    "Error in tlib::tensor_times_matrix: contraction dimension of A and B are not equal.",
    "Error in tlib::tensor_times_matrix: free dimension of C and B are not equal.",
    "Error in tlib::tensor_times_matrix: shape vector of A is not valid.",
    "Error in tlib::tensor_times_matrix: shape vector of C is not valid.",
    "Error in tlib::tensor_times_matrix: layout vector of A is not valid.",
    "Error in tlib::tensor_times_matrix: stride vector of A is not valid.",
    "Error in tlib::tensor_times_matrix: stride vector of C is not valid.",
    "Error in tlib::tensor_times_matrix: pointer to input matrix permutation vector pib should not be zero.",
    "Error in tlib::tensor_times_matrix: tensors must be contiguous, i.e. dense in their layout.",
    "Error in tlib::tensor_times_matrix: shape vector of C must equal that of A outside the contraction mode.",
    "Error in tlib::tensor_times_matrix: the requested precision is not available for this element type.",
    "Error in tlib::tensor_times_matrix: HIP runtime error.",
    "Error in tlib::tensor_times_matrix: device memory allocation failed.",
    "Error in tlib::tensor_times_matrix: invalid argument.",
    // the reference prints this text for a tag triple it does not define
    // (include/tlib/detail/ttm.h:399) and does nothing
    "Error in tlib::tensor_times_matrix: This type of ttm is not yet defined",
};
constexpr int k_num_messages = sizeof(k_messages) / sizeof(k_messages[0]);

// Every mode with extent > 1 must carry the dense stride of its layout
// position; extent-1 modes are never addressed, so their stride is free
// (this also admits the reference's all-ones strides for vector shapes,
// include/tlib/detail/strides.h:44-48).
bool is_dense(const size_t *n, const size_t *w, const size_t *pi, unsigned p)
{
    uint64_t expect = 1;
    for (unsigned r = 0; r < p; ++r) {
        size_t const mode = pi[r] - 1;
        if (n[mode] > 1 && w[mode] != expect) return false;
        expect *= n[mode];
    }
    return true;
}
} // namespace

const char *status_text(ttm_status_t st)
{
    int const i = static_cast<int>(st);
    if (i < 0 || i >= k_num_messages) return "Error in tlib::tensor_times_matrix: unknown status.";
    return k_messages[i];
}

void set_error(ttm_status_t st, const char *detail)
{
    g_last_error = status_text(st);
    if (detail && *detail) {
        g_last_error += " (";
        g_last_error += detail;
        g_last_error += ")";
    }
}

void clear_error() { g_last_error.clear(); }

const char *last_error() { return g_last_error.c_str(); }

ttm_status_t plan(unsigned q, unsigned p, const void *a, const size_t *na, const size_t *wa,
                  const size_t *pia, const void *b, const size_t *nb, const size_t *pib,
                  const void *c, const size_t *nc, const size_t *wc, Fold &f)
{
    using namespace tlib::ttm::detail;
    auto fail = [](ttm_status_t st) {
        set_error(st);
        return st;
    };
    // include/tlib/ttm.h:65-90, same order
    if (p == 0) return fail(TTM_ERR_ORDER_ZERO);
    if (q == 0 || q > p) return fail(TTM_ERR_MODE);
    if (!a) return fail(TTM_ERR_NULL_A);
    if (!b) return fail(TTM_ERR_NULL_B);
    if (!c) return fail(TTM_ERR_NULL_C);
    if (!na) return fail(TTM_ERR_NULL_NA);
    if (!nb) return fail(TTM_ERR_NULL_NB);
    if (!nc) return fail(TTM_ERR_NULL_NC);
    if (!wa) return fail(TTM_ERR_NULL_WA);
    if (!wc) return fail(TTM_ERR_NULL_WC);
    if (!pia) return fail(TTM_ERR_NULL_PIA);
    if (na[q - 1] != nb[1]) return fail(TTM_ERR_CONTRACTION_DIM);
    if (nc[q - 1] != nb[0]) return fail(TTM_ERR_FREE_DIM);
    if (!is_valid_shape(na, na + p)) return fail(TTM_ERR_SHAPE_A);
    if (!is_valid_shape(nc, nc + p)) return fail(TTM_ERR_SHAPE_C);
    if (!is_valid_layout(pia, pia + p)) return fail(TTM_ERR_LAYOUT_A);
    if (!is_valid_strides(pia, pia + p, wa)) return fail(TTM_ERR_STRIDES_A);
    if (!is_valid_strides(pia, pia + p, wc)) return fail(TTM_ERR_STRIDES_C);
    // beyond the reference
    if (!pib) return fail(TTM_ERR_NULL_PIB);
    for (unsigned r = 0; r < p; ++r)
        if (r != q - 1 && nc[r] != na[r]) return fail(TTM_ERR_SHAPE_C_MISMATCH);
    if (!is_dense(na, wa, pia, p) || !is_dense(nc, wc, pia, p)) return fail(TTM_ERR_NOT_DENSE);

    unsigned const qh = static_cast<unsigned>(inverse_mode(pia, pia + p, q));
    f = Fold{};
    f.qhat = qh;
    f.nq = na[q - 1];
    f.m = nb[0];
    f.b_colmajor = pib[0] == 1;
    for (unsigned r = 0; r + 1 < qh; ++r) f.w *= na[pia[r] - 1];
    for (unsigned r = qh; r < p; ++r) f.outer *= na[pia[r] - 1];
    for (int k = 1; k <= 8 && f.ref_case == 0; ++k) {
        bool hit = false;
        switch (k) {
        case 1: hit = is_case<1>(p, q, pia); break;
        case 2: hit = is_case<2>(p, q, pia); break;
        case 3: hit = is_case<3>(p, q, pia); break;
        case 4: hit = is_case<4>(p, q, pia); break;
        case 5: hit = is_case<5>(p, q, pia); break;
        case 6: hit = is_case<6>(p, q, pia); break;
        case 7: hit = is_case<7>(p, q, pia); break;
        case 8: hit = is_case<8>(p, q, pia); break;
        }
        if (hit) f.ref_case = k;
    }
    clear_error();
    return TTM_SUCCESS;
}

} // namespace ttm_amd
