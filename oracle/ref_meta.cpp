// ref_meta.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A thin extern "C" shim over the reference's own metadata headers, compiled
// where they lie (REF_INC=/root/reference/include) into oracle/_ref/.  Only
// the metadata part of the reference builds in this image: detail/ttm.h and
// detail/mtm.h need a CBLAS header that the image does not have, so the
// contraction itself is not built here (the C oracle restates it instead).
//
// The shim lets the test-suite compare the oracle's and the product's
// shape / layout / strides / inverse_mode / is_case logic with the real
// reference on exhaustive small inputs.  Functions that throw in the
// reference return a sentinel here (0 or -1).
#include <cassert>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include <tlib/detail/cases.h>
#include <tlib/detail/layout.h>
#include <tlib/detail/shape.h>
#include <tlib/detail/strides.h>
#include <tlib/detail/tensor.h>

using namespace tlib::ttm;

extern "C" {

int ref_is_valid_shape(const size_t *n, size_t p) { return detail::is_valid_shape(n, n + p); }
int ref_is_scalar(const size_t *n, size_t p) { return detail::is_scalar(n, n + p); }
int ref_is_vector(const size_t *n, size_t p) { return detail::is_vector(n, n + p); }
int ref_is_matrix(const size_t *n, size_t p) { return detail::is_matrix(n, n + p); }
int ref_is_tensor(const size_t *n, size_t p) { return detail::is_tensor(n, n + p); }
int ref_is_valid_layout(const size_t *pi, size_t p) { return detail::is_valid_layout(pi, pi + p); }
int ref_is_first_order(const size_t *pi, size_t p) { return detail::is_first_order(pi, pi + p); }
int ref_is_last_order(const size_t *pi, size_t p) { return detail::is_last_order(pi, pi + p); }

int ref_compute_k_order_layout(size_t *pi, size_t p, size_t k)
{
    try {
        detail::compute_k_order_layout(pi, pi + p, k);
        return 0;
    } catch (std::exception const &) {
        return -1;
    }
}

size_t ref_inverse_mode(const size_t *pi, size_t p, size_t mode)
{
    try {
        return detail::inverse_mode(pi, pi + p, mode);
    } catch (std::exception const &) {
        return 0;
    }
}

int ref_compute_strides(const size_t *n, const size_t *pi, size_t p, size_t *w)
{
    try {
        detail::compute_strides(n, n + p, pi, w);
        return 0;
    } catch (std::exception const &) {
        return -1;
    }
}

int ref_is_valid_strides(const size_t *pi, size_t p, const size_t *w)
{
    try {
        return detail::is_valid_strides(pi, pi + p, w) ? 1 : 0;
    } catch (std::exception const &) {
        return -1;
    }
}

int ref_is_case(unsigned case_nr, unsigned p, size_t q, const size_t *pi)
{
    switch (case_nr) {
    case 1: return detail::is_case<1u>(p, q, pi);
    case 2: return detail::is_case<2u>(p, q, pi);
    case 3: return detail::is_case<3u>(p, q, pi);
    case 4: return detail::is_case<4u>(p, q, pi);
    case 5: return detail::is_case<5u>(p, q, pi);
    case 6: return detail::is_case<6u>(p, q, pi);
    case 7: return detail::is_case<7u>(p, q, pi);
    case 8: return detail::is_case<8u>(p, q, pi);
    default: return 0;
    }
}

// tensor(n, pi): 0 on success (strides written to w), -1 if it throws.
int ref_tensor_strides(const size_t *n, const size_t *pi, size_t p, size_t *w)
{
    try {
        auto t = tensor<float>(std::vector<size_t>(n, n + p), std::vector<size_t>(pi, pi + p));
        auto s = t.strides();
        for (size_t r = 0; r < p; ++r) w[r] = s[r];
        return 0;
    } catch (std::exception const &) {
        return -1;
    }
}

} // extern "C"
