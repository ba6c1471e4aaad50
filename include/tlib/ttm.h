// Drop-in replacement for the reference's public header include/tlib/ttm.h.
//
// The three call forms of the reference compile unchanged against this
// header and run on an MI355X through the C ABI of libttm_amd.so:
//   1. tlib::ttm::ttm(ep, sp, fp, q, p, a, na, wa, pia, b, nb, pib, c, nc, wc)
//        (reference include/tlib/ttm.h:54-93)
//   2. tlib::ttm::ttm(q, A, B, ep, sp, fp) -> tensor      (:100-125)
//   3. A(q) * B                                             (:133-138)
// All policy tag triples are accepted; they select CPU threading strategies
// in the reference and have no effect on the GPU planner.
//
// Semantics kept from the reference: validation in the same order with the
// same std::runtime_error texts, C fully overwritten, and the call returns
// when C is complete.  Buffers may be host memory (staged to the device per
// call) or device memory (used in place).
#pragma once

#include <cstddef>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "../ttm_amd.h"
#include "detail/cases.h"
#include "detail/layout.h"
#include "detail/shape.h"
#include "detail/strides.h"
#include "detail/tags.h"
#include "detail/tensor.h"

namespace tlib::ttm
{

namespace detail
{
inline void throw_if_failed(ttm_status_t st)
{
    if (st != TTM_SUCCESS) throw std::runtime_error(ttm_last_error());
}

// Index arrays are forwarded as they are when their element type is
// std::size_t and copied otherwise (the reference templates the size type).
template <class size_t>
struct size_array
{
    std::vector<std::size_t> copy;
    std::size_t const *ptr = nullptr;
    size_array(size_t const *src, std::size_t len)
    {
        if constexpr (std::is_same_v<size_t, std::size_t>) {
            ptr = src;
        } else if (src != nullptr) {
            copy.assign(src, src + len);
            ptr = copy.data();
        }
    }
};
} // namespace detail

/**
 * Mode-q tensor-times-matrix product
 *   C([k1, j, k2]) = sum_iq A([k1, iq, k2]) * B(j, iq)
 * with k1 = (i1..i(q-1)), k2 = (i(q+1)..ip).  Arguments as the reference
 * (include/tlib/ttm.h:24-61): q and p one-based, na/wa/pia the extents,
 * strides and layout of A (length p), nb = {m, n_q} and pib the layout of B
 * ({1,2} column-major, {2,1} row-major), nc/wc the extents and strides of C,
 * which has A's layout.
 */
template <class value_t, class size_t, class parallel_policy, class slicing_policy, class fusion_policy>
inline void ttm(parallel_policy, slicing_policy, fusion_policy, unsigned const q, unsigned const p,
                const value_t *a, size_t const *const na, size_t const *const wa, size_t const *const pia,
                const value_t *b, size_t const *const nb, size_t const *const pib,
                value_t *c, size_t const *const nc, size_t const *const wc)
{
    static_assert(std::is_same_v<value_t, float> || std::is_same_v<value_t, double>,
                  "tlib::ttm::ttm on MI355X supports float and double elements.");
    auto const len = static_cast<std::size_t>(p);
    detail::size_array<size_t> na_(na, len), wa_(wa, len), pia_(pia, len), nc_(nc, len), wc_(wc, len);
    detail::size_array<size_t> nb_(nb, 2), pib_(pib, 2);
    ttm_status_t st;
    if constexpr (std::is_same_v<value_t, float>)
        st = ttm_sttm_host(q, p, a, na_.ptr, wa_.ptr, pia_.ptr, b, nb_.ptr, pib_.ptr, c, nc_.ptr, wc_.ptr);
    else
        st = ttm_dttm_host(q, p, a, na_.ptr, wa_.ptr, pia_.ptr, b, nb_.ptr, pib_.ptr, c, nc_.ptr, wc_.ptr);
    detail::throw_if_failed(st);
}

/**
 * Tensor-object form (reference include/tlib/ttm.h:100-125): allocates C in
 * A's layout with extent nb[0] in mode q and returns it.
 */
template <class value_t, class execution_policy, class slicing_policy, class fusion_policy>
inline auto ttm(std::size_t q, ttm::tensor<value_t> const &a, ttm::tensor<value_t> const &b,
                execution_policy ep, slicing_policy sp, fusion_policy fp)
{
    auto const p = a.order();
    if (p == 0)
        throw std::runtime_error("Error in tlib::tensor_times_matrix: input tensor order should be greater zero.");
    if (q == 0 || q > p)
        throw std::runtime_error(
            "Error in tlib::tensor_times_matrix: contraction mode should be greater zero or less than or equal to p.");
    if (b.order() != 2)
        throw std::runtime_error("Error in tlib::tensor_times_matrix: input matrix B should have order 2.");

    auto nc = a.shape();
    auto const &nb = b.shape();
    nc[q - 1] = nb.at(0);
    auto c = tensor<value_t>(nc, a.layout());
    ttm(ep, sp, fp, static_cast<unsigned>(q), static_cast<unsigned>(p),
        a.data().data(), a.shape().data(), a.strides().data(), a.layout().data(),
        b.data().data(), b.shape().data(), b.layout().data(),
        c.data().data(), c.shape().data(), c.strides().data());
    return c;
}

} // namespace tlib::ttm

/**
 * C = A(q) * B (reference include/tlib/ttm.h:133-138).
 */
template <class value_t>
inline auto operator*(tlib::ttm::tensor_view<value_t> const &a, tlib::ttm::tensor<value_t> const &b)
{
    return tlib::ttm::ttm(a.contraction_mode(), a.get_tensor(), b, tlib::ttm::parallel_policy::combined,
                          tlib::ttm::slicing_policy::combined, tlib::ttm::fusion_policy::all);
}
