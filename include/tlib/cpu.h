// Host executor of the mode-q tensor-times-matrix product, for callers
// without a GPU: the reference's call forms, with its policy tag triples
// choosing how the work is cut and scheduled over host threads
// (include/tlib/detail/ttm.h:405-1474 of the reference), through
// ttm_sttm_cpu / ttm_dttm_cpu of libttm_amd.so.
//
//   tlib::ttm::cpu::ttm(ep, sp, fp, q, p, a, na, wa, pia, b, nb, pib, c, nc, wc)
//   tlib::ttm::cpu::ttm(parallel_loop_blas, sp, all, q, p, ratio, a, ...)
//   tlib::ttm::cpu::ttm(q, A, B, ep, sp, fp) -> tensor
//
// Same validation and std::runtime_error texts as tlib::ttm::ttm.  A tag
// triple the reference does not define prints the reference's message and
// leaves C untouched, as the reference does (include/tlib/detail/ttm.h:391-400).
// Every triple and thread count gives the same bits.
#pragma once

#include <cstddef>
#include <iostream>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "../ttm_amd.h"
#include "ttm.h"

namespace tlib::ttm::cpu
{

namespace detail
{
constexpr ttm_cpu_parallel_t policy(parallel_policy::sequential_t) { return TTM_CPU_SEQUENTIAL; }
constexpr ttm_cpu_parallel_t policy(parallel_policy::parallel_blas_t) { return TTM_CPU_PARALLEL_BLAS; }
constexpr ttm_cpu_parallel_t policy(parallel_policy::parallel_loop_t) { return TTM_CPU_PARALLEL_LOOP; }
constexpr ttm_cpu_parallel_t policy(parallel_policy::parallel_taskloop_t) { return TTM_CPU_PARALLEL_TASKLOOP; }
constexpr ttm_cpu_parallel_t policy(parallel_policy::parallel_task_t) { return TTM_CPU_PARALLEL_TASK; }
constexpr ttm_cpu_parallel_t policy(parallel_policy::parallel_loop_blas_t) { return TTM_CPU_PARALLEL_LOOP_BLAS; }
constexpr ttm_cpu_parallel_t policy(parallel_policy::batched_gemm_t) { return TTM_CPU_BATCHED_GEMM; }
constexpr ttm_cpu_parallel_t policy(parallel_policy::combined_t) { return TTM_CPU_COMBINED; }
constexpr ttm_cpu_slicing_t policy(slicing_policy::slice_t) { return TTM_CPU_SLICE; }
constexpr ttm_cpu_slicing_t policy(slicing_policy::subtensor_t) { return TTM_CPU_SUBTENSOR; }
constexpr ttm_cpu_slicing_t policy(slicing_policy::combined_t) { return TTM_CPU_SLICING_COMBINED; }
constexpr ttm_cpu_fusion_t policy(fusion_policy::none_t) { return TTM_CPU_FUSION_NONE; }
constexpr ttm_cpu_fusion_t policy(fusion_policy::outer_t) { return TTM_CPU_FUSION_OUTER; }
constexpr ttm_cpu_fusion_t policy(fusion_policy::all_t) { return TTM_CPU_FUSION_ALL; }

template <class value_t, class size_t>
inline void run(ttm_cpu_parallel_t par, ttm_cpu_slicing_t sl, ttm_cpu_fusion_t fu, double ratio, unsigned q,
                unsigned p, const value_t *a, size_t const *na, size_t const *wa, size_t const *pia,
                const value_t *b, size_t const *nb, size_t const *pib, value_t *c, size_t const *nc,
                size_t const *wc)
{
    static_assert(std::is_same_v<value_t, float> || std::is_same_v<value_t, double>,
                  "tlib::ttm::cpu::ttm supports float and double elements.");
    auto const len = static_cast<std::size_t>(p);
    ttm::detail::size_array<size_t> na_(na, len), wa_(wa, len), pia_(pia, len), nc_(nc, len), wc_(wc, len);
    ttm::detail::size_array<size_t> nb_(nb, 2), pib_(pib, 2);
    ttm_status_t st;
    if constexpr (std::is_same_v<value_t, float>)
        st = ttm_sttm_cpu(q, p, a, na_.ptr, wa_.ptr, pia_.ptr, b, nb_.ptr, pib_.ptr, c, nc_.ptr, wc_.ptr, par, sl, fu,
                          0, ratio);
    else
        st = ttm_dttm_cpu(q, p, a, na_.ptr, wa_.ptr, pia_.ptr, b, nb_.ptr, pib_.ptr, c, nc_.ptr, wc_.ptr, par, sl, fu,
                          0, ratio);
    if (st == TTM_ERR_POLICY) {
        std::cout << "This type of ttm is not yet defined" << std::endl;
        return;
    }
    ttm::detail::throw_if_failed(st);
}
} // namespace detail

/** The reference's low-level call (include/tlib/ttm.h:54-93) on the host. */
template <class value_t, class size_t, class parallel_policy, class slicing_policy, class fusion_policy>
inline void ttm(parallel_policy ep, slicing_policy sp, fusion_policy fp, unsigned const q, unsigned const p,
                const value_t *a, size_t const *const na, size_t const *const wa, size_t const *const pia,
                const value_t *b, size_t const *const nb, size_t const *const pib,
                value_t *c, size_t const *const nc, size_t const *const wc)
{
    detail::run(detail::policy(ep), detail::policy(sp), detail::policy(fp), 0.0, q, p, a, na, wa, pia, b, nb, pib, c,
                nc, wc);
}

/** parallel_loop_blas with the share `ratio` of the threads on the outer
 * loop (include/tlib/detail/ttm.h:841-912, 1259-1320). */
template <class value_t, class size_t, class slicing_policy>
inline void ttm(parallel_policy::parallel_loop_blas_t ep, slicing_policy sp, fusion_policy::all_t fp,
                unsigned const q, unsigned const p, double ratio,
                const value_t *a, size_t const *const na, size_t const *const wa, size_t const *const pia,
                const value_t *b, size_t const *const nb, size_t const *const pib,
                value_t *c, size_t const *const nc, size_t const *const wc)
{
    detail::run(detail::policy(ep), detail::policy(sp), detail::policy(fp), ratio, q, p, a, na, wa, pia, b, nb, pib,
                c, nc, wc);
}

/** Tensor-object form (include/tlib/ttm.h:100-125) on the host. */
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
    nc[q - 1] = b.shape().at(0);
    auto c = tensor<value_t>(nc, a.layout());
    cpu::ttm(ep, sp, fp, static_cast<unsigned>(q), static_cast<unsigned>(p), a.data().data(), a.shape().data(),
             a.strides().data(), a.layout().data(), b.data().data(), b.shape().data(), b.layout().data(),
             c.data().data(), c.shape().data(), c.strides().data());
    return c;
}

} // namespace tlib::ttm::cpu
