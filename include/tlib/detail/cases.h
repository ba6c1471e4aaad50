// Classification of a (p, q, pi) problem into the eight cases of the
// reference (include/tlib/detail/cases.h:25-37): p == 1 (case 1), the four
// order-2 cases (2..5), q fastest (6), q slowest (7) and q in the middle (8).
// On the GPU every case folds into the same batched product; the
// classification is kept for API compatibility and for reporting.
#pragma once

#include <cstddef>

namespace tlib::ttm::detail
{

template <unsigned case_nr>
inline constexpr bool is_case(unsigned p, std::size_t q, std::size_t const *const pi)
{
    static_assert(case_nr >= 1u && case_nr <= 8u, "tlib::ttm::detail::is_case: only 8 cases from 1 to 8 are covered.");
    if constexpr (case_nr == 1u) return p == 1u;
    if constexpr (case_nr == 2u) return p == 2u && q == 1u && pi[0] == 1u;
    if constexpr (case_nr == 3u) return p == 2u && q == 2u && pi[0] == 1u;
    if constexpr (case_nr == 4u) return p == 2u && q == 1u && pi[0] == 2u;
    if constexpr (case_nr == 5u) return p == 2u && q == 2u && pi[0] == 2u;
    if constexpr (case_nr == 6u) return p >= 3u && pi[0] == q;
    if constexpr (case_nr == 7u) return p >= 3u && pi[p - 1] == q;
    if constexpr (case_nr == 8u) return p >= 3u && pi[0] != q && pi[p - 1] != q;
    return false;
}

} // namespace tlib::ttm::detail
