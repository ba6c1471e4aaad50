// Layout (mode permutation) utilities of the drop-in API (names as
// include/tlib/detail/layout.h:29-144 of the reference).  A layout pi is a
// one-based permutation of 1..p; pi[0] is the fastest-varying mode.
#pragma once

#include <algorithm>
#include <cstddef>
#include <iterator>
#include <stdexcept>
#include <type_traits>
#include <vector>

namespace tlib::ttm::detail
{

// True for a permutation of 1..p with p > 0 (reference layout.h:29-55).
template <class InputIt>
inline bool is_valid_layout(InputIt begin, InputIt end)
{
    auto const d = std::distance(begin, end);
    if (d <= 0) return false;
    auto const p = static_cast<std::size_t>(d);
    std::vector<bool> seen(p + 1, false);
    for (auto it = begin; it != end; ++it) {
        auto const v = static_cast<std::size_t>(*it);
        if (v == 0u || v > p || seen[v]) return false;
        seen[v] = true;
    }
    return true;
}

// Writes the k-order layout {k, k-1, .., 1, k+1, .., p}; k == 0 selects the
// last order {p, .., 1} and k > p is clamped to p (reference layout.h:57-78).
template <class OutputIt, class size_t>
inline void compute_k_order_layout(OutputIt begin, OutputIt end, size_t k)
{
    auto const d = std::distance(begin, end);
    if (d <= 0)
        throw std::runtime_error(
            "Error in tlib::ttm::detail::compute_k_order: range provided by begin and end not correct!");
    auto const p = static_cast<std::size_t>(d);
    std::size_t kk = static_cast<std::size_t>(k);
    if (kk == 0u) kk = p;
    kk = std::min(kk, p);
    using value_type = std::remove_reference_t<decltype(*begin)>;
    std::size_t r = 0;
    for (; r < kk; ++r) begin[r] = static_cast<value_type>(kk - r);
    for (; r < p; ++r) begin[r] = static_cast<value_type>(r + 1);
}

template <class size_t>
inline auto generate_k_order_layout(size_t p, size_t k)
{
    std::vector<size_t> v(p);
    compute_k_order_layout(v.begin(), v.end(), k);
    return v;
}

// pi == {1, 2, .., p} (reference layout.h:86-93).
template <class InputIt>
inline bool is_first_order(InputIt begin, InputIt end)
{
    std::size_t expect = 1;
    for (auto it = begin; it != end; ++it, ++expect)
        if (static_cast<std::size_t>(*it) != expect) return false;
    return true;
}

// pi == {p, .., 1} (reference layout.h:95-102).
template <class InputIt>
inline bool is_last_order(InputIt begin, InputIt end)
{
    auto expect = static_cast<std::size_t>(std::distance(begin, end));
    for (auto it = begin; it != end; ++it, --expect)
        if (static_cast<std::size_t>(*it) != expect) return false;
    return true;
}

template <class OutputIt>
inline void compute_first_order_layout(OutputIt begin, OutputIt end)
{
    compute_k_order_layout(begin, end, 1u);
}

template <class OutputIt>
inline void compute_last_order_layout(OutputIt begin, OutputIt end)
{
    compute_k_order_layout(begin, end, 0u);
}

// One-based position of the one-based `mode` in the layout, i.e. pi^{-1}(mode)
// (reference layout.h:120-144, same error texts).
template <class InputIt, class SizeType>
inline auto inverse_mode(InputIt layout_begin, InputIt layout_end, SizeType mode)
{
    using value_type = typename std::iterator_traits<InputIt>::value_type;
    if (!is_valid_layout(layout_begin, layout_end))
        throw std::runtime_error("Error in tlib::ttm::detail::inverse_mode(): input layout is not valid.");
    auto const p = static_cast<value_type>(std::distance(layout_begin, layout_end));
    if (mode == 0u || mode > SizeType(p))
        throw std::runtime_error(
            "Error in tlib::ttm::detail::inverse_mode(): mode should be one-based and equal to or less than layout size.");
    value_type r = 0;
    while (r < p && layout_begin[r] != value_type(mode)) ++r;
    return value_type(r + 1);
}

} // namespace tlib::ttm::detail
