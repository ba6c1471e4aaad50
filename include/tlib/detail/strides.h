// Strides of dense tensors in a given layout (names as
// include/tlib/detail/strides.h:31-101 of the reference).
#pragma once

#include <array>
#include <cstddef>
#include <iterator>
#include <stdexcept>
#include <vector>

#include "layout.h"
#include "shape.h"

namespace tlib::ttm::detail
{

// Dense strides w[pi_r] = w[pi_{r-1}] * n[pi_{r-1}], w[pi_1] = 1.  Vectors and
// scalars get all-ones strides, exactly as the reference does
// (reference strides.h:31-57, same error texts).
template <class InputIt1, class InputIt2, class OutputIt>
inline void compute_strides(InputIt1 shape_begin, InputIt1 shape_end, InputIt2 layout_begin,
                            OutputIt strides_begin)
{
    if (!is_valid_shape(shape_begin, shape_end))
        throw std::runtime_error("Error in tlib::detail::compute_strides(): input shape is not valid.");
    auto const p = static_cast<std::size_t>(std::distance(shape_begin, shape_end));
    if (!is_valid_layout(layout_begin, layout_begin + p))
        throw std::runtime_error("Error in tlib::detail::compute_strides(): input layout is not valid.");
    for (std::size_t r = 0; r < p; ++r) strides_begin[r] = 1u;
    if (is_vector(shape_begin, shape_end) || is_scalar(shape_begin, shape_end)) return;
    if (!is_matrix(shape_begin, shape_end) && !is_tensor(shape_begin, shape_end)) return;
    for (std::size_t r = 1; r < p; ++r) {
        auto const cur = static_cast<std::size_t>(layout_begin[r]) - 1u;
        auto const prev = static_cast<std::size_t>(layout_begin[r - 1]) - 1u;
        strides_begin[cur] = strides_begin[prev] * shape_begin[prev];
    }
}

template <class size_type>
inline auto generate_strides(std::vector<size_type> const &shape, std::vector<size_type> const &layout)
{
    auto strides = std::vector<size_type>(shape.size());
    compute_strides(shape.begin(), shape.end(), layout.begin(), strides.begin());
    return strides;
}

template <class size_type, std::size_t N>
inline auto generate_strides(std::array<size_type, N> const &shape, std::array<size_type, N> const &layout)
{
    static_assert(N > 0, "Static error in tlib::detail::generate_strides(): N, i.e. length of array should be greater than zero.");
    auto strides = std::array<size_type, N>{};
    compute_strides(shape.begin(), shape.end(), layout.begin(), strides.begin());
    return strides;
}

// Strides must be non-decreasing along the layout (reference strides.h:76-101).
template <class InputIt1, class InputIt2>
inline bool is_valid_strides(InputIt1 layout_begin, InputIt1 layout_end, InputIt2 stride_begin)
{
    if (!is_valid_layout(layout_begin, layout_end))
        throw std::runtime_error("Error in tlib::detail::is_valid_strides(): input layout is not valid.");
    auto const p = static_cast<std::size_t>(std::distance(layout_begin, layout_end));
    if (p == 1 && stride_begin[0] == 1) return true;
    for (std::size_t r = 0; r + 1 < p; ++r) {
        auto const lo = static_cast<std::size_t>(layout_begin[r]) - 1u;
        auto const hi = static_cast<std::size_t>(layout_begin[r + 1]) - 1u;
        if (stride_begin[lo] > stride_begin[hi]) return false;
    }
    return true;
}

} // namespace tlib::ttm::detail
