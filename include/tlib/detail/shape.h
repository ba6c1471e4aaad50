// Shape predicates of the drop-in API (namespace and names as
// include/tlib/detail/shape.h:31-90 of the reference).
#pragma once

#include <algorithm>
#include <cassert>
#include <cstddef>
#include <iterator>
#include <stdexcept>
#include <vector>

namespace tlib::ttm::detail
{

// Non-empty and no zero extent (reference shape.h:31-35).
template <class InputIt>
inline bool is_valid_shape(InputIt begin, InputIt end)
{
    if (begin == end) return false;
    for (auto it = begin; it != end; ++it)
        if (*it == 0u) return false;
    return true;
}

// All extents are one (reference shape.h:39-47).
template <class InputIt>
inline bool is_scalar(InputIt begin, InputIt end)
{
    if (!is_valid_shape(begin, end)) return false;
    for (auto it = begin; it != end; ++it)
        if (*it != 1u) return false;
    return true;
}

// One of the first two extents is one, the other larger, all later ones are
// one; a single extent counts when it is larger than one (reference shape.h:50-67).
template <class InputIt>
inline bool is_vector(InputIt begin, InputIt end)
{
    if (!is_valid_shape(begin, end)) return false;
    auto const p = static_cast<std::size_t>(std::distance(begin, end));
    if (p == 1) return *begin > 1u;
    auto const n0 = begin[0], n1 = begin[1];
    for (std::size_t r = 2; r < p; ++r)
        if (begin[r] != 1u) return false;
    return (n0 > 1u || n1 > 1u) && (n0 == 1u || n1 == 1u);
}

// First two extents larger than one, all later ones one (reference shape.h:69-79).
template <class InputIt>
inline bool is_matrix(InputIt begin, InputIt end)
{
    if (!is_valid_shape(begin, end)) return false;
    auto const p = static_cast<std::size_t>(std::distance(begin, end));
    if (p < 2) return false;
    if (!(begin[0] > 1u && begin[1] > 1u)) return false;
    for (std::size_t r = 2; r < p; ++r)
        if (begin[r] != 1u) return false;
    return true;
}

// Order three or more with some extent beyond the second larger than one
// (reference shape.h:82-90).
template <class InputIt>
inline bool is_tensor(InputIt begin, InputIt end)
{
    if (!is_valid_shape(begin, end)) return false;
    auto const p = static_cast<std::size_t>(std::distance(begin, end));
    if (p < 3) return false;
    for (std::size_t r = 2; r < p; ++r)
        if (begin[r] > 1u) return true;
    return false;
}

} // namespace tlib::ttm::detail
