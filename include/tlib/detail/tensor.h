// Host tensor container and contraction view of the drop-in API (same class
// names, constructors and accessors as include/tlib/detail/tensor.h:39-121 of
// the reference; the printer reproduces its MATLAB-style text format,
// tensor.h:129-189).  Storage is a std::vector on the host; the product
// stages it to the GPU per call.
#pragma once

#include <cstddef>
#include <functional>
#include <numeric>
#include <ostream>
#include <stdexcept>
#include <vector>

#include "layout.h"
#include "shape.h"
#include "strides.h"

namespace tlib::ttm
{

template <class value_t>
class tensor;

// A tensor bound to a contraction mode, produced by tensor::operator()(q)
// and consumed by operator*(tensor_view, tensor).
template <class _value_t>
struct tensor_view
{
public:
    using value_t = _value_t;
    using tensor_t = tensor<value_t>;
    tensor_view() = delete;
    tensor_view(tensor_view const &) = delete;
    tensor_view &operator=(tensor_view const &) = delete;
    inline tensor_t const &get_tensor() const { return _tensor; }
    inline std::size_t contraction_mode() const { return _q; }

private:
    friend tensor_view tensor_t::operator()(std::size_t q) const;
    tensor_view(tensor_t const &t, std::size_t q) : _tensor(t), _q(q) {}
    tensor_t const &_tensor;
    std::size_t _q;
};

template <class value_t>
class tensor
{
public:
    using shape_t = std::vector<std::size_t>;
    using layout_t = std::vector<std::size_t>;
    using strides_t = std::vector<std::size_t>;
    using vector_t = std::vector<value_t>;
    using tensor_view_t = tensor_view<value_t>;

    tensor() = delete;

    // Zero-initialised dense tensor of shape n in layout pi.
    tensor(shape_t const &n, layout_t const &pi)
        : _n(n), _pi(pi),
          _data(std::accumulate(n.begin(), n.end(), std::size_t{1}, std::multiplies<>()))
    {
        if (n.size() != pi.size())
            throw std::runtime_error("Error in tlib::tensor: shape vector and layout vector must have the same length.");
        if (!detail::is_valid_shape(n.begin(), n.end()))
            throw std::runtime_error("Error in tlib::tensor: shape vector of tensor is not valid.");
        if (!detail::is_valid_layout(pi.begin(), pi.end()))
            throw std::runtime_error("Error in tlib::tensor: layout vector of tensor is not valid.");
    }

    // First-order (column-major) layout by default.
    explicit tensor(shape_t const &n) : tensor(n, detail::generate_k_order_layout(n.size(), std::size_t{1})) {}

    tensor &operator=(value_t v)
    {
        std::fill(_data.begin(), _data.end(), v);
        return *this;
    }

    tensor_view_t operator()(std::size_t contraction_mode) const
    {
        if (1ul > contraction_mode || contraction_mode > this->order())
            throw std::runtime_error(
                "Error in tlib::tensor: specified contraction mode should be greater than one and equal to or less than the order.");
        return tensor_view_t(*this, contraction_mode);
    }

    auto begin() const { return _data.begin(); }
    auto end() const { return _data.end(); }
    decltype(auto) begin() { return _data.begin(); }
    decltype(auto) end() { return _data.end(); }

    auto const &data() const { return _data; }
    auto &data() { return _data; }
    auto const &shape() const { return _n; }
    auto const &layout() const { return _pi; }
    auto strides() const { return detail::generate_strides(_n, _pi); }
    auto order() const { return _n.size(); }

    auto const &operator[](std::size_t k) const { return _data[k]; }
    auto &operator[](std::size_t k) { return _data[k]; }

private:
    shape_t _n;
    layout_t _pi;
    vector_t _data;
};

} // namespace tlib::ttm

namespace tlib::ttm::detail
{
// Recursive printer for order >= 3: one "cat(r,..." block per index of mode r.
template <class value_type>
void print_slices(std::ostream &out, tlib::ttm::tensor<value_type> const &a, std::size_t offset, unsigned r)
{
    auto const w = a.strides();
    auto const &n = a.shape();
    auto const p = a.order();
    if (r >= 3) {
        out << "cat(" << r << ",..." << std::endl;
        for (std::size_t ir = 0; ir < n[r - 1]; ++ir)
            print_slices(out, a, offset + ir * w[r - 1], r - 1);
        out << (r == p ? ");" : "), ...") << std::endl;
    } else if (r == 2) {
        out << "[ ... " << std::endl;
        for (std::size_t i1 = 0; i1 < n[0]; ++i1) {
            auto const row = offset + i1 * w[0];
            for (std::size_t i2 = 0; i2 + 1 < n[1]; ++i2) out << a[row + i2 * w[1]] << ", ";
            out << a[row + (n[1] - 1) * w[1]] << "; ..." << std::endl;
        }
        out << "],..." << std::endl;
    }
}
} // namespace tlib::ttm::detail

template <class value_type>
std::ostream &operator<<(std::ostream &out, tlib::ttm::tensor<value_type> const &a)
{
    auto const w = a.strides();
    auto const &n = a.shape();
    auto const p = a.order();
    if (p == 1u) {
        out << "[ " << std::endl;
        for (std::size_t i1 = 0; i1 < n[0]; ++i1) out << a[i1 * w[0]] << ", ";
        out << "];" << std::endl;
    } else if (p == 2u) {
        out << "[ ... " << std::endl;
        for (std::size_t i2 = 0; i2 < n[1]; ++i2) {
            for (std::size_t i1 = 0; i1 < n[0]; ++i1) out << a[i2 * w[1] + i1 * w[0]] << ", ";
            out << "..." << std::endl;
        }
        out << "];" << std::endl;
    } else {
        tlib::ttm::detail::print_slices(out, a, 0, static_cast<unsigned>(p));
    }
    return out;
}
