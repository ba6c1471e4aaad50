// Device-resident tensors for the drop-in API (no reference counterpart: the
// reference's tensor, include/tlib/detail/tensor.h:59-121, lives in host
// memory, so every product of the tensor-object and operator forms stages it
// through PCIe).  device_tensor keeps the data in HBM across calls:
//
//   tlib::ttm::device_tensor<float> A(host_tensor);      // upload once
//   auto C = A(2) * B;                                    // B: device_tensor, C stays on the device
//   auto D = tlib::ttm::ttm(1, C, B2);                    // chains without host round trips
//   tlib::ttm::tensor<float> h = D.to_host();             // download when needed
//   std::cout << D;                                       // prints like the host tensor
//
// Same shape / layout / strides / order accessors and the same error texts as
// the host tensor; the products return when C is complete.  Memory comes
// from the C ABI (ttm_device_alloc and friends), so this header needs no HIP
// include.
#pragma once

#include <cstddef>
#include <functional>
#include <numeric>
#include <ostream>
#include <stdexcept>
#include <type_traits>
#include <utility>
#include <vector>

#include "ttm.h"

namespace tlib::ttm
{

template <class value_t>
class device_tensor;

// A device tensor bound to a contraction mode (device_tensor::operator()(q)).
template <class value_t>
struct device_tensor_view
{
    device_tensor<value_t> const &operand;
    std::size_t q;
};

template <class value_t>
class device_tensor
{
    static_assert(std::is_same_v<value_t, float> || std::is_same_v<value_t, double>,
                  "device_tensor holds float or double elements");

public:
    using shape_t = std::vector<std::size_t>;
    using layout_t = std::vector<std::size_t>;

    // Zero-initialised dense tensor of shape n in layout pi (as tensor(n, pi)).
    device_tensor(shape_t const &n, layout_t const &pi) : _n(n), _pi(pi)
    {
        if (n.size() != pi.size())
            throw std::runtime_error("Error in tlib::tensor: shape vector and layout vector must have the same length.");
        if (!detail::is_valid_shape(n.begin(), n.end()))
            throw std::runtime_error("Error in tlib::tensor: shape vector of tensor is not valid.");
        if (!detail::is_valid_layout(pi.begin(), pi.end()))
            throw std::runtime_error("Error in tlib::tensor: layout vector of tensor is not valid.");
        allocate();
        detail::throw_if_failed(ttm_device_zero(_data, bytes()));
    }

    explicit device_tensor(shape_t const &n) : device_tensor(n, detail::generate_k_order_layout(n.size(), std::size_t{1})) {}

    // Uninitialised storage for a result that a product overwrites completely.
    struct uninitialized_t
    {
    };
    device_tensor(shape_t const &n, layout_t const &pi, uninitialized_t) : _n(n), _pi(pi) { allocate(); }

    // Upload a host tensor (same shape, layout and values).
    explicit device_tensor(tensor<value_t> const &h) : _n(h.shape()), _pi(h.layout())
    {
        allocate();
        detail::throw_if_failed(ttm_device_copy(_data, h.data().data(), bytes()));
    }

    device_tensor(device_tensor const &o) : _n(o._n), _pi(o._pi)
    {
        allocate();
        detail::throw_if_failed(ttm_device_copy(_data, o._data, bytes()));
    }
    device_tensor(device_tensor &&o) noexcept : _n(std::move(o._n)), _pi(std::move(o._pi)), _data(o._data)
    {
        o._data = nullptr;
    }
    device_tensor &operator=(device_tensor o) noexcept
    {
        std::swap(_n, o._n);
        std::swap(_pi, o._pi);
        std::swap(_data, o._data);
        return *this;
    }
    ~device_tensor() { (void)ttm_device_free(_data); }

    // Download into a host tensor.
    tensor<value_t> to_host() const
    {
        tensor<value_t> h(_n, _pi);
        detail::throw_if_failed(ttm_device_copy(h.data().data(), _data, bytes()));
        return h;
    }

    device_tensor_view<value_t> operator()(std::size_t contraction_mode) const
    {
        if (1ul > contraction_mode || contraction_mode > this->order())
            throw std::runtime_error(
                "Error in tlib::tensor: specified contraction mode should be greater than one and equal to or less than the order.");
        return device_tensor_view<value_t>{*this, contraction_mode};
    }

    value_t *data() { return _data; }
    value_t const *data() const { return _data; }
    auto const &shape() const { return _n; }
    auto const &layout() const { return _pi; }
    auto strides() const { return detail::generate_strides(_n, _pi); }
    auto order() const { return _n.size(); }
    std::size_t size() const { return std::accumulate(_n.begin(), _n.end(), std::size_t{1}, std::multiplies<>()); }

private:
    std::size_t bytes() const { return size() * sizeof(value_t); }
    void allocate()
    {
        void *p = nullptr;
        detail::throw_if_failed(ttm_device_alloc(bytes(), &p));
        _data = static_cast<value_t *>(p);
    }

    shape_t _n;
    layout_t _pi;
    value_t *_data = nullptr;
};

/**
 * Tensor-object form on device tensors (the reference's include/tlib/ttm.h:100-125
 * with every operand in HBM): C in A's layout with extent nb[0] in mode q.
 * B's layout gives its orientation ({1,2} column-major, {2,1} row-major).
 */
template <class value_t>
inline device_tensor<value_t> ttm(std::size_t q, device_tensor<value_t> const &a, device_tensor<value_t> const &b)
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
    // every element of C is written by the product (no zero fill needed)
    device_tensor<value_t> c(nc, a.layout(), typename device_tensor<value_t>::uninitialized_t{});
    auto const wa = a.strides(), wc = c.strides();
    ttm_status_t st;
    if constexpr (std::is_same_v<value_t, float>)
        st = ttm_sttm(static_cast<unsigned>(q), static_cast<unsigned>(p), a.data(), a.shape().data(), wa.data(),
                      a.layout().data(), b.data(), b.shape().data(), b.layout().data(), c.data(), c.shape().data(),
                      wc.data(), nullptr);
    else
        st = ttm_dttm(static_cast<unsigned>(q), static_cast<unsigned>(p), a.data(), a.shape().data(), wa.data(),
                      a.layout().data(), b.data(), b.shape().data(), b.layout().data(), c.data(), c.shape().data(),
                      wc.data(), nullptr);
    detail::throw_if_failed(st);
    detail::throw_if_failed(ttm_device_synchronize());
    return c;
}

} // namespace tlib::ttm

/**
 * C = A(q) * B on device tensors (as include/tlib/ttm.h:133-138).
 */
template <class value_t>
inline auto operator*(tlib::ttm::device_tensor_view<value_t> const &a, tlib::ttm::device_tensor<value_t> const &b)
{
    return tlib::ttm::ttm(a.q, a.operand, b);
}

/**
 * Prints the values in the host tensor's format (downloads first).
 */
template <class value_t>
std::ostream &operator<<(std::ostream &out, tlib::ttm::device_tensor<value_t> const &a)
{
    return out << a.to_host();
}
