// Exercises the three call forms of the drop-in header (include/tlib/ttm.h)
// the way the reference's examples use them (example/interface{1,2,3}.cpp),
// checks the results against the closed form of the reference's tests
// (A numbered along mode q, B = ones -> T(last) - T(first - 1)), and checks
// that invalid arguments throw std::runtime_error with the reference's text.
// Prints "OK <n>" and returns 0 when every check passes.
#include <tlib/ttm.h>

#include <cmath>
#include <cstdio>
#include <iostream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

using namespace tlib::ttm;

static int g_fail = 0, g_checks = 0;
#define CHECK(cond)                                                                   \
    do {                                                                              \
        ++g_checks;                                                                   \
        if (!(cond)) {                                                                \
            ++g_fail;                                                                 \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);     \
        }                                                                             \
    } while (0)

template <class T>
static void operator_form()
{
    // example/interface1.cpp: A 4x3x2 first-order, iota from 1, B (n+1) x n ones, q = 1
    auto na = std::vector<std::size_t>{4, 3, 2};
    auto pia = detail::generate_k_order_layout(na.size(), std::size_t{1});
    auto nb = std::vector<std::size_t>{na[0] + 1, na[0]};
    auto pib = detail::generate_k_order_layout(nb.size(), std::size_t{1});
    auto A = tensor<T>(na, pia);
    auto B = tensor<T>(nb, pib);
    std::iota(A.begin(), A.end(), T(1));
    std::fill(B.begin(), B.end(), T(1));
    auto C = A(1) * B;
    CHECK(C.shape() == (std::vector<std::size_t>{5, 3, 2}));
    CHECK(C.layout() == pia);
    // column (i2, i3) of A is 4*(i2 + 3 i3) + 1 .. + 4
    for (std::size_t i3 = 0; i3 < 2; ++i3)
        for (std::size_t i2 = 0; i2 < 3; ++i2) {
            T const first = T(4 * (i2 + 3 * i3) + 1);
            T const expect = 4 * first + 6;
            for (std::size_t j = 0; j < 5; ++j) CHECK(C[j + 5 * (i2 + 3 * i3)] == expect);
        }
    std::ostringstream os;
    os << C;
    CHECK(os.str().find("cat(3,...") == 0);
}

template <class T>
static void tensor_form()
{
    // example/interface2.cpp style: every q and layout of an order-3 tensor,
    // B row- and column-major, policies of every kind.
    auto na = std::vector<std::size_t>{3, 4, 2};
    for (std::size_t k = 0; k <= 3; ++k) {
        auto pia = detail::generate_k_order_layout(na.size(), k);
        for (std::size_t q = 1; q <= 3; ++q) {
            auto A = tensor<T>(na, pia);
            auto wa = A.strides();
            // number A along mode q (fibres consecutive), as the reference tests do
            std::size_t nfib = 0;
            std::vector<std::size_t> idx(3, 0);
            for (std::size_t t = 0; t < 24; ++t) {
                std::size_t rem = t, off = 0;
                for (std::size_t r = 0; r < 3; ++r) {
                    idx[r] = rem % na[r];
                    rem /= na[r];
                    off += idx[r] * wa[r];
                }
                // fibre id over the non-q modes (natural order), l = idx[q-1]
                std::size_t fib = 0, rad = 1;
                for (std::size_t r = 0; r < 3; ++r)
                    if (r != q - 1) {
                        fib += idx[r] * rad;
                        rad *= na[r];
                    }
                nfib = rad;
                A[off] = T(fib * na[q - 1] + idx[q - 1] + 1);
            }
            (void)nfib;
            for (int layout_b = 0; layout_b < 2; ++layout_b) {
                auto nb = std::vector<std::size_t>{2 * na[q - 1], na[q - 1]};
                auto pib = layout_b ? std::vector<std::size_t>{2, 1} : std::vector<std::size_t>{1, 2};
                auto B = tensor<T>(nb, pib);
                B = T(1);
                auto C1 = ttm(q, A, B, parallel_policy::parallel_loop, slicing_policy::subtensor, fusion_policy::all);
                auto C2 = ttm(q, A, B, parallel_policy::sequential, slicing_policy::slice, fusion_policy::none);
                auto const wc = C1.strides();
                auto const &nc = C1.shape();
                CHECK(C1.data() == C2.data());
                for (std::size_t t = 0; t < C1.data().size(); ++t) {
                    std::size_t rem = t, off = 0, fib = 0, rad = 1;
                    for (std::size_t r = 0; r < 3; ++r) {
                        idx[r] = rem % nc[r];
                        rem /= nc[r];
                        off += idx[r] * wc[r];
                        if (r != q - 1) {
                            fib += idx[r] * rad;
                            rad *= nc[r];
                        }
                    }
                    double const first = double(fib * na[q - 1] + 1), last = first + double(na[q - 1]) - 1;
                    double const expect = last * (last + 1) / 2 - (first - 1) * first / 2;
                    CHECK(double(C1[off]) == expect);
                }
            }
        }
    }
}

template <class T>
static void lowlevel_form()
{
    // example/interface3.cpp: raw std::vector buffers, q = 2, B column-major ones
    using shape_t = std::vector<std::size_t>;
    auto na = shape_t{4, 3, 2};
    auto p = na.size();
    auto q = 2ul;
    auto pia = detail::generate_k_order_layout(p, 1ul);
    auto wa = detail::generate_strides(na, pia);
    auto pib = shape_t{1, 2};
    auto nb = shape_t{na[q - 1] + 1, na[q - 1]};
    auto nc = na;
    nc[q - 1] = nb[0];
    auto wc = detail::generate_strides(nc, pia);
    auto A = std::vector<T>(24);
    auto B = std::vector<T>(nb[0] * nb[1], T(1));
    auto C = std::vector<T>(4 * 4 * 2, T(-7));
    std::iota(A.begin(), A.end(), T(1));
    ttm(parallel_policy::parallel_blas, slicing_policy::slice, fusion_policy::none, unsigned(q), unsigned(p),
        A.data(), na.data(), wa.data(), pia.data(), B.data(), nb.data(), pib.data(), C.data(), nc.data(), wc.data());
    for (std::size_t i3 = 0; i3 < 2; ++i3)
        for (std::size_t j = 0; j < 4; ++j)
            for (std::size_t i1 = 0; i1 < 4; ++i1) {
                // A(i1, :, i3) = i1 + 1 + 4*l + 12*i3, l = 0..2
                T const expect = T(3 * (i1 + 1 + 12 * i3) + 12);
                CHECK(C[i1 + 4 * j + 16 * i3] == expect);
            }

    // a size type other than std::size_t
    std::vector<unsigned> na_u{4, 3, 2}, wa_u{1, 4, 12}, pia_u{1, 2, 3}, nb_u{4, 3}, pib_u{1, 2}, nc_u{4, 4, 2},
        wc_u{1, 4, 16};
    auto C2 = std::vector<T>(32, T(0));
    ttm(parallel_policy::combined, slicing_policy::combined, fusion_policy::all, 2u, 3u, A.data(), na_u.data(),
        wa_u.data(), pia_u.data(), B.data(), nb_u.data(), pib_u.data(), C2.data(), nc_u.data(), wc_u.data());
    CHECK(C2 == C);

    // reference error texts (include/tlib/ttm.h:65-90)
    auto expect_throw = [&](auto &&fn, const char *text) {
        try {
            fn();
            CHECK(false);
        } catch (std::runtime_error const &e) {
            CHECK(std::string(e.what()) == text);
        }
    };
    expect_throw([&] {
        ttm(parallel_policy::combined, slicing_policy::combined, fusion_policy::all, 0u, unsigned(p), A.data(),
            na.data(), wa.data(), pia.data(), B.data(), nb.data(), pib.data(), C.data(), nc.data(), wc.data());
    }, "Error in tlib::tensor_times_matrix: contraction mode should be greater zero or less than or equal to p.");
    expect_throw([&] {
        ttm(parallel_policy::combined, slicing_policy::combined, fusion_policy::all, unsigned(q), unsigned(p),
            (T *)nullptr, na.data(), wa.data(), pia.data(), B.data(), nb.data(), pib.data(), C.data(), nc.data(),
            wc.data());
    }, "Error in tlib::tensor_times_matrix: pointer to input tensor A should not be zero.");
    auto bad_nb = shape_t{4, 5};
    expect_throw([&] {
        ttm(parallel_policy::combined, slicing_policy::combined, fusion_policy::all, unsigned(q), unsigned(p),
            A.data(), na.data(), wa.data(), pia.data(), B.data(), bad_nb.data(), pib.data(), C.data(), nc.data(),
            wc.data());
    }, "Error in tlib::tensor_times_matrix: contraction dimension of A and B are not equal.");
    auto bad_wa = shape_t{12, 4, 1};
    expect_throw([&] {
        ttm(parallel_policy::combined, slicing_policy::combined, fusion_policy::all, unsigned(q), unsigned(p),
            A.data(), na.data(), bad_wa.data(), pia.data(), B.data(), nb.data(), pib.data(), C.data(), nc.data(),
            wc.data());
    }, "Error in tlib::tensor_times_matrix: stride vector of A is not valid.");
}

int main()
{
    operator_form<float>();
    operator_form<double>();
    tensor_form<float>();
    tensor_form<double>();
    lowlevel_form<float>();
    lowlevel_form<double>();
    if (g_fail) {
        std::printf("FAILED %d of %d\n", g_fail, g_checks);
        return 1;
    }
    std::printf("OK %d\n", g_checks);
    return 0;
}
