// device_tensor (include/tlib/device.h): the operator and tensor-object forms
// on HBM-resident tensors give bit-for-bit the host-tensor results (the same
// kernels run on the same values), a two-product chain stays on the device,
// the printer writes the host tensor's text, and the reference's error texts
// are kept.  Prints "OK <n>" and returns 0 when every check passes.
#include <tlib/device.h>

#include <algorithm>
#include <cstdio>
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
static tensor<T> filled(std::vector<std::size_t> const &n, std::vector<std::size_t> const &pi, unsigned seed)
{
    tensor<T> t(n, pi);
    unsigned s = seed;
    for (auto &x : t) {
        s = s * 1664525u + 1013904223u;
        x = T((s >> 8) % 2001) / T(1000) - T(1); // values in [-1, 1], exact in float
    }
    return t;
}

template <class T>
static void products()
{
    std::vector<std::size_t> const na{12, 9, 7}, pia{2, 3, 1};
    auto const A = filled<T>(na, pia, 1u);
    for (std::size_t q = 1; q <= 3; ++q) {
        for (auto const &pib : {std::vector<std::size_t>{1, 2}, std::vector<std::size_t>{2, 1}}) {
            auto const B = filled<T>({5, na[q - 1]}, pib, 7u + static_cast<unsigned>(q));
            auto const Ch = A(q) * B;                       // host tensors, staged per call
            device_tensor<T> const Ad(A), Bd(B);
            auto const Cd = Ad(q) * Bd;                     // device tensors
            auto const Cb = Cd.to_host();
            CHECK(Cb.shape() == Ch.shape());
            CHECK(Cb.layout() == Ch.layout());
            CHECK(std::equal(Cb.begin(), Cb.end(), Ch.begin()));
            auto const Ct = tlib::ttm::ttm(q, Ad, Bd);       // tensor-object form
            auto const Ctb = Ct.to_host();
            CHECK(std::equal(Ctb.begin(), Ctb.end(), Ch.begin()));
        }
    }
    // two products without leaving the device: (A x1 B1) x3 B3
    auto const B1 = filled<T>({4, 12}, {2, 1}, 21u);
    auto const B3 = filled<T>({3, 7}, {1, 2}, 22u);
    device_tensor<T> const Ad(A), B1d(B1), B3d(B3);
    auto const Dd = tlib::ttm::ttm(3, tlib::ttm::ttm(1, Ad, B1d), B3d);
    auto const Dh = (A(1) * B1)(3) * B3;
    auto const Db = Dd.to_host();
    CHECK(Db.shape() == (std::vector<std::size_t>{4, 9, 3}));
    CHECK(std::equal(Db.begin(), Db.end(), Dh.begin()));
    // printing matches the host tensor's text
    std::ostringstream od, oh;
    od << Dd;
    oh << Dh;
    CHECK(od.str() == oh.str());
}

static void errors()
{
    device_tensor<float> const A({4, 5, 6});
    device_tensor<float> const B({3, 5});
    try {
        (void)A(4);
        CHECK(false);
    } catch (std::runtime_error const &e) {
        CHECK(std::string(e.what()).find("contraction mode should be greater than one") != std::string::npos);
    }
    try {
        (void)(A(1) * B);  // n_q mismatch: 4 != 5
        CHECK(false);
    } catch (std::runtime_error const &e) {
        CHECK(std::string(e.what()) ==
              "Error in tlib::tensor_times_matrix: contraction dimension of A and B are not equal.");
    }
    device_tensor<float> const zero({2, 3}, {1, 2});
    auto const h = zero.to_host();
    CHECK(std::all_of(h.begin(), h.end(), [](float x) { return x == 0.f; }));
    device_tensor<float> copy = zero;
    CHECK(copy.shape() == zero.shape() && copy.data() != zero.data());
}

int main()
{
    try {
        products<float>();
        products<double>();
        errors();
    } catch (std::exception const &e) {
        std::fprintf(stderr, "unexpected exception: %s\n", e.what());
        return 2;
    }
    if (g_fail) {
        std::printf("FAILED %d of %d\n", g_fail, g_checks);
        return 1;
    }
    std::printf("OK %d\n", g_checks);
    return 0;
}
