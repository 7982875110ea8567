// Tests of the C++ drop-in layer (ompblas_dropin.hpp) in the shape of the
// reference's Catch2 suites: test/test.mtm.cpp (8 layout combinations,
// sizes 2..31), test/test.mtv.cpp (first / last order, sizes 2..511 and
// 32 Ki), test/test.dot_prod.cpp (1..s closed form), test/test.outer_prod.cpp
// and test/test.trans.cpp (sizes 2..31, every layout, in and out of place).  Inputs are
// rand() % 100 as in test/test_utils.hpp; the expected values come from
// integer brute force (exact while sums stay below 2^24) instead of BLIS,
// and from a relative bound for the 32 Ki case (the reference's Approx).
// Host buffers exercise the staging path, device buffers the pass-through.
//
// Usage: test_dropin [--quick]   (exit status 0 = all passed)
//        test_dropin --time-host   (host-buffer rates, PCIe staging included)
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "ompblas_dropin.hpp"

namespace {

int g_checks = 0, g_failures = 0;

#define CHECK(cond, ...)                                   \
    do {                                                   \
        ++g_checks;                                        \
        if (!(cond)) {                                     \
            ++g_failures;                                  \
            if (g_failures <= 20) {                        \
                std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
                std::printf(__VA_ARGS__);                  \
                std::printf("\n");                         \
            }                                              \
        }                                                  \
    } while (0)

// A matrix in one storage order with uBLAS-style extents and strides.
struct Mat {
    std::size_t n[2];
    std::size_t w[2];
    std::vector<float> v;
    Mat(std::size_t rows, std::size_t cols, bool first) : v(rows * cols) {
        n[0] = rows;
        n[1] = cols;
        w[0] = first ? 1 : cols;
        w[1] = first ? rows : 1;
    }
    float& at(std::size_t i, std::size_t j) { return v[i * w[0] + j * w[1]]; }
};

void rand_fill(std::vector<float>& v) {
    for (auto& x : v) x = static_cast<float>(std::rand() % 100);
}

// Device copy of a host vector (freed by the destructor).
struct Dev {
    float* p = nullptr;
    std::size_t n = 0;
    explicit Dev(const std::vector<float>& h) : n(h.size()) {
        void* d = nullptr;
        amt::dropin::check(omb_device_alloc(&d, n * sizeof(float)), "alloc");
        p = static_cast<float*>(d);
        amt::dropin::check(omb_memcpy_h2d(p, h.data(), n * sizeof(float), nullptr), "h2d");
        amt::dropin::check(omb_stream_synchronize(nullptr), "sync");
    }
    std::vector<float> get() const {
        std::vector<float> h(n);
        amt::dropin::check(omb_memcpy_d2h(h.data(), p, n * sizeof(float), nullptr), "d2h");
        amt::dropin::check(omb_stream_synchronize(nullptr), "sync");
        return h;
    }
    ~Dev() { (void)omb_device_free(p); }
};

template <class Tag>
void mtm_case(bool fc, bool fa, bool fb, std::size_t sz, bool device, const char* name) {
    Mat A(sz, sz, fa), B(sz, sz, fb), C(sz, sz, fc);
    rand_fill(A.v);
    rand_fill(B.v);
    rand_fill(C.v);  // C += A*B: a nonzero start checks the accumulation
    std::vector<int64_t> want(sz * sz);
    for (std::size_t i = 0; i < sz; ++i)
        for (std::size_t j = 0; j < sz; ++j) {
            int64_t s = static_cast<int64_t>(C.at(i, j));
            for (std::size_t k = 0; k < sz; ++k)
                s += static_cast<int64_t>(A.at(i, k)) * static_cast<int64_t>(B.at(k, j));
            want[i * sz + j] = s;
        }
    if (device) {
        Dev a(A.v), b(B.v), c(C.v);
        amt::mtm_helper(c.p, C.n, C.w, static_cast<const float*>(a.p), A.n, A.w,
                        static_cast<const float*>(b.p), B.n, B.w, Tag{});
        C.v = c.get();
    } else {
        amt::mtm_helper(C.v.data(), C.n, C.w, static_cast<const float*>(A.v.data()), A.n, A.w,
                        static_cast<const float*>(B.v.data()), B.n, B.w, Tag{});
    }
    for (std::size_t i = 0; i < sz; ++i)
        for (std::size_t j = 0; j < sz; ++j)
            CHECK(static_cast<int64_t>(C.at(i, j)) == want[i * sz + j],
                  "mtm %s sz=%zu %s (%zu,%zu): got %.1f want %lld", name, sz,
                  device ? "device" : "host", i, j, C.at(i, j), (long long)want[i * sz + j]);
}

void test_mtm(bool quick) {
    const char* names[8] = {"FFF", "FFL", "FLF", "FLL", "LFF", "LFL", "LLF", "LLL"};
    for (int l = 0; l < 8; ++l) {
        const bool fc = !(l & 4), fa = !(l & 2), fb = !(l & 1);
        for (std::size_t sz = 2; sz < 32; sz += quick ? 7 : 1) {
            if (fc)
                mtm_case<amt::layout::first_order>(fc, fa, fb, sz, false, names[l]);
            else
                mtm_case<amt::layout::last_order>(fc, fa, fb, sz, false, names[l]);
        }
        for (std::size_t sz : {2, 17, 31})
            mtm_case<amt::layout::last_order>(fc, fa, fb, sz, true, names[l]);
    }
}

// first order: c += A v (A column-major); last order: c = A v (row-major).
void mtv_case(bool first, std::size_t sz, bool device) {
    Mat A(sz, sz, first);
    rand_fill(A.v);
    std::vector<float> v(sz), c(sz);
    rand_fill(v);
    rand_fill(c);
    std::vector<int64_t> want(sz);
    for (std::size_t i = 0; i < sz; ++i) {
        int64_t s = first ? static_cast<int64_t>(c[i]) : 0;
        for (std::size_t k = 0; k < sz; ++k)
            s += static_cast<int64_t>(A.at(i, k)) * static_cast<int64_t>(v[k]);
        want[i] = s;
    }
    const std::size_t nv[2] = {1, sz}, nc[2] = {1, sz};
    auto call = [&](float* cp, const float* ap, const float* vp) {
        if (first)
            amt::mtv_helper(cp, nc, ap, A.n, A.w, vp, nv, 8, amt::layout::first_order{});
        else
            amt::mtv_helper(cp, nc, ap, A.n, A.w, vp, nv, 8, amt::layout::last_order{});
    };
    if (device) {
        Dev a(A.v), dv(v), dc(c);
        call(dc.p, a.p, dv.p);
        c = dc.get();
    } else {
        call(c.data(), A.v.data(), v.data());
    }
    for (std::size_t i = 0; i < sz; ++i)
        CHECK(static_cast<int64_t>(c[i]) == want[i], "mtv %s sz=%zu %s row %zu: got %.1f want %lld",
              first ? "first" : "last", sz, device ? "device" : "host", i, c[i],
              (long long)want[i]);
}

void test_mtv(bool quick) {
    for (int first = 0; first < 2; ++first) {
        for (std::size_t sz = 2; sz < 512; sz += quick ? 37 : 1) mtv_case(first, sz, false);
        for (std::size_t sz : {2, 255, 511}) mtv_case(first, sz, true);
    }
}

// test/test.mtv.cpp:128-211: 32 Ki x 32 Ki (4 GiB), compared with a relative
// bound like Catch2's Approx (the sums exceed 2^24).
void test_mtv_large(bool first) {
    const std::size_t sz = 32 * 1024;
    Mat A(sz, sz, first);
    rand_fill(A.v);
    std::vector<float> v(sz), c(sz, 0.f);
    rand_fill(v);
    std::vector<double> want(sz, 0.0);
    if (first) {  // walk A in storage order
        for (std::size_t k = 0; k < sz; ++k)
            for (std::size_t i = 0; i < sz; ++i) want[i] += static_cast<double>(A.v[i + k * sz]) * v[k];
    } else {
        for (std::size_t i = 0; i < sz; ++i)
            for (std::size_t k = 0; k < sz; ++k) want[i] += static_cast<double>(A.v[i * sz + k]) * v[k];
    }
    Dev a(A.v), dv(v), dc(c);
    A.v.clear();
    A.v.shrink_to_fit();
    const std::size_t nv[2] = {1, sz}, nc[2] = {1, sz};
    if (first)
        amt::mtv_helper(dc.p, nc, static_cast<const float*>(a.p), A.n, A.w,
                        static_cast<const float*>(dv.p), nv, 8, amt::layout::first_order{});
    else
        amt::mtv_helper(dc.p, nc, static_cast<const float*>(a.p), A.n, A.w,
                        static_cast<const float*>(dv.p), nv, 8, amt::layout::last_order{});
    c = dc.get();
    const double eps = 100.0 * 1.1920929e-07;  // Catch2 Approx default
    for (std::size_t i = 0; i < sz; ++i)
        CHECK(std::fabs(c[i] - want[i]) <= eps * (1.0 + std::fabs(want[i])),
              "mtv large %s row %zu: got %.1f want %.1f", first ? "first" : "last", i, c[i], want[i]);
}

// test/test.dot_prod.cpp:10-28: a = 1..s, a.a = s(s+1)(2s+1)/6.
void test_dot(bool quick) {
    const std::size_t smax = (1u << 17) + 1;
    std::vector<float> full(smax);
    for (std::size_t i = 0; i < smax; ++i) full[i] = static_cast<float>(i + 1);
    Dev dfull(full);
    float* r_dev = nullptr;
    {
        void* p = nullptr;
        amt::dropin::check(omb_device_alloc(&p, sizeof(float)), "alloc");
        r_dev = static_cast<float*>(p);
    }
    const double u = std::ldexp(1.0, -24);
    const std::size_t step = quick ? 4099 : 97;
    for (std::size_t s = 2; s <= smax; s += (s < 300 ? 1 : step)) {
        const std::size_t n[2] = {1, s};
        float r = 0.f;
        const bool device = (s % 3 == 0);
        if (device) {
            amt::dot_prod_helper(r_dev, static_cast<const float*>(dfull.p), n,
                                 static_cast<const float*>(dfull.p), n, 8);
            amt::dropin::check(omb_memcpy_d2h(&r, r_dev, sizeof(float), nullptr), "d2h");
            amt::dropin::check(omb_stream_synchronize(nullptr), "sync");
        } else {
            amt::dot_prod_helper(&r, static_cast<const float*>(full.data()), n,
                                 static_cast<const float*>(full.data()), n, 8);
        }
        const double exact = static_cast<double>(s) * (s + 1) * (2 * s + 1) / 6.0;
        if (exact < 16777216.0)
            CHECK(r == exact, "dot s=%zu: got %.1f want %.1f", s, r, exact);
        else
            CHECK(std::fabs(r - exact) <= (s * u / (1 - s * u)) * exact, "dot s=%zu: got %.6e want %.6e",
                  s, r, exact);
    }
    (void)omb_device_free(r_dev);
}

// test/test.outer_prod.cpp: C += a b^T through the closure's helper calls
// (include/outer.hpp:104-120): first-order C as columns, last-order C as
// rows with a and b swapped.
void outer_case(bool first, std::size_t m, std::size_t n, bool device) {
    Mat C(m, n, first);
    std::vector<float> a(m), b(n);
    rand_fill(a);
    rand_fill(b);
    rand_fill(C.v);
    std::vector<int64_t> want(m * n);
    for (std::size_t i = 0; i < m; ++i)
        for (std::size_t j = 0; j < n; ++j)
            want[i * n + j] = static_cast<int64_t>(C.at(i, j)) +
                              static_cast<int64_t>(a[i]) * static_cast<int64_t>(b[j]);
    const std::size_t na[2] = {1, m}, nb[2] = {1, n};
    const std::size_t wc_first[2] = {1, m}, wc_last[2] = {n, 1};
    auto call = [&](float* cp, const float* ap, const float* bp) {
        if (first)
            amt::outer_prod_helper(cp, C.n, wc_first, ap, na, bp, nb, 8);
        else
            amt::outer_prod_helper(cp, C.n, wc_last, bp, nb, ap, na, 8);
    };
    if (device) {
        Dev dc(C.v), da(a), db(b);
        call(dc.p, da.p, db.p);
        C.v = dc.get();
    } else {
        call(C.v.data(), a.data(), b.data());
    }
    for (std::size_t i = 0; i < m; ++i)
        for (std::size_t j = 0; j < n; ++j)
            CHECK(static_cast<int64_t>(C.at(i, j)) == want[i * n + j],
                  "outer %s %zux%zu %s (%zu,%zu)", first ? "first" : "last", m, n,
                  device ? "device" : "host", i, j);
}

// test/test.trans.cpp: out of place for every layout pair, and in place.
void transpose_case(bool fa, bool fc, std::size_t m, std::size_t n, bool device) {
    Mat A(m, n, fa), C(n, m, fc);
    rand_fill(A.v);
    std::fill(C.v.begin(), C.v.end(), -1.f);
    if (device) {
        Dev da(A.v), dc(C.v);
        amt::transpose_helper(dc.p, C.n, C.w, static_cast<const float*>(da.p), A.n, A.w,
                              amt::tag::outplace{});
        C.v = dc.get();
    } else {
        amt::transpose_helper(C.v.data(), C.n, C.w, static_cast<const float*>(A.v.data()), A.n,
                              A.w, amt::tag::outplace{});
    }
    for (std::size_t i = 0; i < m; ++i)
        for (std::size_t j = 0; j < n; ++j)
            CHECK(C.at(j, i) == A.at(i, j), "transpose %c%c %zux%zu %s (%zu,%zu)", fa ? 'F' : 'L',
                  fc ? 'F' : 'L', m, n, device ? "device" : "host", i, j);
}

void transpose_inplace_case(std::size_t n, bool device) {
    Mat A(n, n, true);
    rand_fill(A.v);
    const std::vector<float> orig = A.v;
    const std::size_t wc[2] = {1, n}, wa[2] = {1, n};
    if (device) {
        Dev da(A.v);
        amt::transpose_helper(da.p, A.n, wc, da.p, A.n, wa, amt::tag::inplace{});
        A.v = da.get();
    } else {
        amt::transpose_helper(A.v.data(), A.n, wc, A.v.data(), A.n, wa, amt::tag::inplace{});
    }
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = 0; j < n; ++j)
            CHECK(A.v[i + j * n] == orig[j + i * n], "transpose in place n=%zu %s (%zu,%zu)", n,
                  device ? "device" : "host", i, j);
}

void test_outer_transpose(bool quick) {
    for (int first = 0; first < 2; ++first) {
        for (std::size_t sz = 2; sz < 32; sz += quick ? 7 : 1) outer_case(first, sz, sz, false);
        outer_case(first, 3, 70, false);
        outer_case(first, 129, 5, true);
    }
    for (int l = 0; l < 4; ++l) {
        for (std::size_t sz = 2; sz < 32; sz += quick ? 7 : 1) transpose_case(l & 1, l & 2, sz, sz, false);
        transpose_case(l & 1, l & 2, 70, 3, false);
        transpose_case(l & 1, l & 2, 129, 200, true);
    }
    for (std::size_t sz = 2; sz < 32; sz += quick ? 7 : 1) transpose_inplace_case(sz, false);
    transpose_inplace_case(65, true);
    transpose_inplace_case(300, false);
}

// Tensor-level closures over amt::matrix_view (host memory: the closures run
// the helpers, which stage it): shapes the reference's closure tests use,
// rectangular ones, and the shape errors raised when a closure is made.
void test_tensor_level() {
    using V = amt::matrix_view<float>;
    using CV = amt::matrix_view<const float>;
    for (std::size_t m : {3, 17, 40}) {
        const std::size_t k = m + 5, n = 2 * m + 1;
        Mat A(m, k, true), B(k, n, false), C(m, n, false);
        rand_fill(A.v);
        rand_fill(B.v);
        std::vector<int64_t> want(m * n, 0);
        for (std::size_t i = 0; i < m; ++i)
            for (std::size_t j = 0; j < n; ++j)
                for (std::size_t q = 0; q < k; ++q)
                    want[i * n + j] += static_cast<int64_t>(A.at(i, q)) * static_cast<int64_t>(B.at(q, j));
        auto cv = V::last_order(C.v.data(), m, n);
        amt::mtm(cv, CV::first_order(A.v.data(), m, k), CV::last_order(B.v.data(), k, n))();
        for (std::size_t i = 0; i < m; ++i)
            for (std::size_t j = 0; j < n; ++j)
                CHECK(static_cast<int64_t>(C.at(i, j)) == want[i * n + j], "tensor mtm %zu (%zu,%zu)", m, i, j);
        // mtv (first order accumulates), vtm (true transpose)
        std::vector<float> x(k), y(m, 0.f), z(k, 0.f), u(m);
        rand_fill(x);
        rand_fill(u);
        auto yv = V::last_order(y.data(), 1, m);
        amt::mtv(yv, CV::first_order(A.v.data(), m, k), CV::last_order(x.data(), 1, k))();
        auto zv = V::last_order(z.data(), 1, k);
        amt::vtm(zv, CV::first_order(A.v.data(), m, k), CV::last_order(u.data(), 1, m))();
        for (std::size_t i = 0; i < m; ++i) {
            int64_t s = 0;
            for (std::size_t q = 0; q < k; ++q) s += static_cast<int64_t>(A.at(i, q)) * static_cast<int64_t>(x[q]);
            CHECK(static_cast<int64_t>(y[i]) == s, "tensor mtv %zu row %zu", m, i);
        }
        for (std::size_t q = 0; q < k; ++q) {
            int64_t s = 0;
            for (std::size_t i = 0; i < m; ++i) s += static_cast<int64_t>(A.at(i, q)) * static_cast<int64_t>(u[i]);
            CHECK(static_cast<int64_t>(z[q]) == s, "tensor vtm %zu col %zu", m, q);
        }
        float r = 0.f;
        amt::matrix_view<float> rv{&r, {1, 1}, {1, 1}};
        amt::dot_prod(rv, CV::last_order(x.data(), 1, k), CV::last_order(x.data(), 1, k))();
        int64_t dd = 0;
        for (float t : x) dd += static_cast<int64_t>(t) * static_cast<int64_t>(t);
        CHECK(static_cast<int64_t>(r) == dd, "tensor dot %zu", m);
        // the tuning entries of src/dot_tuning.cpp: blocks accepted, same result
        float rt = -1.f;
        amt::dot_prod_tuning_param(rt, CV::last_order(x.data(), 1, k),
                                   CV::last_order(x.data(), 1, k), 1024, 4096, 65536, 8)();
        CHECK(rt == r, "tensor dot tuning %zu", m);
        float rh = -1.f;
        amt::dot_prod_helper_tuning_param(&rh, x.data(), x.data(), k, 8, std::size_t(64),
                                          std::size_t(128), std::size_t(256));
        CHECK(rh == r, "dot helper tuning %zu", m);
        // outer product into both orders, transpose out of and in place
        for (int first = 0; first < 2; ++first) {
            Mat O(m, k, first);
            std::fill(O.v.begin(), O.v.end(), 1.f);
            auto ov = first ? V::first_order(O.v.data(), m, k) : V::last_order(O.v.data(), m, k);
            amt::outer_prod(ov, CV::last_order(u.data(), 1, m), CV::last_order(x.data(), 1, k))();
            for (std::size_t i = 0; i < m; ++i)
                for (std::size_t q = 0; q < k; ++q)
                    CHECK(O.at(i, q) == 1.f + u[i] * x[q], "tensor outer %d %zu (%zu,%zu)", first, m, i, q);
        }
        Mat T(k, m, false);
        auto tv = V::last_order(T.v.data(), k, m);
        amt::transpose(tv, CV::first_order(A.v.data(), m, k))();
        for (std::size_t i = 0; i < m; ++i)
            for (std::size_t q = 0; q < k; ++q) CHECK(T.at(q, i) == A.at(i, q), "tensor transpose %zu", m);
        Mat S(m, m, true);
        rand_fill(S.v);
        const std::vector<float> s0 = S.v;
        auto sv = V::first_order(S.v.data(), m, m);
        amt::transpose(sv)();
        for (std::size_t i = 0; i < m; ++i)
            for (std::size_t j = 0; j < m; ++j) CHECK(S.v[i + j * m] == s0[j + i * m], "tensor transpose in place %zu", m);
    }
    // shape errors at closure creation
    std::vector<float> buf(64);
    auto expect_throw = [&](auto&& make, const char* what) {
        bool threw = false;
        try {
            make();
        } catch (const std::runtime_error& e) {
            threw = std::strstr(e.what(), what) != nullptr;
        }
        CHECK(threw, "expected '%s'", what);
    };
    auto m34 = V::last_order(buf.data(), 3, 4);
    auto m52 = CV::last_order(buf.data(), 5, 2);
    auto m32 = V::last_order(buf.data(), 3, 2);
    expect_throw([&] { amt::mtm(m32, CV::last_order(buf.data(), 3, 4), m52); }, "dimension mismatch");
    auto v3 = V::last_order(buf.data(), 1, 3);
    expect_throw([&] { amt::mtv(v3, CV::last_order(buf.data(), 3, 4), CV::last_order(buf.data(), 1, 5)); },
                 "dimension mismatch");
    expect_throw([&] { amt::mtv(m34, CV::last_order(buf.data(), 3, 4), CV::last_order(buf.data(), 1, 4)); },
                 "must be vector");
    expect_throw([&] { amt::dot_prod(v3, CV::last_order(buf.data(), 1, 4), CV::last_order(buf.data(), 1, 5)); },
                 "same size");
    expect_throw([&] { amt::outer_prod(m34, CV::last_order(buf.data(), 1, 3), CV::last_order(buf.data(), 1, 5)); },
                 "dimension mismatch");
    expect_throw([&] { amt::transpose(m34); }, "square");
}

void test_errors() {
    // only float is computed
    double c = 0, a[4] = {1, 2, 3, 4};
    const std::size_t n[2] = {2, 2}, w[2] = {2, 1};
    bool threw = false;
    try {
        amt::mtm_helper(&c, n, w, static_cast<const double*>(a), n, w, static_cast<const double*>(a),
                        n, w, amt::layout::last_order{});
    } catch (const std::runtime_error& e) {
        threw = std::strstr(e.what(), "only float") != nullptr;
    }
    CHECK(threw, "double mtm_helper must throw");
    // strides where neither is 1 are refused with the library's message
    std::vector<float> A(64, 1.f), B(64, 1.f), C(64, 0.f);
    const std::size_t na[2] = {3, 3}, wbad[2] = {6, 2}, wok[2] = {3, 1};
    threw = false;
    try {
        amt::mtm_helper(C.data(), na, wok, static_cast<const float*>(A.data()), na, wbad,
                        static_cast<const float*>(B.data()), na, wok, amt::layout::last_order{});
    } catch (const std::runtime_error& e) {
        threw = std::strstr(e.what(), "amt::mtm_helper") != nullptr;
    }
    CHECK(threw, "general strides must throw from mtm_helper");
}

}  // namespace

// Wall time of the drop-in helpers on host buffers: each call stages its
// operands to the device and the result back, as a caller of the reference
// helpers with host memory would see it.  Mean of 3 calls after one.
template <class F>
double host_call_seconds(F&& f) {
    f();
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < 3; ++i) f();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / 3;
}

// Host operands are staged through the drop-in layer's buffer pool: the
// second call of the same shapes reuses the cached blocks (the cache does
// not grow), results stay exact, release_staging() empties the cache, and
// keep_bytes = 0 frees every block on return.
void test_staging_pool() {
    namespace lay = amt::layout;
    auto& pool = amt::dropin::staging_pool::get();
    amt::dropin::release_staging();
    CHECK(pool.cached_bytes() == 0, "pool not empty after release_staging");
    const std::size_t m = 67, k = 45, n = 53;
    std::vector<float> a(m * k), b(k * n), c(m * n), want(m * n, 0.f);
    for (auto& v : a) v = (float)(std::rand() % 100);
    for (auto& v : b) v = (float)(std::rand() % 100);
    for (std::size_t i = 0; i < m; ++i)
        for (std::size_t j = 0; j < n; ++j) {
            double t = 0.0;
            for (std::size_t q = 0; q < k; ++q) t += (double)a[i * k + q] * b[q * n + j];
            want[i * n + j] = (float)t;
        }
    const std::size_t na[2] = {m, k}, wa[2] = {k, 1}, nb[2] = {k, n}, wb[2] = {n, 1},
                      nc[2] = {m, n}, wc[2] = {n, 1};
    std::size_t cached[2] = {0, 0};
    for (int call = 0; call < 2; ++call) {
        std::fill(c.begin(), c.end(), 0.f);
        amt::mtm_helper(c.data(), nc, wc, a.data(), na, wa, b.data(), nb, wb, lay::last_order{});
        CHECK(c == want, "staged mtm call %d: wrong result", call);
        cached[call] = pool.cached_bytes();
    }
    CHECK(cached[0] > 0, "no staging block cached after a host call");
    CHECK(cached[1] == cached[0], "second call grew the cache (%zu -> %zu)", cached[0], cached[1]);
    amt::dropin::release_staging();
    CHECK(pool.cached_bytes() == 0, "pool not empty after release_staging");
    const std::size_t keep = pool.keep_bytes;
    pool.keep_bytes = 0;
    std::fill(c.begin(), c.end(), 0.f);
    amt::mtm_helper(c.data(), nc, wc, a.data(), na, wa, b.data(), nb, wb, lay::last_order{});
    CHECK(c == want, "staged mtm with keep_bytes = 0: wrong result");
    CHECK(pool.cached_bytes() == 0, "keep_bytes = 0 still cached %zu bytes", pool.cached_bytes());
    pool.keep_bytes = keep;
}

int time_host() {
    namespace lay = amt::layout;
    const std::size_t n = 8192;
    std::vector<float> a(n * n), b(n * n), c(n * n);
    for (std::size_t i = 0; i < n * n; ++i) {
        a[i] = (float)(std::rand() % 100) / 100.f;
        b[i] = (float)(std::rand() % 100) / 100.f;
    }
    const std::size_t na[2] = {n, n}, wrow[2] = {n, 1};
    for (int mode : {OMB_MODE_PLAIN_F32, OMB_MODE_SPLIT_BF16, OMB_MODE_SPLIT_FP16}) {
        amt::dropin::settings().gemm_mode = mode;
        const double t = host_call_seconds([&] {
            amt::mtm_helper(c.data(), na, wrow, a.data(), na, wrow, b.data(), na, wrow,
                            lay::last_order{});
        });
        std::printf("host mtm 8192^3 %s: %.3f s, %.2f TFLOP/s (768 MiB to the device, 256 MiB back)\n",
                    mode == OMB_MODE_PLAIN_F32   ? "plain"
                    : mode == OMB_MODE_SPLIT_BF16 ? "split-bf16"
                                                  : "split-fp16",
                    t, 2.0 * n * n * n / t / 1e12);
    }
    amt::dropin::settings().gemm_mode = OMB_MODE_PLAIN_F32;
    const std::size_t m = 16384;
    std::vector<float> ga(m * m, 0.5f), x(m, 0.25f), y(m, 0.f);
    const std::size_t nga[2] = {m, m}, wga[2] = {m, 1}, nx[2] = {m, 1}, ny[2] = {m, 1};
    const double tv = host_call_seconds([&] {
        amt::mtv_helper(y.data(), ny, ga.data(), nga, wga, x.data(), nx, 0, lay::last_order{});
    });
    std::printf("host mtv 16384^2 row-major: %.4f s, %.1f GB/s of A\n", tv, 4.0 * m * m / tv / 1e9);
    const std::size_t nd = std::size_t(1) << 28;
    std::vector<float> dx(nd, 0.5f), dy(nd, 0.25f);
    float r = 0.f;
    const std::size_t ndx[2] = {nd, 1};
    const double td = host_call_seconds(
        [&] { amt::dot_prod_helper(&r, dx.data(), ndx, dy.data(), ndx, 0); });
    std::printf("host dot 2^28: %.4f s, %.1f GB/s of x and y\n", td, 8.0 * nd / td / 1e9);
    std::fflush(stdout);
    return 0;
}

int main(int argc, char** argv) {
    const bool quick = argc > 1 && std::strcmp(argv[1], "--quick") == 0;
    if (argc > 1 && std::strcmp(argv[1], "--time-host") == 0) {
        try {
            amt::dropin::check(omb_init(), "omb_init");
            return time_host();
        } catch (const std::exception& e) {
            std::printf("FAIL: exception: %s\n", e.what());
            return 2;
        }
    }
    std::srand(0);
    try {
        amt::dropin::check(omb_init(), "omb_init");
        test_mtm(quick);
        std::printf("mtm: %d checks, %d failures\n", g_checks, g_failures);
        test_mtv(quick);
        std::printf("mtv: %d checks, %d failures\n", g_checks, g_failures);
        test_dot(quick);
        std::printf("dot: %d checks, %d failures\n", g_checks, g_failures);
        if (!quick) {
            test_mtv_large(true);
            test_mtv_large(false);
            std::printf("mtv 32Ki: %d checks, %d failures\n", g_checks, g_failures);
        }
        test_outer_transpose(quick);
        std::printf("outer/transpose: %d checks, %d failures\n", g_checks, g_failures);
        test_tensor_level();
        std::printf("tensor-level closures: %d checks, %d failures\n", g_checks, g_failures);
        test_errors();
        test_staging_pool();
        std::printf("staging pool: %d checks, %d failures\n", g_checks, g_failures);
        amt::dropin::settings().gemm_mode = OMB_MODE_SPLIT_BF16;
        for (std::size_t sz : {5, 31}) mtm_case<amt::layout::last_order>(false, true, false, sz, false, "LFL split");
        amt::dropin::settings().gemm_mode = OMB_MODE_SPLIT_FP16;
        for (std::size_t sz : {5, 31}) mtm_case<amt::layout::last_order>(false, true, false, sz, false, "LFL split-fp16");
        amt::dropin::settings().gemm_mode = OMB_MODE_PLAIN_F32;
    } catch (const std::exception& e) {
        std::printf("FAIL: exception: %s\n", e.what());
        return 2;
    }
    std::printf("%s: %d checks, %d failures\n", g_failures ? "FAILED" : "PASSED", g_checks,
                g_failures);
    std::fflush(stdout);
    return g_failures ? 1 : 0;
}
