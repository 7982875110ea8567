// C++ drop-in for the reference's pointer-level BLAS helpers.
//
//   amt::mtm_helper      include/mtm.hpp:116-122   C += A*B
//   amt::mtv_helper      include/mtv.hpp:15-22     first_order (column-major A): c += A*b
//                        include/mtv.hpp:72-79     last_order  (row-major A):    c  = A*b
//   amt::dot_prod_helper include/dot.hpp:143-149   *c = sum a[i]*b[i]
//   amt::outer_prod_helper include/outer.hpp:15-48 C lines += a * b[i]
//   amt::transpose_helper include/trans.hpp:23-71  C = A^T (out of place / square in place)
//
// Same template signatures, argument meanings (n* = extents {rows, cols},
// w* = strides {row stride, column stride}) and accumulate / overwrite
// semantics as the reference; the work runs on the gfx950 kernels behind the
// C ABI in ompblas.h.  Pointers may be host or device memory: device
// pointers pass straight through, host buffers are staged to the device
// (outputs are copied back).  Every call returns with its result in place,
// like the reference's synchronous OpenMP calls.
//
// Deviations, by necessity: failures throw std::runtime_error carrying the
// library's message (the reference's mtv/dot helpers are noexcept and cannot
// fail); `max_threads` is accepted and ignored (the GPU has no thread team to
// size); only `float` is computed -- other value types throw.
//
// The reference's layout tags are boost::numeric::ublas::layout::first_order
// and last_order.  This header defines stand-ins with the same names under
// amt::layout; code that keeps using Boost's tags adds one line per tag:
//   template <> struct amt::layout_traits<boost::numeric::ublas::layout::first_order>
//       : amt::layout_traits<amt::layout::first_order> {};
#ifndef OMPBLAS_DROPIN_HPP
#define OMPBLAS_DROPIN_HPP

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "ompblas.h"

namespace amt {

namespace layout {
struct first_order {};
struct last_order {};
}  // namespace layout

template <class Tag>
struct layout_traits;  // `first` is true for column-major ("first order")
template <>
struct layout_traits<layout::first_order> {
    static constexpr bool first = true;
};
template <>
struct layout_traits<layout::last_order> {
    static constexpr bool first = false;
};

namespace dropin {

// Process-wide choices of the drop-in layer: the HIP stream its calls use
// (nullptr = the default stream) and the arithmetic of mtm_helper.
struct options {
    omb_stream_t stream = nullptr;
    int32_t gemm_mode = OMB_MODE_PLAIN_F32;
};
inline options& settings() {
    static options o;
    return o;
}

inline void check(omb_status st, const char* what) {
    if (st != OMB_OK) throw std::runtime_error(std::string(what) + ": " + omb_last_error());
}

// Elements spanned by a rows x cols view with strides (rs, cs).
inline std::size_t span(int64_t rows, int64_t cols, int64_t rs, int64_t cs) {
    if (rows <= 0 || cols <= 0) return 0;
    return static_cast<std::size_t>((rows - 1) * rs + (cols - 1) * cs + 1);
}

// Device buffers for staging host operands, kept between calls: a request
// takes the smallest free cached block that holds it (blocks are sized in
// 2 MiB steps), else allocates one; a block comes back after its call has
// synchronised.  At most `keep_bytes` of free blocks stay cached (larger
// ones are freed on return); release_staging() frees every free block.
// The pool is never destroyed (device memory is not freed during static
// destruction; the process's end reclaims it).
class staging_pool {
   public:
    using block = std::pair<void*, std::size_t>;  // device pointer, bytes
    static staging_pool& get() {
        static staging_pool* p = new staging_pool();
        return *p;
    }
    void* acquire(std::size_t bytes) {
        const std::size_t want = round(bytes);
        {
            std::lock_guard<std::mutex> lock(mu_);
            auto best = free_.end();
            for (auto it = free_.begin(); it != free_.end(); ++it)
                if (it->second >= want && (best == free_.end() || it->second < best->second))
                    best = it;
            if (best != free_.end()) {
                const block b = *best;
                free_.erase(best);
                cached_ -= b.second;
                used_.push_back(b);
                return b.first;
            }
        }
        void* d = nullptr;
        omb_status st = omb_device_alloc(&d, want);
        if (st != OMB_OK) {  // out of memory: drop the cache and try once more
            release_staging();
            st = omb_device_alloc(&d, want);
        }
        if (st != OMB_OK)
            throw std::runtime_error(std::string("amt drop-in staging: ") + omb_last_error());
        std::lock_guard<std::mutex> lock(mu_);
        used_.emplace_back(d, want);
        return d;
    }
    void release(void* p) {
        std::lock_guard<std::mutex> lock(mu_);
        auto it = std::find_if(used_.begin(), used_.end(),
                               [&](const block& b) { return b.first == p; });
        if (it == used_.end()) return;
        const block b = *it;
        used_.erase(it);
        if (cached_ + b.second > keep_bytes) {
            (void)omb_device_free(b.first);
            return;
        }
        free_.push_back(b);
        cached_ += b.second;
    }
    void release_staging() {
        std::lock_guard<std::mutex> lock(mu_);
        for (const block& b : free_) (void)omb_device_free(b.first);
        free_.clear();
        cached_ = 0;
    }
    std::size_t cached_bytes() {
        std::lock_guard<std::mutex> lock(mu_);
        return cached_;
    }
    std::size_t keep_bytes = std::size_t(4) << 30;

   private:
    static std::size_t round(std::size_t b) {
        const std::size_t g = std::size_t(2) << 20;
        return (b + g - 1) / g * g;
    }
    std::mutex mu_;
    std::vector<block> free_, used_;
    std::size_t cached_ = 0;
};

// Frees the drop-in layer's cached staging buffers (see staging_pool).
inline void release_staging() { staging_pool::get().release_staging(); }

// A device view of `count` elements at `p`: device memory passes through,
// host memory is copied to a staging buffer from the pool (when `in`) and
// back by finish() (when `out`).
template <class T>
class staged {
   public:
    staged(T* p, std::size_t count, bool in, bool out, omb_stream_t s)
        : host_(p), count_(count), out_(out), stream_(s) {
        if (!p || count == 0 || omb_is_device_pointer(p)) {
            dev_ = p;
            return;
        }
        void* d = staging_pool::get().acquire(count * sizeof(T));
        dev_ = static_cast<T*>(d);
        owned_ = true;
        if (in) check(omb_memcpy_h2d(d, host_, count * sizeof(T), stream_), "amt drop-in staging");
    }
    staged(const staged&) = delete;
    staged& operator=(const staged&) = delete;
    ~staged() {
        if (owned_) {
            (void)omb_stream_synchronize(stream_);
            staging_pool::get().release(const_cast<std::remove_const_t<T>*>(dev_));
        }
    }
    T* get() const { return dev_; }
    void finish() {
        if (owned_ && out_)
            check(omb_memcpy_d2h(const_cast<std::remove_const_t<T>*>(host_), dev_,
                                 count_ * sizeof(T), stream_),
                  "amt drop-in staging");
        check(omb_stream_synchronize(stream_), "amt drop-in");
    }

   private:
    T* host_;
    T* dev_ = nullptr;
    std::size_t count_;
    bool out_;
    bool owned_ = false;
    omb_stream_t stream_;
};

}  // namespace dropin

// C += A*B; M = na[0], K = na[1], N = nb[1]; any row/column-major mix
// (include/mtm.hpp:116-206).  OutLayout only picks the reference's
// micro-kernel store order and does not change the result.
template <class OutLayout, class ValueType, class SizeType>
void mtm_helper(ValueType* c, SizeType const* nc, SizeType const* wc, ValueType const* a,
                SizeType const* na, SizeType const* wa, ValueType const* b, SizeType const* nb,
                SizeType const* wb, OutLayout) {
    if constexpr (!std::is_same_v<ValueType, float>) {
        throw std::runtime_error("amt::mtm_helper: only float is implemented on the GPU");
    } else {
        (void)nc;
        const int64_t M = static_cast<int64_t>(na[0]), K = static_cast<int64_t>(na[1]);
        const int64_t N = static_cast<int64_t>(nb[1]);
        const int64_t rsA = wa[0], csA = wa[1], rsB = wb[0], csB = wb[1], rsC = wc[0], csC = wc[1];
        auto& o = dropin::settings();
        dropin::staged<const float> A(a, dropin::span(M, K, rsA, csA), true, false, o.stream);
        dropin::staged<const float> B(b, dropin::span(K, N, rsB, csB), true, false, o.stream);
        dropin::staged<float> C(c, dropin::span(M, N, rsC, csC), true, true, o.stream);
        dropin::check(omb_sgemm(M, N, K, 1.f, A.get(), rsA, csA, B.get(), rsB, csB, 1.f, C.get(),
                                rsC, csC, o.gemm_mode, o.stream),
                      "amt::mtm_helper");
        C.finish();
    }
}

// first_order: c += A*b with M = nc[0]*nc[1], N = nb[0]*nb[1], A strides
// (wa[0], wa[1]) (include/mtv.hpp:15-70).  last_order: c = A*b with
// M = na[0], N = nb[0]*nb[1], rows max(wa[0], wa[1]) apart (include/mtv.hpp:72-98).
template <class ValueType, class SizeType, class Order>
void mtv_helper(ValueType* c, SizeType const* nc, ValueType const* a, SizeType const* na,
                SizeType const* wa, ValueType const* b, SizeType const* nb, int max_threads,
                Order) {
    (void)max_threads;
    if constexpr (!std::is_same_v<ValueType, float>) {
        throw std::runtime_error("amt::mtv_helper: only float is implemented on the GPU");
    } else {
        const bool first = layout_traits<Order>::first;
        const int64_t N = static_cast<int64_t>(nb[0] * nb[1]);
        int64_t M, rs, cs;
        if (first) {
            M = static_cast<int64_t>(nc[0] * nc[1]);
            rs = wa[0];
            cs = wa[1];
        } else {
            M = static_cast<int64_t>(na[0]);
            rs = static_cast<int64_t>(wa[0] > wa[1] ? wa[0] : wa[1]);
            cs = 1;
        }
        auto& o = dropin::settings();
        dropin::staged<const float> A(a, dropin::span(M, N, rs, cs), true, false, o.stream);
        dropin::staged<const float> X(b, static_cast<std::size_t>(N), true, false, o.stream);
        dropin::staged<float> Y(c, static_cast<std::size_t>(M), first, true, o.stream);
        dropin::check(omb_sgemv(M, N, 1.f, A.get(), rs, cs, X.get(), first ? 1.f : 0.f, Y.get(),
                                o.stream),
                      "amt::mtv_helper");
        Y.finish();
    }
}

// *c = sum_i a[i]*b[i], n = na[0]*na[1] (include/dot.hpp:143-174).
template <class ValueType, class SizeType>
void dot_prod_helper(ValueType* c, ValueType const* a, SizeType const* na, ValueType const* b,
                     SizeType const* nb, int max_threads) {
    (void)nb;
    (void)max_threads;
    if constexpr (!std::is_same_v<ValueType, float>) {
        throw std::runtime_error("amt::dot_prod_helper: only float is implemented on the GPU");
    } else {
        const int64_t n = static_cast<int64_t>(na[0] * na[1]);
        auto& o = dropin::settings();
        dropin::staged<const float> X(a, static_cast<std::size_t>(n), true, false, o.stream);
        dropin::staged<const float> Y(b, static_cast<std::size_t>(n), true, false, o.stream);
        dropin::staged<float> R(c, 1, false, true, o.stream);
        dropin::check(omb_sdot(n, X.get(), Y.get(), R.get(), o.stream), "amt::dot_prod_helper");
        R.finish();
    }
}

// The reference's tuning entry (include/dot.hpp:59-89): the same *c = a . b
// with caller-chosen cache block sizes for the CPU loops.  The GPU kernel's
// grid depends only on n, so the blocks are accepted and not used.
template <class ValueType, class SizeType>
void dot_prod_helper_tuning_param(ValueType* c, ValueType const* a, ValueType const* b,
                                  SizeType n, int max_threads, SizeType const block1,
                                  SizeType const block2, SizeType const block3) {
    (void)block1;
    (void)block2;
    (void)block3;
    const SizeType na[2] = {n, 1};
    dot_prod_helper(c, a, na, b, na, max_threads);
}

// Tags of transpose_helper (the reference's amt::tag, include/utils.hpp:71-73).
namespace tag {
struct inplace {};
struct outplace {};
}  // namespace tag

// C line i (WC = max(wc) apart) += a * b[i] for i < nb[0]*nb[1], lines of
// na[0]*na[1] elements (include/outer.hpp:15-48): a rank-1 update of the
// lines viewed as a row-major NB x NA matrix.
template <class ValueType, class SizeType>
void outer_prod_helper(ValueType* c, SizeType const* nc, SizeType const* wc, ValueType const* a,
                       SizeType const* na, ValueType const* b, SizeType const* nb,
                       int max_threads) {
    (void)nc;
    (void)max_threads;
    if constexpr (!std::is_same_v<ValueType, float>) {
        throw std::runtime_error("amt::outer_prod_helper: only float is implemented on the GPU");
    } else {
        const int64_t NA = static_cast<int64_t>(na[0] * na[1]);
        const int64_t NB = static_cast<int64_t>(nb[0] * nb[1]);
        const int64_t WC = static_cast<int64_t>(wc[0] > wc[1] ? wc[0] : wc[1]);
        auto& o = dropin::settings();
        dropin::staged<const float> X(b, static_cast<std::size_t>(NB), true, false, o.stream);
        dropin::staged<const float> Y(a, static_cast<std::size_t>(NA), true, false, o.stream);
        dropin::staged<float> C(c, dropin::span(NB, NA, WC, 1), true, true, o.stream);
        dropin::check(omb_sger(NB, NA, 1.f, X.get(), Y.get(), C.get(), WC, 1, o.stream),
                      "amt::outer_prod_helper");
        C.finish();
    }
}

// Out of place: C = A^T, A extents na, strides wa; C strides wc
// (include/trans.hpp:23-45).
template <class ValueType, class SizeType>
void transpose_helper(ValueType* c, SizeType const* nc, SizeType const* wc, ValueType const* a,
                      SizeType const* na, SizeType const* wa, tag::outplace) {
    (void)nc;
    if constexpr (!std::is_same_v<ValueType, float>) {
        throw std::runtime_error("amt::transpose_helper: only float is implemented on the GPU");
    } else {
        const int64_t M = static_cast<int64_t>(na[0]), N = static_cast<int64_t>(na[1]);
        auto& o = dropin::settings();
        dropin::staged<const float> A(a, dropin::span(M, N, wa[0], wa[1]), true, false, o.stream);
        dropin::staged<float> C(c, dropin::span(N, M, wc[0], wc[1]), false, true, o.stream);
        dropin::check(omb_stranspose(M, N, A.get(), wa[0], wa[1], C.get(), wc[0], wc[1], o.stream),
                      "amt::transpose_helper");
        C.finish();
    }
}

// In place: the square matrix at a (na[0] x na[0], rows max(wa) apart) is
// replaced by its transpose (include/trans.hpp:47-71; c == a there).
template <class ValueType, class SizeType>
void transpose_helper(ValueType* c, SizeType const* nc, SizeType const* wc, ValueType* a,
                      SizeType const* na, SizeType const* wa, tag::inplace) {
    (void)c;
    (void)nc;
    (void)wc;
    if constexpr (!std::is_same_v<ValueType, float>) {
        throw std::runtime_error("amt::transpose_helper: only float is implemented on the GPU");
    } else {
        const int64_t n = static_cast<int64_t>(na[0]);
        if (na[1] != na[0])
            throw std::runtime_error("amt::transpose_helper: in-place transpose needs a square matrix");
        const int64_t ld = static_cast<int64_t>(wa[0] > wa[1] ? wa[0] : wa[1]);
        auto& o = dropin::settings();
        dropin::staged<float> A(a, dropin::span(n, n, ld, 1), true, true, o.stream);
        dropin::check(omb_stranspose_inplace(n, A.get(), ld, o.stream), "amt::transpose_helper");
        A.finish();
    }
}

// ------------------------------------------------------------------------
// Tensor-level closures (include/mtm.hpp:208-267, include/mtv.hpp:100-236,
// include/dot.hpp:176-224, include/outer.hpp:51-122, include/trans.hpp:82-166):
// validate shapes when the closure is made (std::runtime_error, as the
// reference) and run the helper when it is called.  Any 2-D tensor type with
// data(), extents() and strides() (indexable, size() == 2) works, uBLAS
// tensors included; amt::matrix_view below is a minimal one over caller
// memory.  Unlike the reference, the closures copy extents and strides, so
// only the data must outlive them.

// Row/column extents and strides of a caller-owned 2-D array.
template <class T>
struct matrix_view {
    T* ptr;
    std::array<std::size_t, 2> n;
    std::array<std::size_t, 2> w;
    T* data() const { return ptr; }
    const std::array<std::size_t, 2>& extents() const { return n; }
    const std::array<std::size_t, 2>& strides() const { return w; }
    // rows x cols with the unit stride along the rows (first order) or the
    // columns (last order)
    static matrix_view first_order(T* p, std::size_t rows, std::size_t cols) {
        return {p, {rows, cols}, {1, rows}};
    }
    static matrix_view last_order(T* p, std::size_t rows, std::size_t cols) {
        return {p, {rows, cols}, {cols, 1}};
    }
};

namespace dropin {
template <class Tn>
std::array<std::size_t, 2> ext2(const Tn& t) {
    const auto& e = t.extents();
    if (e.size() != 2) throw std::runtime_error("amt: only rank-2 tensors are supported");
    return {static_cast<std::size_t>(e[0]), static_cast<std::size_t>(e[1])};
}
template <class Tn>
std::array<std::size_t, 2> str2(const Tn& t) {
    const auto& w = t.strides();
    return {static_cast<std::size_t>(w[0]), static_cast<std::size_t>(w[1])};
}
inline bool is_vec(const std::array<std::size_t, 2>& n) { return n[0] == 1 || n[1] == 1; }
inline bool first_order_of(const std::array<std::size_t, 2>& n,
                           const std::array<std::size_t, 2>& w) {
    return w[0] == 1 && !(w[1] == 1 && n[1] > 1);
}
}  // namespace dropin

template <class TC, class TA, class TB>
auto mtm(TC& c, const TA& a, const TB& b, std::optional<std::size_t> num_threads = std::nullopt) {
    (void)num_threads;
    const auto na = dropin::ext2(a), nb = dropin::ext2(b), nc = dropin::ext2(c);
    if (!(na[0] == nc[0] && na[1] == nb[0] && nc[1] == nb[1]))
        throw std::runtime_error("amt::mtm: dimension mismatch");
    const auto wa = dropin::str2(a), wb = dropin::str2(b), wc = dropin::str2(c);
    auto* cp = c.data();
    const auto* ap = a.data();
    const auto* bp = b.data();
    return [=] {
        mtm_helper(cp, nc.data(), wc.data(), ap, na.data(), wa.data(), bp, nb.data(), wb.data(),
                   layout::last_order{});
    };
}

// c += A b (first-order A) or c = A b (last-order A).
template <class TC, class TA, class TB>
auto mtv(TC& c, const TA& a, const TB& b, std::optional<std::size_t> num_threads = std::nullopt) {
    (void)num_threads;
    const auto na = dropin::ext2(a), nb = dropin::ext2(b), nc = dropin::ext2(c);
    if (!(dropin::is_vec(nb) && dropin::is_vec(nc)))
        throw std::runtime_error("amt::mtv: c and b must be vector, and a must be a matrix");
    if (nc[0] * nc[1] != na[0] || nb[0] * nb[1] != na[1])
        throw std::runtime_error("amt::mtv: dimension mismatch");
    const auto wa = dropin::str2(a);
    const bool first = dropin::first_order_of(na, wa);
    auto* cp = c.data();
    const auto* ap = a.data();
    const auto* bp = b.data();
    return [=] {
        if (first)
            mtv_helper(cp, nc.data(), ap, na.data(), wa.data(), bp, nb.data(), 1, layout::first_order{});
        else
            mtv_helper(cp, nc.data(), ap, na.data(), wa.data(), bp, nb.data(), 1, layout::last_order{});
    };
}

// c = A^T b, through the helper of the other order on the transposed view:
// first-order A overwrites c, last-order A accumulates into it (as the
// reference's vtm, but the true transpose for every shape).
template <class TC, class TA, class TB>
auto vtm(TC& c, const TA& a, const TB& b, std::optional<std::size_t> num_threads = std::nullopt) {
    (void)num_threads;
    const auto na = dropin::ext2(a), nb = dropin::ext2(b), nc = dropin::ext2(c);
    if (!(dropin::is_vec(nb) && dropin::is_vec(nc)))
        throw std::runtime_error("amt::vtm: c and b must be vector, and a must be a matrix");
    if (nc[0] * nc[1] != na[1] || nb[0] * nb[1] != na[0])
        throw std::runtime_error("amt::vtm: dimension mismatch");
    const auto wa = dropin::str2(a);
    const std::array<std::size_t, 2> nt = {na[1], na[0]}, wt = {wa[1], wa[0]};
    const bool first_t = dropin::first_order_of(nt, wt);
    auto* cp = c.data();
    const auto* ap = a.data();
    const auto* bp = b.data();
    return [=] {
        if (first_t)
            mtv_helper(cp, nc.data(), ap, nt.data(), wt.data(), bp, nb.data(), 1, layout::first_order{});
        else
            mtv_helper(cp, nc.data(), ap, nt.data(), wt.data(), bp, nb.data(), 1, layout::last_order{});
    };
}

template <class TC, class TA, class TB>
auto dot_prod(TC& c, const TA& a, const TB& b,
              std::optional<std::size_t> num_threads = std::nullopt) {
    (void)num_threads;
    const auto na = dropin::ext2(a), nb = dropin::ext2(b);
    if (!(dropin::is_vec(na) && dropin::is_vec(nb)))
        throw std::runtime_error("amt::dot_prod: both tensor must be vector");
    if (na[0] * na[1] != nb[0] * nb[1])
        throw std::runtime_error("amt::dot_prod: both vector must be of same size");
    auto* cp = c.data();
    const auto* ap = a.data();
    const auto* bp = b.data();
    return [=] { dot_prod_helper(cp, ap, na.data(), bp, nb.data(), 1); };
}

// include/dot.hpp:91-140: c is a scalar of the vectors' value type,
// captured by reference as in the reference.
template <class Out, class TA, class TB>
auto dot_prod_tuning_param(Out& c, const TA& a, const TB& b, std::size_t b1, std::size_t b2,
                           std::size_t b3, std::optional<std::size_t> num_threads) {
    (void)num_threads;
    const auto na = dropin::ext2(a), nb = dropin::ext2(b);
    if (!(dropin::is_vec(na) && dropin::is_vec(nb)))
        throw std::runtime_error("amt::dot_prod: both tensor must be vector");
    if (na[0] * na[1] != nb[0] * nb[1])
        throw std::runtime_error("amt::dot_prod: both vector must be of same size");
    const auto* ap = a.data();
    const auto* bp = b.data();
    const std::size_t n = na[0] * na[1];
    return [&c, ap, bp, n, b1, b2, b3] {
        dot_prod_helper_tuning_param(&c, ap, bp, n, 1, b1, b2, b3);
    };
}

template <class TC, class TA, class TB>
auto outer_prod(TC& c, const TA& a, const TB& b,
                std::optional<std::size_t> num_threads = std::nullopt) {
    (void)num_threads;
    const auto na = dropin::ext2(a), nb = dropin::ext2(b), nc = dropin::ext2(c);
    if (!(dropin::is_vec(na) && dropin::is_vec(nb)))
        throw std::runtime_error("amt::outer_prod: both tensor must be vector");
    if (nc[0] != na[0] * na[1] || nc[1] != nb[0] * nb[1])
        throw std::runtime_error("amt::outer_prod: dimension mismatch");
    const auto wc0 = dropin::str2(c);
    const bool first = dropin::first_order_of(nc, wc0);
    auto* cp = c.data();
    const auto* ap = a.data();
    const auto* bp = b.data();
    // first order: C's columns are the lines; last order: its rows, a and b swapped
    const std::array<std::size_t, 2> wc = first ? std::array<std::size_t, 2>{1, wc0[1]}
                                                : std::array<std::size_t, 2>{wc0[0], 1};
    return [=] {
        if (first)
            outer_prod_helper(cp, nc.data(), wc.data(), ap, na.data(), bp, nb.data(), 1);
        else
            outer_prod_helper(cp, nc.data(), wc.data(), bp, nb.data(), ap, na.data(), 1);
    };
}

template <class TC, class TA>
auto transpose(TC& c, const TA& a, std::optional<std::size_t> num_threads = std::nullopt) {
    (void)num_threads;
    const auto na = dropin::ext2(a), nc = dropin::ext2(c);
    if (!(na[0] == nc[1] && na[1] == nc[0]))
        throw std::runtime_error("amt::transpose: dimension mismatch");
    const auto wa = dropin::str2(a), wc = dropin::str2(c);
    auto* cp = c.data();
    const auto* ap = a.data();
    return [=] { transpose_helper(cp, nc.data(), wc.data(), ap, na.data(), wa.data(), tag::outplace{}); };
}

template <class TA>
auto transpose(TA& a, std::optional<std::size_t> num_threads = std::nullopt) {
    (void)num_threads;
    const auto na = dropin::ext2(a);
    if (na[0] != na[1])
        throw std::runtime_error("amt::transpose: in-place transpose needs a square matrix");
    const auto wa = dropin::str2(a);
    auto* ap = a.data();
    return [=] { transpose_helper(ap, na.data(), wa.data(), ap, na.data(), wa.data(), tag::inplace{}); };
}

}  // namespace amt

#endif  // OMPBLAS_DROPIN_HPP
