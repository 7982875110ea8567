// Request-serving loop: continuous batching of mixed GEMM / GEMV / dot
// requests over device-resident operands.  (The reference has no serving
// layer; each request is one of its mtm / mtv / dot calls.)
//
// One admission engine serves two front ends: omb_serve_run replays a trace
// of arrival times on the calling thread, and the live server
// (omb_server_*) runs it on a worker thread fed by submissions.
//   - a request becomes eligible at its arrival (or submission) time;
//   - it is admitted when a slot (one HIP stream each) is free and its
//     workspace (split-BF16 planes, dot partials) fits in the HBM budget,
//     taken first-fit from one pre-allocated arena -- eligible requests are
//     scanned in arrival order and one that does not fit yet does not hold
//     back smaller ones behind it;
//   - when at least one request per slot is waiting, small GEMM and GEMV
//     requests that are eligible together (two or more) run as one grouped
//     launch on one slot (the persistent kernel of the grouped plans, over a
//     per-slot descriptor block) instead of a stream each -- only those
//     whose direct call does not split K into coarse blocks, so that every
//     result keeps its solo bits;
//   - admitted requests run concurrently with those already on the GPU;
//   - completions are observed by polling one event per slot, which frees
//     the slot and the workspace for the next admissions.
// Kernels are the same deterministic ones as the direct entry points and
// each request owns its workspace, so a result does not depend on what ran
// beside it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "common.hpp"
#include "host_util.hpp"

namespace omb {
namespace {

using Clock = std::chrono::steady_clock;

// One device buffer handed out first-fit in 256-byte granules.
class Arena {
   public:
    ~Arena() {
        if (base_) (void)hipFree(base_);
    }
    omb_status init(size_t bytes) {
        fit_.reset(bytes);
        if (hipMalloc(&base_, fit_.capacity()) != hipSuccess) {
            (void)hipGetLastError();
            base_ = nullptr;
            set_error("serve: arena allocation of " + std::to_string(fit_.capacity()) +
                      " bytes failed");
            return OMB_ENOMEM;
        }
        return OMB_OK;
    }
    size_t capacity() const { return fit_.capacity(); }
    bool alloc(size_t bytes, size_t* off) { return fit_.alloc(bytes, off); }
    void release(size_t off, size_t bytes) { fit_.release(off, bytes); }
    char* base() const { return static_cast<char*>(base_); }

   private:
    void* base_ = nullptr;
    FirstFit fit_;
};

// Planes a split-BF16 request packs: the 3-product form reads only the
// first two, so it packs and stages two, as omb_sgemm does.
int split_planes(const omb_request& r) { return r.mode == OMB_MODE_SPLIT_BF16_P3 ? 2 : 3; }

int64_t packed_elems(const omb_request& r, int64_t rows) {
    return split_planes(r) == 2 ? split16_packed_elems(rows, r.k) : split_packed_elems(rows, r.k);
}

// Packed split-BF16 operands of a GEMM request, rounded to 256 bytes (the
// split partial sums, when K is split, follow them).
size_t packed_bytes(const omb_request& r) {
    return (size_t)round_up((packed_elems(r, r.m) + packed_elems(r, r.n)) * (int64_t)sizeof(__bf16),
                            256);
}

size_t need_bytes(const omb_request& r) {
    switch (r.kind) {
        case OMB_REQ_GEMM: {
            if (r.m == 0 || r.n == 0 || r.k == 0) return 0;
            if (r.mode == OMB_MODE_PLAIN_F32) {  // partial sums when K is split
                int64_t kc = 0;
                const int64_t splits = sgemm_f32_splits(r.m, r.n, r.k, &kc);
                return splits > 1 ? (size_t)(splits * r.m * r.n) * sizeof(float) : 0;
            }
            if (r.mode == OMB_MODE_SPLIT_FP16) return split16_ws_bytes(r.m, r.n, r.k);
            int64_t kc = 0;
            const int64_t splits = sgemm_split_splits(r.m, r.n, r.k, &kc);
            return packed_bytes(r) +
                   (splits > 1 ? (size_t)(splits * r.m * r.n) * sizeof(float) : 0);
        }
        case OMB_REQ_DOT: return r.n > 0 ? (size_t)sdot_blocks(r.n) * sizeof(double) : 0;
        default: return 0;
    }
}

double flop_of(const omb_request& r) {
    switch (r.kind) {
        case OMB_REQ_GEMM: return 2.0 * r.m * r.n * r.k;
        case OMB_REQ_GEMV: return 2.0 * r.m * r.n;
        default: return 2.0 * r.n;
    }
}

omb_status check_request(const omb_request& r, int i) {
    bool ok = true;
    switch (r.kind) {
        case OMB_REQ_GEMM:
            ok = r.m >= 0 && r.n >= 0 && r.k >= 0 &&
                 (r.mode == OMB_MODE_PLAIN_F32 || r.mode == OMB_MODE_SPLIT_BF16 ||
                  r.mode == OMB_MODE_SPLIT_BF16_P3 || r.mode == OMB_MODE_SPLIT_FP16) &&
                 (r.m == 0 || r.n == 0 || (r.c && (r.k == 0 || (r.a && r.b))));
            break;
        case OMB_REQ_GEMV:
            ok = r.m >= 0 && r.n >= 0 && (r.m == 0 || (r.c && (r.n == 0 || (r.a && r.b))));
            break;
        case OMB_REQ_DOT: ok = r.n >= 0 && r.c && (r.n == 0 || (r.a && r.b)); break;
        default: ok = false;
    }
    if (!ok || !(r.arrival_s >= 0.0) || !std::isfinite(r.arrival_s)) {
        set_error("omb_serve_run: invalid request " + std::to_string(i));
        return OMB_EINVAL;
    }
    return OMB_OK;
}

// One request on stream s; all operands row-major and contiguous.
omb_status launch_request(const omb_request& r, char* ws, hipStream_t s) {
    switch (r.kind) {
        case OMB_REQ_GEMM: {
            if (r.m == 0 || r.n == 0) return OMB_OK;
            if (r.k == 0) return launch_scale(r.m, r.n, 0.f, r.c, r.n, s);
            if (r.mode == OMB_MODE_PLAIN_F32)
                return launch_sgemm_f32(r.m, r.n, r.k, 1.f, r.a, r.k, true, r.b, r.n, true, 0.f, r.c,
                                        r.n, s, reinterpret_cast<float*>(ws));
            if (r.mode == OMB_MODE_SPLIT_FP16)
                return launch_sgemm_split16_ws(r.m, r.n, r.k, 1.f, r.a, r.k, 1, r.b, r.n, 1, 0.f,
                                               r.c, r.n, ws, s);
            const int planes = split_planes(r);
            __bf16* ap = reinterpret_cast<__bf16*>(ws);
            __bf16* bp = ap + packed_elems(r, r.m);
            // A, and B^T of row-major B, in one launch
            OMB_TRY(launch_split_pack_pair(r.m, r.n, r.k, r.a, r.k, 1, ap, r.b, 1, r.n, bp, s,
                                           planes));
            return launch_sgemm_packed(r.m, r.n, r.k, 1.f, ap, bp, 0.f, r.c, r.n,
                                       planes == 2 ? 3 : 6, s,
                                       reinterpret_cast<float*>(ws + packed_bytes(r)), nullptr,
                                       nullptr, planes);
        }
        case OMB_REQ_GEMV:
            if (r.m == 0) return OMB_OK;
            if (r.n == 0) return launch_scale(r.m, 1, 0.f, r.c, 1, s);
            return launch_sgemv_row(r.m, r.n, 1.f, r.a, r.n, r.b, 0.f, r.c, s);
        default:
            return launch_sdot_ws(r.n, r.a, r.b, r.c, reinterpret_cast<double*>(ws), s);
    }
}

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    int req = -1;                 // solo request, or -1
    std::vector<int> batch;       // requests of a grouped launch
    GroupBatch* gb = nullptr;     // the slot's descriptor block (made on first use)
    size_t off = 0, bytes = 0;
    bool busy() const { return req >= 0 || !batch.empty(); }
};

// Requests small enough to share a grouped launch: at most 1e8 FLOP per
// GEMM and 4 MiB of A per GEMV (launch-bound on their own), at most 32 per
// launch.  Measured on the C5 trace (3000 requests, split-BF16): larger
// limits (3e8-4e9 FLOP, 8-32 MiB, 512 per launch) put mid-size GEMMs on
// the grouped 64x64 tiles, which are slower than their direct kernels, and
// made p99 worse at 48k-64k requests/s.  A GEMM joins only if its direct
// call keeps the unsplit fold (no K split, or one 128-k / 256-k
// accumulation block per split), since a grouped tile never splits K.
constexpr double kBatchGemmFlop = 1e8;
constexpr double kBatchGemvBytes = 4.0 * (1 << 20);
constexpr int kBatchMax = 32;
constexpr int64_t kBatchMaxItems = 65536;  // the descriptor block's item capacity

bool batchable(const omb_request& r) {
    switch (r.kind) {
        case OMB_REQ_GEMM: {
            if (r.m <= 0 || r.n <= 0 || r.k <= 0 || !batch_mode_supported(r.mode)) return false;
            if (2.0 * r.m * r.n * r.k > kBatchGemmFlop) return false;
            int64_t kc = 0;
            if (r.mode == OMB_MODE_PLAIN_F32) {
                const int64_t sp = sgemm_f32_splits(r.m, r.n, r.k, &kc);
                return sp == 1 || kc == 128;
            }
            const int64_t sp = sgemm_split_splits(r.m, r.n, r.k, &kc);
            return sp == 1 || kc == 256;
        }
        case OMB_REQ_GEMV:
            return r.m > 0 && r.n > 0 && 4.0 * r.m * r.n <= kBatchGemvBytes;
        default: return false;
    }
}

double pct(const std::vector<double>& sorted, double q) {
    if (sorted.empty()) return 0.0;
    size_t i = (size_t)std::ceil(q * (double)sorted.size());
    i = i == 0 ? 0 : i - 1;
    return sorted[std::min(i, sorted.size() - 1)];
}

// Requests and results by index: the caller's arrays (trace replay) or the
// live server's own lists.
struct Store {
    virtual omb_request& req(int i) = 0;
    virtual omb_request_result& res(int i) = 0;
    virtual ~Store() = default;
};

// The admission engine shared by omb_serve_run and the live server: slots,
// arena, the waiting list and the two admission scans.  Its owner feeds it
// arrivals and calls complete() and admit() in turn, all on one thread.
class Engine {
   public:
    ~Engine() { close(); }

    // Streams, events, the arena (slots x max_need bytes within the HBM
    // budget) and each slot's descriptor block, made and warmed before the
    // clock starts.
    omb_status open(const omb_serve_config& cfg, size_t max_need, bool want_arena,
                    bool want_batching) {
        size_t budget = (size_t)cfg.hbm_budget;
        if (budget == 0) {
            size_t freeb = 0, total = 0;
            OMB_HIP_TRY(hipMemGetInfo(&freeb, &total));
            budget = freeb > (size_t)cfg.hbm_reserve ? freeb - (size_t)cfg.hbm_reserve : 0;
        }
        // At most `slots` requests hold workspace at once.
        const size_t arena_bytes =
            std::min(budget, (size_t)cfg.slots * std::max<size_t>(max_need, 256));
        if (want_arena) OMB_TRY(arena_.init(arena_bytes));
        slots_.resize(cfg.slots);
        for (auto& s : slots_) {
            if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) {
                (void)hipGetLastError();
                close();
                set_error("serve: stream/event creation failed");
                return OMB_EDEVICE;
            }
        }
        // Each slot's descriptor block for grouped launches, made before the
        // clock starts (a pinned and a device allocation take milliseconds);
        // if they cannot be had, requests run a stream each.
        batching_ = want_batching;
        if (batching_) {
            for (auto& s : slots_) {
                if (batch_create(&s.gb) != OMB_OK) {
                    clear_error();
                    for (auto& t : slots_) {
                        batch_destroy(t.gb);
                        t.gb = nullptr;
                    }
                    batching_ = false;
                    break;
                }
            }
        }
        // Each slot's first grouped launch, and its stream's first copy from
        // pinned memory, take milliseconds once (measured 5-9 ms inside the
        // timed loop): they happen here, on every slot, before the clock.
        for (auto& s : slots_) {
            if (batching_ && batch_warm(s.gb, s.stream) != OMB_OK) {
                (void)hipGetLastError();
                clear_error();
            }
        }
        // Operands may still be in production on other streams (the slots do
        // not wait for the null stream): finish all queued device work first.
        if (hipDeviceSynchronize() != hipSuccess) {
            (void)hipGetLastError();
            close();
            set_error("serve: device synchronisation failed");
            return OMB_EDEVICE;
        }
        t0_ = Clock::now();
        return OMB_OK;
    }

    // Waits for every slot's stream, then frees streams, events and blocks.
    void close() {
        for (auto& s : slots_) {
            if (s.stream) (void)hipStreamSynchronize(s.stream);
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.stream) (void)hipStreamDestroy(s.stream);
            batch_destroy(s.gb);
            s.gb = nullptr;
            s.stream = nullptr;
            s.done = nullptr;
        }
        slots_.clear();
    }

    double now_s() const { return std::chrono::duration<double>(Clock::now() - t0_).count(); }
    bool idle() const { return in_flight_ == 0 && waiting_.empty(); }
    int in_flight() const { return in_flight_; }
    bool any_waiting() const { return !waiting_.empty(); }

    // Request r (its ws_bytes filled in) is eligible from now on.
    void arrive(int r) {
        if ((size_t)r >= deferred_.size()) deferred_.resize((size_t)r + 1, 0);
        waiting_.push_back(r);
    }

    // Completions observed by polling each busy slot's event; `t` moves to
    // the time of the last one seen.  Finished requests are appended to
    // `done` (if given).
    bool complete(Store& st, double& t, std::vector<int>* done) {
        bool progress = false;
        for (int si = 0; si < (int)slots_.size(); ++si) {
            Slot& s = slots_[si];
            if (!s.busy()) continue;
            hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) (void)hipGetLastError();
            t = now_s();
            auto complete_one = [&](int r) {
                omb_request_result& res = st.res(r);
                if (q != hipSuccess) res.status = OMB_EDEVICE;
                res.done_s = t;
                res.latency_s = t - st.req(r).arrival_s;
                --in_flight_;
                finish(r, done);
            };
            if (s.req >= 0) complete_one(s.req);
            for (int r : s.batch) complete_one(r);
            last_done_ = t;
            if (s.bytes) arena_.release(s.off, s.bytes);
            s.req = -1;
            s.off = s.bytes = 0;
            s.batch.clear();
            --busy_slots_;
            progress = true;
        }
        return progress;
    }

    // The admission scans over the waiting requests, in arrival order.
    bool admit(Store& st, double t, std::vector<int>* done) {
        bool progress = admit_batch(st, done);
        progress |= admit_each(st, t, done);
        for (int r : waiting_) deferred_[r] = 1;
        return progress;
    }

    void stats(Store& st, int n, omb_serve_stats* out) const {
        std::vector<double> lat;
        double flop = 0.0, sum = 0.0;
        for (int i = 0; i < n; ++i) {
            if (st.res(i).status != OMB_OK) continue;
            lat.push_back(st.res(i).latency_s);
            sum += st.res(i).latency_s;
            flop += flop_of(st.req(i));
        }
        std::sort(lat.begin(), lat.end());
        out->wall_s = last_done_;
        out->p50_s = pct(lat, 0.50);
        out->p90_s = pct(lat, 0.90);
        out->p99_s = pct(lat, 0.99);
        out->max_s = lat.empty() ? 0.0 : lat.back();
        out->mean_s = lat.empty() ? 0.0 : sum / (double)lat.size();
        out->total_flop = flop;
        out->max_in_flight = max_in_flight_;
        out->admitted_late = late_;
        out->batches = batches_;
        out->batched_requests = batched_;
    }

   private:
    void finish(int r, std::vector<int>* done) {
        if (done) done->push_back(r);
    }

    // 3a. Small requests eligible together: one grouped launch on a free
    // slot (one arithmetic mode per launch: the first GEMM's), when at least
    // one request per slot is waiting.  Without such a backlog a stream per
    // request measured the same or better p99; with it the shared launches
    // raise throughput (C5 trace at 96k requests/s: 44.1 -> 45.2-46.8
    // TFLOP/s).
    bool admit_batch(Store& st, std::vector<int>* done) {
        if (!batching_ || busy_slots_ >= (int)slots_.size() || waiting_.size() < slots_.size())
            return false;
        cand_.clear();
        int bmode = -1;
        int64_t items = 0;
        for (int r : waiting_) {
            if ((int)cand_.size() >= kBatchMax) break;
            const omb_request& q = st.req(r);
            if (!batchable(q)) continue;
            if (q.kind == OMB_REQ_GEMM) {
                if (bmode < 0) bmode = q.mode;
                if (q.mode != bmode) continue;
            }
            // at most this many work items (32 x 32 tiles are the smallest)
            const int64_t it = q.kind == OMB_REQ_GEMM ? ceil_div(q.m, 32) * ceil_div(q.n, 32)
                                                      : ceil_div(q.m, 64);
            if (items + it > kBatchMaxItems) continue;
            items += it;
            cand_.push_back(r);
        }
        if (cand_.size() < 2) return false;
        int si = 0;
        while (slots_[si].busy()) ++si;
        Slot& s = slots_[si];
        bg_.clear();
        bv_.clear();
        for (int r : cand_) {
            const omb_request& q = st.req(r);
            if (q.kind == OMB_REQ_GEMM)
                bg_.push_back(omb_gemm_desc{q.m, q.n, q.k, q.a, q.k, q.b, q.n, q.c, q.n, 1.f, 0.f});
            else
                bv_.push_back(omb_gemv_desc{q.m, q.n, q.a, q.n, q.b, q.c, 1.f, 0.f});
        }
        omb_status stt = batch_launch(s.gb, bg_.data(), (int)bg_.size(), bv_.data(),
                                      (int)bv_.size(), bmode < 0 ? OMB_MODE_PLAIN_F32 : bmode,
                                      s.stream);
        if (stt == OMB_OK && hipEventRecord(s.done, s.stream) != hipSuccess) stt = OMB_EDEVICE;
        const double ts = now_s();
        if (stt != OMB_OK) {
            (void)hipGetLastError();
            (void)hipStreamSynchronize(s.stream);
            (void)hipGetLastError();
        }
        for (int r : cand_) {
            omb_request_result& res = st.res(r);
            res.start_s = ts;
            res.slot = si;
            if (deferred_[r]) ++late_;
            if (stt != OMB_OK) {
                res.status = stt;
                res.done_s = ts;
                res.latency_s = ts - st.req(r).arrival_s;
                finish(r, done);
            } else {
                s.batch.push_back(r);
                ++in_flight_;
            }
        }
        if (stt == OMB_OK) {
            ++busy_slots_;
            ++batches_;
            batched_ += (int)cand_.size();
            max_in_flight_ = std::max(max_in_flight_, in_flight_);
        }
        for (auto it = waiting_.begin(); it != waiting_.end();) {
            if (std::find(cand_.begin(), cand_.end(), *it) != cand_.end()) it = waiting_.erase(it);
            else ++it;
        }
        return true;
    }

    // 3b. The others (and a lone small one), each on a free slot with its
    // workspace from the arena; one that does not fit yet does not hold back
    // smaller ones behind it.
    bool admit_each(Store& st, double t, std::vector<int>* done) {
        bool progress = false;
        for (auto it = waiting_.begin(); it != waiting_.end() && busy_slots_ < (int)slots_.size();) {
            const int r = *it;
            const omb_request& q = st.req(r);
            omb_request_result& res = st.res(r);
            const size_t need = (size_t)q.ws_bytes;
            if (need > arena_.capacity()) {  // can never fit: refuse instead of waiting forever
                res.status = OMB_EBUSY;
                res.start_s = res.done_s = t;
                res.latency_s = t - q.arrival_s;
                finish(r, done);
                it = waiting_.erase(it);
                progress = true;
                continue;
            }
            size_t off = 0;
            if (need > 0 && !arena_.alloc(need, &off)) {
                ++it;
                continue;
            }
            int si = 0;
            while (slots_[si].busy()) ++si;
            Slot& s = slots_[si];
            omb_status stt = launch_request(q, need ? arena_.base() + off : nullptr, s.stream);
            if (stt == OMB_OK && hipEventRecord(s.done, s.stream) != hipSuccess) stt = OMB_EDEVICE;
            const double ts = now_s();
            res.start_s = ts;
            res.slot = si;
            if (deferred_[r]) ++late_;
            if (stt != OMB_OK) {
                (void)hipGetLastError();
                // Kernels of this request may already be queued (a split pack
                // before a failed launch): the block goes back to the arena
                // only once the slot's stream has drained.
                (void)hipStreamSynchronize(s.stream);
                (void)hipGetLastError();
                res.status = stt;
                res.done_s = ts;
                res.latency_s = ts - q.arrival_s;
                if (need) arena_.release(off, need);
                finish(r, done);
            } else {
                s.req = r;
                s.off = off;
                s.bytes = need;
                ++in_flight_;
                ++busy_slots_;
                max_in_flight_ = std::max(max_in_flight_, in_flight_);
            }
            it = waiting_.erase(it);
            progress = true;
        }
        return progress;
    }

    Arena arena_;
    std::vector<Slot> slots_;
    bool batching_ = false;
    Clock::time_point t0_ = Clock::now();
    std::list<int> waiting_;
    std::vector<char> deferred_;  // skipped by an admission scan (slot or memory)
    int in_flight_ = 0, max_in_flight_ = 0, late_ = 0;
    int busy_slots_ = 0, batches_ = 0, batched_ = 0;
    double last_done_ = 0.0;
    std::vector<omb_gemm_desc> bg_;
    std::vector<omb_gemv_desc> bv_;
    std::vector<int> cand_;
};

struct TraceStore : Store {
    omb_request* reqs;
    omb_request_result* results;
    TraceStore(omb_request* q, omb_request_result* r) : reqs(q), results(r) {}
    omb_request& req(int i) override { return reqs[i]; }
    omb_request_result& res(int i) override { return results[i]; }
};

bool valid_config(const omb_serve_config* cfg) {
    return cfg && cfg->slots >= 1 && cfg->slots <= 64 && cfg->hbm_budget >= 0 &&
           cfg->hbm_reserve >= 0;
}

}  // namespace
}  // namespace omb

using namespace omb;

extern "C" omb_status omb_serve_run(omb_request* reqs, int32_t n, const omb_serve_config* cfg,
                                    omb_request_result* results, omb_serve_stats* stats) {
    clear_error();
    if (n < 0 || (n > 0 && (!reqs || !results)) || !valid_config(cfg)) {
        set_error("omb_serve_run: invalid arguments");
        return OMB_EINVAL;
    }
    for (int i = 0; i < n; ++i) OMB_TRY(check_request(reqs[i], i));

    std::vector<int> order(n);
    size_t max_need = 0;
    for (int i = 0; i < n; ++i) {
        order[i] = i;
        reqs[i].ws_bytes = (int64_t)need_bytes(reqs[i]);
        max_need = std::max(max_need, (size_t)reqs[i].ws_bytes);
        results[i] = omb_request_result{0.0, 0.0, 0.0, OMB_OK, -1};
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return reqs[a].arrival_s < reqs[b].arrival_s; });

    Engine eng;
    OMB_TRY(eng.open(*cfg, max_need, n > 0, cfg->no_batching == 0 && n > 1));
    TraceStore st(reqs, results);
    std::vector<int> done;
    int next = 0, finished = 0;
    while (finished < n) {
        double t = eng.now_s();
        done.clear();
        bool progress = eng.complete(st, t, &done);
        // arrivals up to now
        while (next < n && reqs[order[next]].arrival_s <= t) eng.arrive(order[next++]);
        progress |= eng.admit(st, t, &done);
        finished += (int)done.size();
        if (!progress) {
            if (eng.idle() && next < n) {
                // idle until the next arrival: sleep most of the gap, spin the rest
                const double gap = reqs[order[next]].arrival_s - eng.now_s();
                if (gap > 200e-6 && !cfg->spin_poll)
                    std::this_thread::sleep_for(std::chrono::duration<double>(gap - 100e-6));
            } else if (!cfg->spin_poll) {
                std::this_thread::yield();
            }
        }
    }
    eng.close();
    if (stats) eng.stats(st, n, stats);
    return OMB_OK;
}

// ------------------------------------------------------------ live server --
// The same engine on a worker thread, fed by omb_server_submit instead of a
// trace: a request is eligible when it is submitted, and its result is
// published when its completion is observed.
struct omb_server {
    omb_serve_config cfg{};
    int device = 0;
    omb::Engine eng;
    // worker-only request and result lists (deque: stable references)
    struct LiveStore : omb::Store {
        std::deque<omb_request> reqs;
        std::deque<omb_request_result> results;
        omb_request& req(int i) override { return reqs[(size_t)i]; }
        omb_request_result& res(int i) override { return results[(size_t)i]; }
    } st;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_in, cv_done;
    std::vector<omb_request> inbox;          // submitted, not yet taken (under mu)
    std::vector<omb_request_result> published;  // by ticket (under mu)
    std::vector<char> is_done;               // by ticket (under mu)
    int64_t submitted = 0;                   // tickets handed out (under mu)
    bool closing = false;                    // (under mu)

    void run() {
        (void)hipSetDevice(device);
        std::vector<omb_request> taken;
        std::vector<int> done;
        for (;;) {
            double t = eng.now_s();
            done.clear();
            bool progress = eng.complete(st, t, &done);
            bool stop = false;
            {
                std::lock_guard<std::mutex> lock(mu);
                taken.swap(inbox);
                stop = closing;
            }
            for (omb_request& q : taken) {
                const int r = (int)st.reqs.size();
                st.reqs.push_back(q);
                st.results.push_back(omb_request_result{0.0, 0.0, 0.0, OMB_OK, -1});
                eng.arrive(r);
                progress = true;
            }
            taken.clear();
            progress |= eng.admit(st, t, &done);
            if (!done.empty()) {
                std::lock_guard<std::mutex> lock(mu);
                for (int r : done) {
                    published[(size_t)r] = st.results[(size_t)r];
                    is_done[(size_t)r] = 1;
                }
                cv_done.notify_all();
            }
            if (stop && eng.idle()) {
                std::lock_guard<std::mutex> lock(mu);
                if (inbox.empty()) break;
                continue;
            }
            if (!progress) {
                if (eng.idle() && !cfg.spin_poll) {
                    // nothing on the GPU and nothing waiting: sleep until a
                    // submission (or close) wakes the worker (with spin_poll
                    // it keeps polling, as the trace loop does)
                    std::unique_lock<std::mutex> lock(mu);
                    cv_in.wait_for(lock, std::chrono::milliseconds(10),
                                   [&] { return !inbox.empty() || closing; });
                } else if (!cfg.spin_poll) {
                    std::this_thread::yield();
                }
            }
        }
    }
};

extern "C" omb_status omb_server_open(const omb_serve_config* cfg, int64_t max_request_bytes,
                                      omb_server** out) {
    clear_error();
    if (!out || !valid_config(cfg) || max_request_bytes < 0) {
        set_error("omb_server_open: invalid arguments");
        return OMB_EINVAL;
    }
    *out = nullptr;
    std::unique_ptr<omb_server> s(new omb_server());
    s->cfg = *cfg;
    OMB_HIP_TRY(hipGetDevice(&s->device));
    OMB_TRY(s->eng.open(*cfg, (size_t)max_request_bytes, true, cfg->no_batching == 0));
    s->worker = std::thread([p = s.get()] { p->run(); });
    *out = s.release();
    return OMB_OK;
}

extern "C" omb_status omb_server_submit(omb_server* s, const omb_request* r, int64_t* ticket) {
    clear_error();
    if (!s || !r || !ticket) {
        set_error("omb_server_submit: invalid arguments");
        return OMB_EINVAL;
    }
    omb_request q = *r;
    q.arrival_s = 0.0;  // set to the submission time below
    OMB_TRY(check_request(q, 0));
    q.ws_bytes = (int64_t)need_bytes(q);
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->closing) {
        set_error("omb_server_submit: the server is closing");
        return OMB_EINVAL;
    }
    if (s->submitted >= (int64_t)INT32_MAX) {
        set_error("omb_server_submit: too many requests for one server");
        return OMB_EUNSUPPORTED;
    }
    q.arrival_s = s->eng.now_s();
    *ticket = s->submitted++;
    s->inbox.push_back(q);
    s->published.push_back(omb_request_result{0.0, 0.0, 0.0, OMB_OK, -1});
    s->is_done.push_back(0);
    s->cv_in.notify_one();
    return OMB_OK;
}

extern "C" omb_status omb_server_poll(omb_server* s, int64_t ticket, int32_t* done,
                                      omb_request_result* result) {
    clear_error();
    if (!s || !done) {
        set_error("omb_server_poll: invalid arguments");
        return OMB_EINVAL;
    }
    std::lock_guard<std::mutex> lock(s->mu);
    if (ticket < 0 || ticket >= s->submitted) {
        set_error("omb_server_poll: unknown ticket");
        return OMB_EINVAL;
    }
    *done = s->is_done[(size_t)ticket];
    if (*done && result) *result = s->published[(size_t)ticket];
    return OMB_OK;
}

extern "C" omb_status omb_server_wait(omb_server* s, int64_t ticket, omb_request_result* result) {
    clear_error();
    if (!s) {
        set_error("omb_server_wait: invalid arguments");
        return OMB_EINVAL;
    }
    std::unique_lock<std::mutex> lock(s->mu);
    if (ticket < 0 || ticket >= s->submitted) {
        set_error("omb_server_wait: unknown ticket");
        return OMB_EINVAL;
    }
    s->cv_done.wait(lock, [&] { return s->is_done[(size_t)ticket] != 0; });
    if (result) *result = s->published[(size_t)ticket];
    return OMB_OK;
}

extern "C" omb_status omb_server_close(omb_server* s, omb_serve_stats* stats) {
    clear_error();
    if (!s) {
        set_error("omb_server_close: invalid arguments");
        return OMB_EINVAL;
    }
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->closing = true;
    }
    s->cv_in.notify_one();
    s->worker.join();
    s->eng.close();
    if (stats) s->eng.stats(s->st, (int)s->st.reqs.size(), stats);
    delete s;
    return OMB_OK;
}

extern "C" int64_t omb_serve_request_bytes(const omb_request* r) {
    clear_error();
    if (!r || check_request(*r, 0) != OMB_OK) {
        set_error("omb_serve_request_bytes: invalid request");
        return -1;
    }
    return (int64_t)need_bytes(*r);
}
