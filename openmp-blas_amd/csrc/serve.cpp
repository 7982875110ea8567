// Request-serving loop: continuous batching of mixed GEMM / GEMV / dot
// requests over device-resident operands.  (The reference has no serving
// layer; each request is one of its mtm / mtv / dot calls.)
//
// The loop replays a trace of arrival times on the calling thread:
//   - a request becomes eligible at its arrival time;
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
#include <list>
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

}  // namespace
}  // namespace omb

using namespace omb;

extern "C" omb_status omb_serve_run(omb_request* reqs, int32_t n, const omb_serve_config* cfg,
                                    omb_request_result* results, omb_serve_stats* stats) {
    clear_error();
    if (n < 0 || (n > 0 && (!reqs || !results)) || !cfg || cfg->slots < 1 || cfg->slots > 64 ||
        cfg->hbm_budget < 0 || cfg->hbm_reserve < 0) {
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

    size_t budget = (size_t)cfg->hbm_budget;
    if (budget == 0) {
        size_t freeb = 0, total = 0;
        OMB_HIP_TRY(hipMemGetInfo(&freeb, &total));
        budget = freeb > (size_t)cfg->hbm_reserve ? freeb - (size_t)cfg->hbm_reserve : 0;
    }
    // At most `slots` requests hold workspace at once.
    const size_t arena_bytes = std::min(budget, (size_t)cfg->slots * std::max<size_t>(max_need, 256));
    Arena arena;
    if (n > 0) OMB_TRY(arena.init(arena_bytes));

    std::vector<Slot> slots(cfg->slots);
    auto cleanup = [&]() {
        for (auto& s : slots) {
            if (s.stream) (void)hipStreamSynchronize(s.stream);
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.stream) (void)hipStreamDestroy(s.stream);
            batch_destroy(s.gb);
            s.gb = nullptr;
        }
    };
    for (auto& s : slots) {
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            cleanup();
            set_error("omb_serve_run: stream/event creation failed");
            return OMB_EDEVICE;
        }
    }
    // Each slot's descriptor block for grouped launches, made before the
    // clock starts (a pinned and a device allocation take milliseconds); if
    // they cannot be had, requests run a stream each.
    bool batching = cfg->no_batching == 0 && n > 1;
    if (batching) {
        for (auto& s : slots) {
            if (batch_create(&s.gb) != OMB_OK) {
                clear_error();
                for (auto& t : slots) {
                    batch_destroy(t.gb);
                    t.gb = nullptr;
                }
                batching = false;
                break;
            }
        }
    }
    // Each slot's first grouped launch, and its stream's first copy from
    // pinned memory, take milliseconds once (measured 5-9 ms inside the timed
    // loop): they happen here, on every slot, before the clock starts.
    for (auto& s : slots) {
        if (batching && batch_warm(s.gb, s.stream) != OMB_OK) {
            (void)hipGetLastError();
            clear_error();
        }
    }
    // Operands may still be in production on other streams (the slots do not
    // wait for the null stream): finish all queued device work first.
    if (hipDeviceSynchronize() != hipSuccess) {
        (void)hipGetLastError();
        cleanup();
        set_error("omb_serve_run: device synchronisation failed");
        return OMB_EDEVICE;
    }

    const auto t0 = Clock::now();
    auto now_s = [&]() { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    std::list<int> waiting;
    std::vector<char> deferred(n, 0);  // skipped by an admission scan (slot or memory)
    int next = 0, finished = 0, in_flight = 0, max_in_flight = 0, late = 0;
    int busy_slots = 0, batches = 0, batched = 0;
    std::vector<omb_gemm_desc> bg;
    std::vector<omb_gemv_desc> bv;
    std::vector<int> cand;
    double last_done = 0.0;
    while (finished < n) {
        bool progress = false;
        double t = now_s();
        // 1. completions
        for (int si = 0; si < (int)slots.size(); ++si) {
            Slot& s = slots[si];
            if (!s.busy()) continue;
            hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) (void)hipGetLastError();
            t = now_s();
            auto complete = [&](int r) {
                omb_request_result& res = results[r];
                if (q != hipSuccess) res.status = OMB_EDEVICE;
                res.done_s = t;
                res.latency_s = t - reqs[r].arrival_s;
                --in_flight;
                ++finished;
            };
            if (s.req >= 0) complete(s.req);
            for (int r : s.batch) complete(r);
            last_done = t;
            if (s.bytes) arena.release(s.off, s.bytes);
            s.req = -1;
            s.off = s.bytes = 0;
            s.batch.clear();
            --busy_slots;
            progress = true;
        }
        // 2. arrivals
        while (next < n && reqs[order[next]].arrival_s <= t) waiting.push_back(order[next++]);
        // 3a. small requests eligible together: one grouped launch on a free
        // slot (one arithmetic mode per launch: the first GEMM's), when at
        // least one request per slot is waiting.  Without such a backlog a
        // stream per request measured the same or better p99; with it the
        // shared launches raise throughput (C5 trace at 96k requests/s:
        // 44.1 -> 45.2-46.8 TFLOP/s).
        if (batching && busy_slots < (int)slots.size() &&
            waiting.size() >= slots.size()) {
            cand.clear();
            int bmode = -1;
            int64_t items = 0;
            for (int r : waiting) {
                if ((int)cand.size() >= kBatchMax) break;
                const omb_request& q = reqs[r];
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
                cand.push_back(r);
            }
            if (cand.size() >= 2) {
                int si = 0;
                while (slots[si].busy()) ++si;
                Slot& s = slots[si];
                omb_status st = OMB_OK;
                bg.clear();
                bv.clear();
                for (int r : cand) {
                    const omb_request& q = reqs[r];
                    if (q.kind == OMB_REQ_GEMM)
                        bg.push_back(omb_gemm_desc{q.m, q.n, q.k, q.a, q.k, q.b, q.n, q.c, q.n, 1.f,
                                                   0.f});
                    else
                        bv.push_back(omb_gemv_desc{q.m, q.n, q.a, q.n, q.b, q.c, 1.f, 0.f});
                }
                if (st == OMB_OK)
                    st = batch_launch(s.gb, bg.data(), (int)bg.size(), bv.data(), (int)bv.size(),
                                      bmode < 0 ? OMB_MODE_PLAIN_F32 : bmode, s.stream);
                if (st == OMB_OK && hipEventRecord(s.done, s.stream) != hipSuccess) st = OMB_EDEVICE;
                const double ts = now_s();
                if (st != OMB_OK) {
                    (void)hipGetLastError();
                    (void)hipStreamSynchronize(s.stream);
                    (void)hipGetLastError();
                }
                for (int r : cand) {
                    results[r].start_s = ts;
                    results[r].slot = si;
                    if (deferred[r]) ++late;
                    if (st != OMB_OK) {
                        results[r].status = st;
                        results[r].done_s = ts;
                        results[r].latency_s = ts - reqs[r].arrival_s;
                        ++finished;
                    } else {
                        s.batch.push_back(r);
                        ++in_flight;
                    }
                }
                if (st == OMB_OK) {
                    ++busy_slots;
                    ++batches;
                    batched += (int)cand.size();
                    max_in_flight = std::max(max_in_flight, in_flight);
                }
                for (auto it = waiting.begin(); it != waiting.end();) {
                    if (std::find(cand.begin(), cand.end(), *it) != cand.end()) it = waiting.erase(it);
                    else ++it;
                }
                progress = true;
            }
        }
        // 3b. the others (and a lone small one), in arrival order
        for (auto it = waiting.begin(); it != waiting.end() && busy_slots < (int)slots.size();) {
            const int r = *it;
            const size_t need = (size_t)reqs[r].ws_bytes;
            if (need > arena.capacity()) {  // can never fit: refuse instead of waiting forever
                results[r].status = OMB_EBUSY;
                results[r].start_s = results[r].done_s = t;
                results[r].latency_s = t - reqs[r].arrival_s;
                ++finished;
                it = waiting.erase(it);
                progress = true;
                continue;
            }
            size_t off = 0;
            if (need > 0 && !arena.alloc(need, &off)) {
                ++it;
                continue;
            }
            int si = 0;
            while (slots[si].busy()) ++si;
            Slot& s = slots[si];
            omb_status st = launch_request(reqs[r], need ? arena.base() + off : nullptr, s.stream);
            if (st == OMB_OK && hipEventRecord(s.done, s.stream) != hipSuccess) st = OMB_EDEVICE;
            const double ts = now_s();
            results[r].start_s = ts;
            results[r].slot = si;
            if (deferred[r]) ++late;
            if (st != OMB_OK) {
                (void)hipGetLastError();
                // Kernels of this request may already be queued (a split pack
                // before a failed launch): the block goes back to the arena
                // only once the slot's stream has drained.
                (void)hipStreamSynchronize(s.stream);
                (void)hipGetLastError();
                results[r].status = st;
                results[r].done_s = ts;
                results[r].latency_s = ts - reqs[r].arrival_s;
                if (need) arena.release(off, need);
                ++finished;
            } else {
                s.req = r;
                s.off = off;
                s.bytes = need;
                ++in_flight;
                ++busy_slots;
                max_in_flight = std::max(max_in_flight, in_flight);
            }
            it = waiting.erase(it);
            progress = true;
        }
        for (int r : waiting) deferred[r] = 1;
        if (!progress) {
            if (in_flight == 0 && waiting.empty() && next < n) {
                // idle until the next arrival: sleep most of the gap, spin the rest
                const double gap = reqs[order[next]].arrival_s - now_s();
                if (gap > 200e-6 && !cfg->spin_poll)
                    std::this_thread::sleep_for(std::chrono::duration<double>(gap - 100e-6));
            } else if (!cfg->spin_poll) {
                std::this_thread::yield();
            }
        }
    }
    cleanup();

    if (stats) {
        std::vector<double> lat;
        double flop = 0.0, sum = 0.0;
        for (int i = 0; i < n; ++i) {
            if (results[i].status != OMB_OK) continue;
            lat.push_back(results[i].latency_s);
            sum += results[i].latency_s;
            flop += flop_of(reqs[i]);
        }
        std::sort(lat.begin(), lat.end());
        stats->wall_s = last_done;
        stats->p50_s = pct(lat, 0.50);
        stats->p90_s = pct(lat, 0.90);
        stats->p99_s = pct(lat, 0.99);
        stats->max_s = lat.empty() ? 0.0 : lat.back();
        stats->mean_s = lat.empty() ? 0.0 : sum / (double)lat.size();
        stats->total_flop = flop;
        stats->max_in_flight = max_in_flight;
        stats->admitted_late = late;
        stats->batches = batches;
        stats->batched_requests = batched;
    }
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
