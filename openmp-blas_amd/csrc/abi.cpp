// C-ABI entry points: argument validation (before any HIP call), layout
// resolution and folding, per-stream workspaces, error strings.
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>

#include "common.hpp"

namespace omb {

namespace {
thread_local std::string t_error;

struct WsKey {
    int dev;
    hipStream_t stream;
    bool operator<(const WsKey& o) const {
        return dev != o.dev ? dev < o.dev : stream < o.stream;
    }
};

std::mutex g_ws_mu;
std::map<WsKey, std::unique_ptr<Workspace>>& ws_map() {
    static auto* m = new std::map<WsKey, std::unique_ptr<Workspace>>();  // lives to exit
    return *m;
}
}  // namespace

void set_error(const std::string& msg) { t_error = msg; }
void clear_error() { t_error.clear(); }

// The returned buffer stays valid until the next call for the same (device,
// stream) asks for more; calls on one stream are issued by one host thread
// at a time, so no call still holds it then.
omb_status get_workspace(hipStream_t stream, size_t bytes, Workspace** out) {
    int dev = 0;
    OMB_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_ws_mu);
    auto& slot = ws_map()[WsKey{dev, stream}];
    if (!slot) slot = std::make_unique<Workspace>();
    Workspace* ws = slot.get();
    if (bytes > ws->bytes) {
        if (ws->buf) OMB_HIP_TRY(hipFree(ws->buf));  // hipFree waits for work in flight
        ws->buf = nullptr;
        ws->bytes = 0;
        size_t want = ((bytes + (1u << 20) - 1) >> 20) << 20;
        if (hipMalloc(&ws->buf, want) != hipSuccess) {
            (void)hipGetLastError();
            ws->buf = nullptr;
            set_error("workspace: hipMalloc of " + std::to_string(want) + " bytes failed");
            return OMB_ENOMEM;
        }
        ws->bytes = want;
    }
    *out = ws;
    return OMB_OK;
}

int device_cu_count() {
    static std::mutex mu;
    static std::map<int, int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return 256;
    }
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(dev);
    if (it != cache.end()) return it->second;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus < 1) {
        (void)hipGetLastError();
        cus = 256;
    }
    cache[dev] = cus;
    return cus;
}

omb_status resolve_layout(const char* name, int64_t rows, int64_t cols, int64_t rs, int64_t cs,
                          bool* unit_col, int64_t* ld) {
    if (rs < 1 || cs < 1) {
        set_error(std::string(name) + ": strides must be >= 1");
        return OMB_EINVAL;
    }
    if (cs == 1 && rs >= (cols > 1 ? cols : 1)) {
        *unit_col = true;
        *ld = rs;
        return OMB_OK;
    }
    if (rs == 1 && cs >= (rows > 1 ? rows : 1)) {
        *unit_col = false;
        *ld = cs;
        return OMB_OK;
    }
    if (cs == 1 && rows <= 1) {
        *unit_col = true;
        *ld = cols > 1 ? cols : 1;
        return OMB_OK;
    }
    if (rs == 1 && cols <= 1) {
        *unit_col = false;
        *ld = rows > 1 ? rows : 1;
        return OMB_OK;
    }
    if (rs == 1 || cs == 1) {
        set_error(std::string(name) + ": leading dimension smaller than the extent");
        return OMB_EINVAL;
    }
    set_error(std::string(name) + ": one stride must be 1 (first_order or last_order layout)");
    return OMB_EUNSUPPORTED;
}

}  // namespace omb

using namespace omb;

extern "C" const char* omb_last_error(void) { return t_error.c_str(); }

extern "C" const char* omb_version(void) { return "ompblas-gfx950 0.1"; }

extern "C" omb_status omb_init(void) {
    clear_error();
    int dev = 0;
    OMB_HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    OMB_HIP_TRY(hipGetDeviceProperties(&prop, dev));
    std::string arch = prop.gcnArchName;
    if (arch.rfind("gfx950", 0) != 0) {
        set_error("omb_init: device is " + arch + "; this library is built for gfx950 only");
        return OMB_EUNSUPPORTED;
    }
    return sdot_scratch_setup(nullptr);  // omb_sdot_noalloc's scratch on this device
}

extern "C" omb_status omb_sgemm(int64_t M, int64_t N, int64_t K, float alpha, const float* A,
                                int64_t rsA, int64_t csA, const float* B, int64_t rsB,
                                int64_t csB, float beta, float* C, int64_t rsC, int64_t csC,
                                int32_t mode, omb_stream_t stream) {
    clear_error();
    hipStream_t s = (hipStream_t)stream;
    if (M < 0 || N < 0 || K < 0) {
        set_error("omb_sgemm: negative extent");
        return OMB_EINVAL;
    }
    if (mode != OMB_MODE_PLAIN_F32 && mode != OMB_MODE_SPLIT_BF16 &&
        mode != OMB_MODE_SPLIT_BF16_P3 && mode != OMB_MODE_SPLIT_FP16) {
        set_error("omb_sgemm: unknown mode " + std::to_string(mode));
        return OMB_EINVAL;
    }
    if (M == 0 || N == 0) return OMB_OK;
    if (!C) {
        set_error("omb_sgemm: C is null");
        return OMB_EINVAL;
    }
    bool c_row;
    int64_t ldc;
    OMB_TRY(resolve_layout("omb_sgemm C", M, N, rsC, csC, &c_row, &ldc));
    if (K > 0 && (!A || !B)) {
        set_error("omb_sgemm: A or B is null");
        return OMB_EINVAL;
    }
    bool a_row = true, b_row = true;
    int64_t lda = 1, ldb = 1;
    if (K > 0) {
        OMB_TRY(resolve_layout("omb_sgemm A", M, K, rsA, csA, &a_row, &lda));
        OMB_TRY(resolve_layout("omb_sgemm B", K, N, rsB, csB, &b_row, &ldb));
    }
    // Fold a column-major C: C^T (N x M, row-major, ld = ldc) = B^T A^T.
    int64_t m = M, n = N;
    const float* pa = A;
    const float* pb = B;
    int64_t ars = rsA, acs = csA, brs = rsB, bcs = csB;
    if (!c_row) {
        std::swap(m, n);
        pa = B;   // A' = B^T : n x k
        pb = A;   // B' = A^T : k x m
        ars = csB; acs = rsB;
        brs = csA; bcs = rsA;
        std::swap(lda, ldb);
        bool ta = a_row, tb = b_row;
        a_row = !tb;   // B^T is row-major iff B is column-major
        b_row = !ta;
    }
    if (K == 0) return launch_scale(m, n, beta, C, ldc, s);
    if (mode == OMB_MODE_PLAIN_F32)
        return launch_sgemm_f32(m, n, K, alpha, pa, lda, a_row, pb, ldb, b_row, beta, C, ldc, s);
    if (mode == OMB_MODE_SPLIT_FP16)
        return launch_sgemm_split16(m, n, K, alpha, pa, ars, acs, pb, brs, bcs, beta, C, ldc, s);
    return launch_sgemm_split(m, n, K, alpha, pa, ars, acs, pb, brs, bcs, beta, C, ldc,
                              mode == OMB_MODE_SPLIT_BF16_P3 ? 3 : 6, s);
}

extern "C" size_t omb_split_pack_bytes(int64_t R, int64_t K) {
    return (size_t)split_packed_elems(R, K) * sizeof(__bf16);
}

extern "C" omb_status omb_split_pack(int64_t R, int64_t K, const float* X, int64_t rs, int64_t cs,
                                     void* packed, omb_stream_t stream) {
    clear_error();
    if (R < 0 || K < 0 || rs < 1 || cs < 1 || (R > 0 && K > 0 && (!X || !packed))) {
        set_error("omb_split_pack: invalid argument");
        return OMB_EINVAL;
    }
    return launch_split_pack(R, K, X, rs, cs, static_cast<__bf16*>(packed), (hipStream_t)stream);
}

extern "C" omb_status omb_sgemm_packed(int64_t M, int64_t N, int64_t K, float alpha,
                                       const void* A_packed, const void* Bt_packed, float beta,
                                       float* C, int64_t rsC, int64_t csC, int32_t nprod,
                                       omb_stream_t stream) {
    clear_error();
    hipStream_t s = (hipStream_t)stream;
    if (M < 0 || N < 0 || K < 0 || (nprod != 3 && nprod != 6)) {
        set_error("omb_sgemm_packed: invalid extent or nprod");
        return OMB_EINVAL;
    }
    if (M == 0 || N == 0) return OMB_OK;
    if (!C || (K > 0 && (!A_packed || !Bt_packed))) {
        set_error("omb_sgemm_packed: null pointer");
        return OMB_EINVAL;
    }
    bool c_row;
    int64_t ldc;
    OMB_TRY(resolve_layout("omb_sgemm_packed C", M, N, rsC, csC, &c_row, &ldc));
    const __bf16* ap = static_cast<const __bf16*>(A_packed);
    const __bf16* bp = static_cast<const __bf16*>(Bt_packed);
    int64_t m = M, n = N;
    if (!c_row) {  // C^T = B A^T: the packed forms of A and B^T swap roles
        std::swap(m, n);
        std::swap(ap, bp);
    }
    if (K == 0) return launch_scale(m, n, beta, C, ldc, s);
    return launch_sgemm_packed(m, n, K, alpha, ap, bp, beta, C, ldc, nprod, s);
}

extern "C" omb_status omb_sgemv(int64_t M, int64_t N, float alpha, const float* A, int64_t rsA,
                                int64_t csA, const float* x, float beta, float* y,
                                omb_stream_t stream) {
    clear_error();
    hipStream_t s = (hipStream_t)stream;
    if (M < 0 || N < 0) {
        set_error("omb_sgemv: negative extent");
        return OMB_EINVAL;
    }
    if (M == 0) return OMB_OK;
    if (!y) {
        set_error("omb_sgemv: y is null");
        return OMB_EINVAL;
    }
    if (N == 0) return launch_scale(M, 1, beta, y, 1, s);
    if (!A || !x) {
        set_error("omb_sgemv: A or x is null");
        return OMB_EINVAL;
    }
    bool row;
    int64_t lda;
    OMB_TRY(resolve_layout("omb_sgemv A", M, N, rsA, csA, &row, &lda));
    if (row) return launch_sgemv_row(M, N, alpha, A, lda, x, beta, y, s);
    return launch_sgemv_col(M, N, alpha, A, lda, x, beta, y, s);
}

extern "C" omb_status omb_sdot(int64_t n, const float* x, const float* y, float* result,
                               omb_stream_t stream) {
    clear_error();
    if (n < 0 || !result || (n > 0 && (!x || !y))) {
        set_error("omb_sdot: invalid argument");
        return OMB_EINVAL;
    }
    return launch_sdot(n, x, y, result, (hipStream_t)stream);
}

extern "C" omb_status omb_sdot_noalloc(int64_t n, const float* x, const float* y, float* result,
                                       omb_stream_t stream) {
    clear_error();
    if (n < 0 || !result || (n > 0 && (!x || !y))) {
        set_error("omb_sdot_noalloc: invalid argument");
        return OMB_EINVAL;
    }
    return launch_sdot_noalloc(n, x, y, result, (hipStream_t)stream);
}

extern "C" omb_status omb_fill_uniform(float* dst, int64_t n, uint64_t seed, uint64_t stream_id,
                                       int64_t offset, omb_stream_t stream) {
    clear_error();
    if (n < 0 || offset < 0 || (n > 0 && !dst)) {
        set_error("omb_fill_uniform: invalid argument");
        return OMB_EINVAL;
    }
    return launch_fill_uniform(dst, n, seed, stream_id, offset, (hipStream_t)stream);
}

extern "C" int32_t omb_is_device_pointer(const void* p) {
    if (!p) return 0;
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged) ? 1 : 0;
}

extern "C" omb_status omb_device_alloc(void** p, size_t bytes) {
    clear_error();
    if (!p) {
        set_error("omb_device_alloc: null out pointer");
        return OMB_EINVAL;
    }
    *p = nullptr;
    if (bytes == 0) return OMB_OK;
    if (hipMalloc(p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        set_error("omb_device_alloc: hipMalloc failed");
        return OMB_ENOMEM;
    }
    return OMB_OK;
}

extern "C" omb_status omb_device_free(void* p) {
    clear_error();
    if (p) OMB_HIP_TRY(hipFree(p));
    return OMB_OK;
}

extern "C" omb_status omb_memcpy_h2d(void* dst, const void* src, size_t bytes,
                                     omb_stream_t stream) {
    clear_error();
    if (bytes == 0) return OMB_OK;
    OMB_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    return OMB_OK;
}

extern "C" omb_status omb_memcpy_d2h(void* dst, const void* src, size_t bytes,
                                     omb_stream_t stream) {
    clear_error();
    if (bytes == 0) return OMB_OK;
    OMB_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    return OMB_OK;
}

extern "C" omb_status omb_stream_synchronize(omb_stream_t stream) {
    clear_error();
    OMB_HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return OMB_OK;
}
