// SDOT for gfx950: *result = sum_i x[i]*y[i].
//
// Counterpart of dot_prod_helper (include/dot.hpp:143-174), whose three
// regimes (serial simd, parallel-for reduction, blocked parallel reduction)
// all reduce FP32 lane partials into per-thread and then per-team sums.
// Here: a workgroup-interleaved grid-stride loop of 128-bit loads (4 float4
// of x and of y in flight per lane) into FP32 lane sums, a fixed butterfly
// per wavefront, the four wavefront sums in order, one partial per
// workgroup; a second one-workgroup kernel adds the partials in a fixed
// order -- in FP64 when n >= 2^26, in FP32 below.  (A second launch instead of a last-workgroup ticket: on this
// GPU an agent-scope fence writes back and invalidates the XCD's L2, which
// costs more than the launch.)  The grid depends on n only and the loads do
// not depend on alignment, so the result is bitwise reproducible for any
// placement of the operands.
#include <algorithm>
#include <mutex>

#include "common.hpp"

namespace omb {
namespace {

constexpr int NT = 256;
constexpr int UNROLL = 4;
constexpr int kMaxBlocks = 1024;
constexpr int64_t kF64Combine = int64_t(1) << 26;

}  // namespace
}  // namespace omb

// Partial sums of omb_sdot_noalloc: one double per workgroup, at most
// kMaxBlocks (8 KiB), in device memory the library owns for the life of the
// process (each device holds its own copy).  Host code never passes this
// name to a kernel: it stands for the host-side shadow, not the device
// address, which set-up obtains with hipGetSymbolAddress per device.
static __device__ __align__(256) double scratch[omb::kMaxBlocks];

namespace omb {
namespace {

__global__ __launch_bounds__(NT) void sdot_kernel(int64_t n, const float* __restrict__ x,
                                                  const float* __restrict__ y, float* result,
                                                  double* partial) {
    const int64_t gtid = (int64_t)blockIdx.x * NT + threadIdx.x;
    const int64_t nthr = (int64_t)gridDim.x * NT;
    float acc = 0.f;
    // Workgroup-interleaved grid stride over 4-float groups (128-bit loads at
    // any alignment): each step a workgroup reads 16 KB of x and of y
    // contiguously (UNROLL groups per thread, 4 KB apart), and the grid
    // sweeps HBM as one front (88% of peak measured at 2^28, against 75%
    // with one contiguous segment per wavefront).
    const int64_t n4 = n >> 2;
    const int64_t step = (int64_t)gridDim.x * NT * UNROLL;
    int64_t i = (int64_t)blockIdx.x * NT * UNROLL + threadIdx.x;
    for (; i + NT * (UNROLL - 1) < n4; i += step) {
        float4 a[UNROLL], b[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            a[u] = ld4u_nt(x + 4 * (i + NT * u));
            b[u] = ld4u_nt(y + 4 * (i + NT * u));
        }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) acc = dot4(a[u], b[u], acc);
    }
    // The last, partial step of this workgroup (its next step starts past n4).
    for (; i < n4; i += NT) acc = dot4(ld4u(x + 4 * i), ld4u(y + 4 * i), acc);
    const int64_t tail = n4 << 2;
    for (int64_t i = tail + gtid; i < n; i += nthr) acc = fmaf(x[i], y[i], acc);

    // Workgroup sum: wavefront butterfly, then the 4 wave sums in order.
    __shared__ float s_wave[NT / 64];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = s_wave[0];
        for (int w = 1; w < NT / 64; ++w) s += s_wave[w];
        if (gridDim.x == 1) *result = s;
        else partial[blockIdx.x] = (double)s;
    }
}

// One workgroup adds the per-workgroup partials: thread t takes partials
// t, t+NT, ... in order, then a fixed tree -- in FP64 or in FP32.
__global__ __launch_bounds__(NT) void sdot_combine_kernel(const double* __restrict__ partial,
                                                          int count, int f64, float* result) {
    __shared__ double s_red[NT];
    if (f64) {
        double s = 0.0;
        for (int b = threadIdx.x; b < count; b += NT) s += partial[b];
        s_red[threadIdx.x] = s;
    } else {
        float s = 0.f;
        for (int b = threadIdx.x; b < count; b += NT) s += (float)partial[b];
        s_red[threadIdx.x] = (double)s;
    }
    __syncthreads();
    for (int w = NT / 2; w >= 1; w >>= 1) {
        if (threadIdx.x < w) {
            if (f64) s_red[threadIdx.x] += s_red[threadIdx.x + w];
            else s_red[threadIdx.x] = (double)((float)s_red[threadIdx.x] + (float)s_red[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *result = (float)s_red[0];
}

__global__ void zero_scalar_kernel(float* r) { *r = 0.f; }

}  // namespace

int64_t sdot_blocks(int64_t n) {
    // One full step (NT * UNROLL float4 of each operand) per workgroup at
    // least; 1024 workgroups (4 per CU) keep enough loads in flight.
    return std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, ceil_div(n, NT * 4 * UNROLL)));
}

omb_status launch_sdot_ws(int64_t n, const float* x, const float* y, float* result,
                          double* part, hipStream_t s) {
    if (n <= 0) {
        hipLaunchKernelGGL(zero_scalar_kernel, dim3(1), dim3(1), 0, s, result);
        OMB_HIP_TRY(hipGetLastError());
        return OMB_OK;
    }
    const int64_t blocks = sdot_blocks(n);
    const int f64 = n >= kF64Combine ? 1 : 0;
    hipLaunchKernelGGL(sdot_kernel, dim3((unsigned)blocks), dim3(NT), 0, s, n, x, y, result, part);
    if (blocks > 1)
        hipLaunchKernelGGL(sdot_combine_kernel, dim3(1), dim3(NT), 0, s, part, (int)blocks, f64,
                           result);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

// Device address of `scratch` on each device, filled once per device with
// that device current (omb_init, or the first omb_sdot_noalloc there).
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};
}  // namespace

omb_status sdot_scratch_setup(double** out) {
    int dev = 0;
    OMB_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("sdot scratch: device ordinal " + std::to_string(dev) + " out of range");
        return OMB_EUNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        OMB_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("sdot scratch: hipGetSymbolAddress returned no address");
            return OMB_EDEVICE;
        }
        g_scratch[dev] = static_cast<double*>(p);
    }
    if (out) *out = g_scratch[dev];
    return OMB_OK;
}

omb_status launch_sdot_noalloc(int64_t n, const float* x, const float* y, float* result,
                               hipStream_t s) {
    double* part = nullptr;
    OMB_TRY(sdot_scratch_setup(&part));
    const int64_t blocks = n > 0 ? sdot_blocks(n) : 0;
    static_assert(sizeof(scratch) == kMaxBlocks * sizeof(double), "one double per workgroup");
    if (blocks > kMaxBlocks) {
        set_error("omb_sdot_noalloc: " + std::to_string(blocks) +
                  " partial sums exceed the static scratch of " + std::to_string(kMaxBlocks));
        return OMB_EUNSUPPORTED;
    }
    return launch_sdot_ws(n, x, y, result, part, s);
}

omb_status launch_sdot(int64_t n, const float* x, const float* y, float* result, hipStream_t s) {
    Workspace* ws = nullptr;
    OMB_TRY(get_workspace(s, n > 0 ? (size_t)sdot_blocks(n) * sizeof(double) : 0, &ws));
    return launch_sdot_ws(n, x, y, result, reinterpret_cast<double*>(ws->buf), s);
}

}  // namespace omb
