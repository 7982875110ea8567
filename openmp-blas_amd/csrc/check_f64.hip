// FP64 measurement kernels: the high-precision product of the same FP32
// inputs, and a deterministic error norm.  These run beside the FP32 path to
// state its error; they are not part of any FP32 result.  Also the input
// generator (uniform[-1,1) from a counter-based hash, bit-identical to the
// host generator) and the VALU issue-ceiling calibration kernel.
#include <algorithm>

#include "common.hpp"

namespace omb {
namespace {

constexpr int NT = 256;

// 64x64 tile of C64, 16-wide k slices in LDS, 4x4 per thread; FP64 FMA.
__global__ __launch_bounds__(NT) void dgemm_check_kernel(int64_t M, int64_t N, int64_t K,
                                                         const float* A, int64_t rsA, int64_t csA,
                                                         const float* B, int64_t rsB, int64_t csB,
                                                         const float* c0, int64_t rsC, int64_t csC,
                                                         double* C64, int64_t tiles_n) {
    __shared__ double As[16][64 + 1];
    __shared__ double Bs[16][64 + 1];
    const int tid = threadIdx.x;
    const int64_t tm = blockIdx.x / tiles_n, tn = blockIdx.x % tiles_n;
    const int64_t m0 = tm * 64, n0 = tn * 64;
    const int tx = tid & 15, ty = tid >> 4;
    double acc[4][4] = {};
    for (int64_t k0 = 0; k0 < K; k0 += 16) {
        for (int e = tid; e < 64 * 16; e += NT) {
            int r = e & 63, kk = e >> 6;
            int64_t m = m0 + r, n = n0 + r, k = k0 + kk;
            As[kk][r] = (m < M && k < K) ? (double)A[m * rsA + k * csA] : 0.0;
            Bs[kk][r] = (n < N && k < K) ? (double)B[k * rsB + n * csB] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a[i] = As[kk][ty + 16 * i];
                b[i] = Bs[kk][tx + 16 * i];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int64_t m = m0 + ty + 16 * i, n = n0 + tx + 16 * j;
            if (m < M && n < N) {
                double v = acc[i][j];
                if (c0) v += (double)c0[m * rsC + n * csC];
                C64[m * N + n] = v;
            }
        }
}

// One wavefront per row.
__global__ __launch_bounds__(NT) void dgemv_check_kernel(int64_t M, int64_t N, const float* A,
                                                         int64_t rsA, int64_t csA, const float* x,
                                                         const float* y0, double* y64) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    if (row >= M) return;
    double s = 0.0;
    for (int64_t j = lane; j < N; j += 64) s = fma((double)A[row * rsA + j * csA], (double)x[j], s);
    s = wave_sum(s);
    if (lane == 0) y64[row] = s + (y0 ? (double)y0[row] : 0.0);
}

__global__ __launch_bounds__(NT) void ddot_partial_kernel(int64_t n, const float* x, const float* y,
                                                          double* part) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
        s = fma((double)x[i], (double)y[i], s);
    __shared__ double red[NT / 64];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < NT / 64; ++w) t += red[w];
        part[blockIdx.x] = t;
    }
}

// Sums `count` groups of `width` doubles in order (width 1: sum; the error
// norm uses width 4 with max for slots 0-1 and sum for slots 2-3).
__global__ void combine_kernel(const double* part, int64_t count, int width, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t b = 0; b < count; ++b) {
        for (int w = 0; w < width; ++w) {
            double v = part[b * width + w];
            if (width == 4 && w < 2) acc[w] = fmax(acc[w], v);
            else acc[w] += v;
        }
    }
    for (int w = 0; w < width; ++w) out[w] = acc[w];
}

__global__ __launch_bounds__(NT) void errnorm_partial_kernel(int64_t M, int64_t N, const float* C,
                                                             int64_t rsC, int64_t csC,
                                                             const double* C64, double* part) {
    double dmax = 0.0, rmax = 0.0, d2 = 0.0, r2 = 0.0;
    const int64_t total = M * N;
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * NT) {
        int64_t m = i / N, n = i - m * N;
        double ref = C64[i];
        double d = (double)C[m * rsC + n * csC] - ref;
        dmax = fmax(dmax, fabs(d));
        rmax = fmax(rmax, fabs(ref));
        d2 = fma(d, d, d2);
        r2 = fma(ref, ref, r2);
    }
    dmax = wave_max(dmax);
    rmax = wave_max(rmax);
    d2 = wave_sum(d2);
    r2 = wave_sum(r2);
    __shared__ double red[NT / 64][4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[w][0] = dmax;
        red[w][1] = rmax;
        red[w][2] = d2;
        red[w][3] = r2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double o[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < NT / 64; ++i) {
            o[0] = fmax(o[0], red[i][0]);
            o[1] = fmax(o[1], red[i][1]);
            o[2] += red[i][2];
            o[3] += red[i][3];
        }
        for (int i = 0; i < 4; ++i) part[blockIdx.x * 4 + i] = o[i];
    }
}

__global__ __launch_bounds__(NT) void fill_uniform_kernel(float* dst, int64_t n, uint64_t key,
                                                          int64_t offset) {
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
        dst[i] = rng_uniform(key, (uint64_t)(i + offset));
}

unsigned grid_for(int64_t n, int64_t cap) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, NT), cap));
}

// VALU issue ceiling: 64 independent FP32 multiply-add chains per thread, in
// the two-VGPR-operand form (v_fmac_f32) that the plain SGEMM kernel's inner
// loop consists of.  Nothing but the chains runs in the loop, so the rate it
// reaches is the most FP32 FMA work the CUs retire at the occupancy of the
// launch.  The single store is taken only when the sum is exactly 0.125 (the
// inputs make that unreachable) and goes to sink[0].
__global__ __launch_bounds__(NT) void fmac_ceiling_kernel(int64_t iters, float* sink) {
    constexpr int CH = 64;
    float acc[CH];
    const float x0 = 1.0f + threadIdx.x * 0x1p-20f, x1 = x0 + 0x1p-12f;
    const float y0 = 0.5f - threadIdx.x * 0x1p-21f, y1 = y0 - 0x1p-13f;
#pragma unroll
    for (int c = 0; c < CH; ++c) acc[c] = threadIdx.x + c;
    for (int64_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int c = 0; c < CH; ++c) acc[c] = fmaf((c & 1) ? x1 : x0, (c & 2) ? y1 : y0, acc[c]);
    }
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < CH; ++c) s += acc[c];
    if (s == 0.125f) sink[0] = s;
}

// MFMA throughput ceiling: nothing but v_mfma_f32_32x32x16_bf16, 8
// independent accumulators per wavefront, operands already in registers (no
// memory traffic in the loop).  With random operands every MFMA multiplies
// other data (8 A and 8 B fragments, 64 pairings per round), as the GEMM's
// MFMAs do; with zero operands nothing toggles in the datapath.  Their rates
// bound what any BF16 MFMA kernel sustains on this GPU with random data and
// without that data-dependent power.  The single store is taken only when the
// sum is exactly 0.125 (unreachable) and goes to sink[0].
typedef __bf16 calib_bf16x8 __attribute__((ext_vector_type(8)));
typedef float calib_f32x16 __attribute__((ext_vector_type(16)));

// `scale` (1 or 0) multiplies the random values: a runtime factor, so the
// compiler cannot fold the zero-operand MFMAs away.
__global__ __launch_bounds__(NT) void mfma_ceiling_kernel(int64_t iters, float scale, float* sink) {
    calib_bf16x8 a[8], b[8];
    const uint64_t key = rng_key(0x5eedULL, blockIdx.x);
#pragma unroll
    for (int v = 0; v < 8; ++v)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint64_t i = ((uint64_t)threadIdx.x * 8 + v) * 16 + e;
            a[v][e] = (__bf16)(scale * rng_uniform(key, i));
            b[v][e] = (__bf16)(scale * rng_uniform(key, i + 8));
        }
    calib_f32x16 acc[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = calib_f32x16{};
    for (int64_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j)
                acc[(i + j) & 7] =
                    __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[(i + j) & 7], 0, 0, 0);
    }
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
        for (int r = 0; r < 16; ++r) s += acc[c][r];
    if (s == 0.125f) sink[0] = s;
}

}  // namespace

omb_status launch_fill_uniform(float* dst, int64_t n, uint64_t seed, uint64_t stream_id,
                               int64_t offset, hipStream_t s) {
    if (n <= 0) return OMB_OK;
    hipLaunchKernelGGL(fill_uniform_kernel, dim3(grid_for(n, 16384)), dim3(NT), 0, s, dst, n,
                       rng_key(seed, stream_id), offset);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

}  // namespace omb

using namespace omb;

extern "C" omb_status omb_check_dgemm(int64_t M, int64_t N, int64_t K, const float* A,
                                      int64_t rsA, int64_t csA, const float* B, int64_t rsB,
                                      int64_t csB, const float* c0, int64_t rsC, int64_t csC,
                                      double* C64, omb_stream_t stream) {
    clear_error();
    if (M < 0 || N < 0 || K < 0 || !C64 || (K > 0 && (!A || !B))) {
        set_error("omb_check_dgemm: invalid argument");
        return OMB_EINVAL;
    }
    if (M == 0 || N == 0) return OMB_OK;
    const int64_t tiles_n = ceil_div(N, 64), tiles = ceil_div(M, 64) * tiles_n;
    hipLaunchKernelGGL(dgemm_check_kernel, dim3((unsigned)tiles), dim3(NT), 0,
                       (hipStream_t)stream, M, N, K, A, rsA, csA, B, rsB, csB, c0, rsC, csC, C64,
                       tiles_n);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

extern "C" omb_status omb_check_dgemv(int64_t M, int64_t N, const float* A, int64_t rsA,
                                      int64_t csA, const float* x, const float* y0, double* y64,
                                      omb_stream_t stream) {
    clear_error();
    if (M < 0 || N < 0 || !y64 || (N > 0 && (!A || !x))) {
        set_error("omb_check_dgemv: invalid argument");
        return OMB_EINVAL;
    }
    if (M == 0) return OMB_OK;
    hipLaunchKernelGGL(dgemv_check_kernel, dim3((unsigned)ceil_div(M, NT / 64)), dim3(NT), 0,
                       (hipStream_t)stream, M, N, A, rsA, csA, x, y0, y64);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

extern "C" omb_status omb_check_ddot(int64_t n, const float* x, const float* y, double* r64,
                                     omb_stream_t stream) {
    clear_error();
    if (n < 0 || !r64 || (n > 0 && (!x || !y))) {
        set_error("omb_check_ddot: invalid argument");
        return OMB_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    const unsigned blocks = grid_for(n, 1024);
    Workspace* ws = nullptr;
    OMB_TRY(get_workspace(s, blocks * sizeof(double), &ws));
    double* part = reinterpret_cast<double*>(ws->buf);
    hipLaunchKernelGGL(ddot_partial_kernel, dim3(blocks), dim3(NT), 0, s, n, x, y, part);
    hipLaunchKernelGGL(combine_kernel, dim3(1), dim3(1), 0, s, part, (int64_t)blocks, 1, r64);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

extern "C" omb_status omb_check_errnorm(int64_t M, int64_t N, const float* C, int64_t rsC,
                                        int64_t csC, const double* C64, double* out,
                                        omb_stream_t stream) {
    clear_error();
    if (M < 0 || N < 0 || !out || (M * N > 0 && (!C || !C64))) {
        set_error("omb_check_errnorm: invalid argument");
        return OMB_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    const unsigned blocks = grid_for(M * N, 1024);
    Workspace* ws = nullptr;
    OMB_TRY(get_workspace(s, (size_t)blocks * 4 * sizeof(double), &ws));
    double* part = reinterpret_cast<double*>(ws->buf);
    hipLaunchKernelGGL(errnorm_partial_kernel, dim3(blocks), dim3(NT), 0, s, M, N, C, rsC, csC,
                       C64, part);
    hipLaunchKernelGGL(combine_kernel, dim3(1), dim3(1), 0, s, part, (int64_t)blocks, 4, out);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

extern "C" omb_status omb_calib_mfma(int64_t blocks, int64_t iters, int32_t random, float* sink,
                                     omb_stream_t stream) {
    clear_error();
    if (blocks < 1 || blocks > 0x7fffffff || iters < 1 || !sink) {
        set_error("omb_calib_mfma: invalid argument");
        return OMB_EINVAL;
    }
    hipLaunchKernelGGL(mfma_ceiling_kernel, dim3((unsigned)blocks), dim3(NT), 0,
                       (hipStream_t)stream, iters, random ? 1.f : 0.f, sink);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

extern "C" omb_status omb_calib_fmac(int64_t blocks, int64_t iters, float* sink,
                                     omb_stream_t stream) {
    clear_error();
    if (blocks < 1 || blocks > 0x7fffffff || iters < 1 || !sink) {
        set_error("omb_calib_fmac: invalid argument");
        return OMB_EINVAL;
    }
    hipLaunchKernelGGL(fmac_ceiling_kernel, dim3((unsigned)blocks), dim3(NT), 0,
                       (hipStream_t)stream, iters, sink);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}
