// SGEMV for gfx950: y = alpha*A*x + beta*y, both storage orders of A.
//
// Row-major A (counterpart of mtv_helper last_order + simd_loop<INNER>,
// include/mtv.hpp:72-98, include/simd_loop.hpp:21-35): one wavefront owns
// ROWS consecutive rows; lanes stream 128-bit loads of each row, share every
// x float4 across the ROWS rows, and finish with a butterfly lane sum.
//
// Column-major A (counterpart of mtv_helper first_order + simd_loop<MTV,16>,
// include/mtv.hpp:15-70, include/simd_loop.hpp:52-78): every thread owns 4
// consecutive rows (one 128-bit load per column); the columns are split into
// chunks across workgroups, each chunk writes a partial y, and a second kernel
// adds the partials in chunk order.  The result depends only on (M, N),
// never on scheduling.
#include <algorithm>

#include "gemv_row.hpp"

namespace omb {
namespace {

constexpr int NT = 256;

template <bool VEC>
__global__ __launch_bounds__(NT) void sgemv_row_kernel(int64_t M, int64_t N, float alpha,
                                                       const float* __restrict__ A, int64_t lda,
                                                       const float* __restrict__ x, float beta,
                                                       float* __restrict__ y) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    const int64_t r0 = wave * kGemvRows;
    if (r0 >= M) return;
    const int nr = (int)std::min<int64_t>(kGemvRows, M - r0);
    gemv_rows_wave<VEC>(r0, nr, N, alpha, A, lda, x, beta, y, lane);
}

constexpr int CRPT = 16;                 // rows per lane (4 float4)
constexpr int COL_ROWS = 64 * CRPT;      // rows per workgroup (1024): 4 KB of each column
constexpr int COL_UNROLL = 4;            // columns in flight per wavefront

// Column-major A: lane l of every wavefront owns rows r0 + 256q + 4l + e
// (q, e < 4), so each column is read as 4 KB contiguous per wavefront; the 4
// wavefronts of a workgroup take interleaved columns of the chunk and their
// partial sums are added in LDS in wavefront order.
template <bool VEC>
__global__ __launch_bounds__(NT) void sgemv_col_kernel(int64_t M, int64_t N, float alpha,
                                                       const float* __restrict__ A, int64_t lda,
                                                       const float* __restrict__ x, float beta,
                                                       float* __restrict__ y, float* partial,
                                                       int64_t mpad, int64_t cols_per_chunk,
                                                       int nchunks) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t rb = blockIdx.x, ch = blockIdx.y;
    const int64_t r0 = rb * COL_ROWS;
    const int64_t c0 = ch * cols_per_chunk;
    const int64_t c1 = std::min<int64_t>(N, c0 + cols_per_chunk);
    float acc[CRPT];
#pragma unroll
    for (int i = 0; i < CRPT; ++i) acc[i] = 0.f;
    if (VEC) {  // M % 4 == 0, lda % 4 == 0, A 16-byte aligned
        bool ok[4];
        const float* ap[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t r = r0 + 256 * q + 4 * lane;
            ok[q] = r < M;
            ap[q] = A + (ok[q] ? r : 0);
        }
        int64_t j = c0 + wave;
        for (; j + 4 * (COL_UNROLL - 1) < c1; j += 4 * COL_UNROLL) {
            float4 av[COL_UNROLL][4];
            float xv[COL_UNROLL];
#pragma unroll
            for (int u = 0; u < COL_UNROLL; ++u) {
                const int64_t col = (j + 4 * u) * lda;
                xv[u] = x[j + 4 * u];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    av[u][q] = ok[q] ? nt_load4(reinterpret_cast<const float4*>(ap[q] + col))
                                     : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < COL_UNROLL; ++u)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    acc[4 * q + 0] = fmaf(av[u][q].x, xv[u], acc[4 * q + 0]);
                    acc[4 * q + 1] = fmaf(av[u][q].y, xv[u], acc[4 * q + 1]);
                    acc[4 * q + 2] = fmaf(av[u][q].z, xv[u], acc[4 * q + 2]);
                    acc[4 * q + 3] = fmaf(av[u][q].w, xv[u], acc[4 * q + 3]);
                }
        }
        for (; j < c1; j += 4) {
            const float xv = x[j];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (!ok[q]) continue;
                float4 av = *reinterpret_cast<const float4*>(ap[q] + j * lda);
                acc[4 * q + 0] = fmaf(av.x, xv, acc[4 * q + 0]);
                acc[4 * q + 1] = fmaf(av.y, xv, acc[4 * q + 1]);
                acc[4 * q + 2] = fmaf(av.z, xv, acc[4 * q + 2]);
                acc[4 * q + 3] = fmaf(av.w, xv, acc[4 * q + 3]);
            }
        }
    } else {
        for (int64_t j = c0 + wave; j < c1; j += 4) {
            const float xv = x[j];
            const float* col = A + j * lda;
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int64_t r = r0 + 256 * q + 4 * lane + e;
                    if (r < M) acc[4 * q + e] = fmaf(col[r], xv, acc[4 * q + e]);
                }
        }
    }
    // Wavefront partials -> LDS, summed in wavefront order by 256 threads
    // (thread t: rows 4t .. 4t+3 of the block).
    __shared__ float red[4][COL_ROWS];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) red[wave][256 * q + 4 * lane + e] = acc[4 * q + e];
    __syncthreads();
    const int64_t r = r0 + 4 * threadIdx.x;
    float s[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int i = 4 * threadIdx.x + e;
        s[e] = ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
    }
    const int nv = r < M ? (int)std::min<int64_t>(4, M - r) : 0;
    if (nchunks == 1) {
        for (int e = 0; e < nv; ++e) {
            float out = alpha * s[e];
            if (beta != 0.f) out += beta * y[r + e];
            y[r + e] = out;
        }
        return;
    }
    // Partial sums of this column chunk; sgemv_col_combine_kernel adds the
    // chunks in order 0..nchunks-1.
    if (nv > 0) {
        float* pp = partial + ch * mpad + r;
        for (int e = 0; e < 4; ++e) pp[e] = s[e];
    }
}

// y[r] = alpha * sum_c partial[c][r] + beta * y[r], chunks added in order.
__global__ __launch_bounds__(NT) void sgemv_col_combine_kernel(int64_t M, float alpha,
                                                               const float* __restrict__ partial,
                                                               int64_t mpad, int nchunks,
                                                               float beta, float* __restrict__ y) {
    const int64_t r = ((int64_t)blockIdx.x * NT + threadIdx.x) * 4;
    if (r >= M) return;
    const int nv = (int)std::min<int64_t>(4, M - r);
    float t[4] = {0.f, 0.f, 0.f, 0.f};
    for (int c = 0; c < nchunks; ++c) {
        const float* pp = partial + (int64_t)c * mpad + r;
#pragma unroll
        for (int e = 0; e < 4; ++e) t[e] += pp[e];
    }
    for (int e = 0; e < nv; ++e) {
        float out = alpha * t[e];
        if (beta != 0.f) out += beta * y[r + e];
        y[r + e] = out;
    }
}

}  // namespace

omb_status launch_sgemv_row(int64_t M, int64_t N, float alpha, const float* A, int64_t lda,
                            const float* x, float beta, float* y, hipStream_t s) {
    if (M <= 0) return OMB_OK;
    const int64_t waves = ceil_div(M, kGemvRows);
    const int64_t blocks = ceil_div(waves, NT / 64);
    if (blocks > 0x7fffffffLL) {
        set_error("sgemv: too many rows");
        return OMB_EUNSUPPORTED;
    }
    if (gemv_row_vec(A, lda, x))
        hipLaunchKernelGGL(sgemv_row_kernel<true>, dim3((unsigned)blocks), dim3(NT), 0, s, M, N,
                           alpha, A, lda, x, beta, y);
    else
        hipLaunchKernelGGL(sgemv_row_kernel<false>, dim3((unsigned)blocks), dim3(NT), 0, s, M, N,
                           alpha, A, lda, x, beta, y);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

omb_status launch_sgemv_col(int64_t M, int64_t N, float alpha, const float* A, int64_t lda,
                            const float* x, float beta, float* y, hipStream_t s) {
    if (M <= 0) return OMB_OK;
    const int64_t nrb = ceil_div(M, COL_ROWS);
    // Enough workgroups to cover the chip (~1024), at least 64 columns each.
    int64_t nchunks = std::max<int64_t>(1, std::min<int64_t>(ceil_div(1024, nrb), ceil_div(N, 64)));
    // (each of the 4 wavefronts then takes every 4th column of its chunk)
    if (N == 0) nchunks = 1;
    const int64_t cpc = N > 0 ? ceil_div(N, nchunks) : 0;
    nchunks = N > 0 ? ceil_div(N, cpc) : 1;
    if (nrb > 0x7fffffffLL) {
        set_error("sgemv: too many row blocks");
        return OMB_EUNSUPPORTED;
    }
    const int64_t mpad = round_up(M, 4);
    Workspace* ws = nullptr;
    OMB_TRY(get_workspace(s, nchunks > 1 ? (size_t)(nchunks * mpad) * sizeof(float) : 0, &ws));
    const bool vec = ((reinterpret_cast<uintptr_t>(A) & 15) == 0) && (lda % 4 == 0) && (M % 4 == 0);
    dim3 grid((unsigned)nrb, (unsigned)nchunks);
    float* part = reinterpret_cast<float*>(ws->buf);
    if (vec)
        hipLaunchKernelGGL(sgemv_col_kernel<true>, grid, dim3(NT), 0, s, M, N, alpha, A, lda, x,
                           beta, y, part, mpad, cpc, (int)nchunks);
    else
        hipLaunchKernelGGL(sgemv_col_kernel<false>, grid, dim3(NT), 0, s, M, N, alpha, A, lda, x,
                           beta, y, part, mpad, cpc, (int)nchunks);
    if (nchunks > 1)
        hipLaunchKernelGGL(sgemv_col_combine_kernel, dim3((unsigned)ceil_div(mpad / 4, NT)),
                           dim3(NT), 0, s, M, alpha, part, mpad, (int)nchunks, beta, y);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

}  // namespace omb
