// SGEMV for gfx950: y = alpha*A*x + beta*y, both storage orders of A.
//
// Row-major A (counterpart of mtv_helper last_order + simd_loop<INNER>,
// include/mtv.hpp:72-98, include/simd_loop.hpp:21-35): one wavefront owns
// ROWS consecutive rows; lanes stream 128-bit loads of each row, share every
// x float4 across the ROWS rows, and finish with a butterfly lane sum.
//
// Column-major A (counterpart of mtv_helper first_order + simd_loop<MTV,16>,
// include/mtv.hpp:15-70, include/simd_loop.hpp:52-78): every thread owns 4
// consecutive rows (one 128-bit load per column).  Tall matrices of whole
// 64-row blocks and 256-column rounds run in one pass of 64-row workgroups;
// otherwise the columns are split into chunks across workgroups, each chunk
// writes a partial y, and a second kernel adds the partials in chunk order.
// The result depends only on (M, N) and the device's CU count, never on
// scheduling.
#include <algorithm>

#include "gemv_row.hpp"

namespace omb {
namespace {

constexpr int NT = 256;

// Rows per wavefront and 128-bit loads in flight per row: 16 rows (one
// workgroup per CU at 16384 rows, measured 86% of HBM peak) for tall
// matrices, fewer rows per wavefront when that would leave CUs idle.  A
// row's order of accumulation is the same for every choice.
template <int ROWS, int UNROLL, bool OFF32 = false>
__global__ __launch_bounds__(NT) void sgemv_row_kernel(int64_t M, int64_t N, float alpha,
                                                       const float* __restrict__ A, int64_t lda,
                                                       const float* __restrict__ x, float beta,
                                                       float* __restrict__ y) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (NT / 64);
    // Grid stride over groups of ROWS rows (a grid that covers every group
    // once measured 4% slower at 16384^2 than this loop over the same grid).
    // The wavefront index is uniform: keep it (and the row pointers) scalar.
    for (int64_t w = (int64_t)blockIdx.x * (NT / 64) +
                     __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
         w * ROWS < M; w += waves) {
        const int64_t r0 = w * ROWS;
        const int nr = (int)std::min<int64_t>(ROWS, M - r0);
        gemv_rows_wave<ROWS, UNROLL, true, OFF32>(r0, nr, N, alpha, A, lda, x, beta, y, lane);
    }
}

constexpr int COL_ROWS = 256;     // rows per workgroup: 4 per lane, one float4 per column
constexpr int COL_UNROLL = 16;    // columns in flight per wavefront (16 KB per wavefront)

// One lane's sums over columns c0+wave, c0+wave+4, ... < c1 of its 4 rows at
// ap: FULL lanes (all 4 rows inside A) use 128-bit loads at any alignment,
// the others guarded scalars (rows past `valid` read as 0).  Four
// accumulator sets (column j + 4u goes to set u % 4) keep each FP32 running
// sum to a quarter of the chunk; both forms add in the same order.
template <bool FULL>
__device__ __forceinline__ void col_lane_sums(const float* __restrict__ ap, int64_t lda,
                                              const float* __restrict__ x, int64_t j, int64_t c1,
                                              int valid, float (&acc)[4][4]) {
    auto load = [&](int64_t col) -> float4 {
        const float* c = ap + col * lda;
        if (FULL) return ld4u_nt(c);
        return make_float4(valid > 0 ? c[0] : 0.f, valid > 1 ? c[1] : 0.f,
                           valid > 2 ? c[2] : 0.f, 0.f);
    };
    for (; j + 4 * (COL_UNROLL - 1) < c1; j += 4 * COL_UNROLL) {
        float4 av[COL_UNROLL];
        float xv[COL_UNROLL];
#pragma unroll
        for (int u = 0; u < COL_UNROLL; ++u) {
            xv[u] = x[j + 4 * u];
            av[u] = load(j + 4 * u);
        }
#pragma unroll
        for (int u = 0; u < COL_UNROLL; ++u) {
            float* a = acc[u & 3];
            a[0] = fmaf(av[u].x, xv[u], a[0]);
            a[1] = fmaf(av[u].y, xv[u], a[1]);
            a[2] = fmaf(av[u].z, xv[u], a[2]);
            a[3] = fmaf(av[u].w, xv[u], a[3]);
        }
    }
    for (int u = 0; j < c1; j += 4, ++u) {
        const float xv = x[j];
        const float4 av = load(j);
        float* a = acc[u & 3];
        a[0] = fmaf(av.x, xv, a[0]);
        a[1] = fmaf(av.y, xv, a[1]);
        a[2] = fmaf(av.z, xv, a[2]);
        a[3] = fmaf(av.w, xv, a[3]);
    }
}

// Column-major A: lane l of every wavefront owns rows r0 + 4l .. r0 + 4l + 3,
// so each column is read as 1 KB contiguous per wavefront; the 4 wavefronts
// of a workgroup take interleaved columns of the chunk and their partial
// sums are added in LDS in wavefront order.  The launcher sizes the grid to
// about one workgroup per CU: fewer, longer column streams keep HBM pages
// open (measured 88% of peak at 16384^2 against 68% with 1024 workgroups).
__global__ __launch_bounds__(NT) void sgemv_col_kernel(int64_t M, int64_t N, float alpha,
                                                       const float* __restrict__ A, int64_t lda,
                                                       const float* __restrict__ x, float beta,
                                                       float* __restrict__ y, float* partial,
                                                       int64_t mpad, int64_t cols_per_chunk,
                                                       int nchunks) {
    // the wavefront index is uniform: its column indices (and x) stay scalar
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t r0 = (int64_t)blockIdx.x * COL_ROWS, ch = blockIdx.y;
    const int64_t c0 = ch * cols_per_chunk;
    const int64_t c1 = std::min<int64_t>(N, c0 + cols_per_chunk);
    const int64_t r = r0 + 4 * lane;  // this lane's 4 rows
    float acc[4][4] = {};
    if (r + 3 < M)
        col_lane_sums<true>(A + r, lda, x, c0 + wave, c1, 4, acc);
    else if (r < M)  // the lane holding the last 1-3 rows
        col_lane_sums<false>(A + r, lda, x, c0 + wave, c1, (int)(M - r), acc);
    // Wavefront partials -> LDS, summed in wavefront order (thread t: row r0 + t).
    __shared__ float red[4][COL_ROWS];
#pragma unroll
    for (int e = 0; e < 4; ++e)
        red[wave][4 * lane + e] = ((acc[0][e] + acc[1][e]) + acc[2][e]) + acc[3][e];
    __syncthreads();
    const int t = threadIdx.x;
    const int64_t row = r0 + t;
    if (row >= M) return;
    const float sum = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
    if (nchunks == 1) {
        float out = alpha * sum;
        if (beta != 0.f) out += beta * y[row];
        y[row] = out;
        return;
    }
    // Partial sum of this column chunk; sgemv_col_combine_kernel adds the
    // chunks in order 0..nchunks-1.
    partial[ch * mpad + row] = sum;
}

// Column-major A in one pass, for M a multiple of 64 with at least one
// 64-row block per CU and N a multiple of 256 (no partial round, no row
// guard: both measurably slow the loop).  Lane l of a wavefront owns rows
// r0 + 4*(l % 16) .. +3 of column group l / 16, so one load instruction reads
// 4 columns of 64 rows (256 bytes each); wavefront w takes column groups
// 4w .. 4w+3 of every 16 columns, 16 rounds in flight.  Each lane keeps
// four accumulator sets (its u-th column of a round goes to set u % 4),
// added in order; then the 16 lane sums of a row are added in LDS in
// (wavefront, group) order.  Measured 0.91-0.92 of HBM peak at 16384^2.
constexpr int R64_UNROLL = 16;
__global__ __launch_bounds__(NT) void sgemv_col_rows64_kernel(int64_t M, int64_t N, float alpha,
                                                              const float* __restrict__ A,
                                                              int64_t lda,
                                                              const float* __restrict__ x,
                                                              float beta,
                                                              float* __restrict__ y) {
    constexpr int STEP = 16;  // columns per round of the 4 wavefronts
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rq = lane & 15, cg = lane >> 4;
    const int64_t r = (int64_t)blockIdx.x * 64 + 4 * rq;
    const float* ap = A + r;
    float acc[4][4] = {};
    for (int64_t j = wave * 4 + cg; j + STEP * (R64_UNROLL - 1) < N; j += STEP * R64_UNROLL) {
        float4 av[R64_UNROLL];
        float xv[R64_UNROLL];
#pragma unroll
        for (int u = 0; u < R64_UNROLL; ++u) {
            xv[u] = x[j + STEP * u];
            av[u] = ld4u_nt(ap + (j + STEP * u) * lda);
        }
#pragma unroll
        for (int u = 0; u < R64_UNROLL; ++u) {
            float* a = acc[u & 3];
            a[0] = fmaf(av[u].x, xv[u], a[0]);
            a[1] = fmaf(av[u].y, xv[u], a[1]);
            a[2] = fmaf(av[u].z, xv[u], a[2]);
            a[3] = fmaf(av[u].w, xv[u], a[3]);
        }
    }
    __shared__ float red[4][4][64];
#pragma unroll
    for (int e = 0; e < 4; ++e)
        red[wave][cg][4 * rq + e] = ((acc[0][e] + acc[1][e]) + acc[2][e]) + acc[3][e];
    __syncthreads();
    if (threadIdx.x < 64) {
        const int t = threadIdx.x;
        float sum = red[0][0][t];
#pragma unroll
        for (int w = 0; w < 4; ++w)
#pragma unroll
            for (int g = 0; g < 4; ++g)
                if (w != 0 || g != 0) sum += red[w][g][t];
        const int64_t row = (int64_t)blockIdx.x * 64 + t;
        float out = alpha * sum;
        if (beta != 0.f) out += beta * y[row];
        y[row] = out;
    }
}

// y[r] = alpha * sum_c partial[c][r] + beta * y[r], chunks added in order
// (one row per thread).
__global__ __launch_bounds__(NT) void sgemv_col_combine_kernel(int64_t M, float alpha,
                                                               const float* __restrict__ partial,
                                                               int64_t mpad, int nchunks,
                                                               float beta, float* __restrict__ y) {
    const int64_t r = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (r >= M) return;
    float t = 0.f;
    for (int c = 0; c < nchunks; ++c) t += partial[(int64_t)c * mpad + r];
    float out = alpha * t;
    if (beta != 0.f) out += beta * y[r];
    y[r] = out;
}

}  // namespace

omb_status launch_sgemv_row(int64_t M, int64_t N, float alpha, const float* A, int64_t lda,
                            const float* x, float beta, float* y, hipStream_t s) {
    if (M <= 0) return OMB_OK;
    const int64_t cus = device_cu_count();
    const int64_t waves_per_wg = NT / 64;
    int rows = 16;
    if (M < 16 * waves_per_wg * cus) rows = M < 4 * waves_per_wg * cus ? 1 : 4;
    // one workgroup per CU for 16-row groups, up to 4 for the smaller ones
    const int64_t blocks = std::min<int64_t>(ceil_div(ceil_div(M, rows), waves_per_wg),
                                             rows == 16 ? cus : 4 * cus);
    const dim3 grid((unsigned)blocks), block(NT);
    if (rows == 16 && N < kGemvOff32Cols)
        hipLaunchKernelGGL((sgemv_row_kernel<16, 1, true>), grid, block, 0, s, M, N, alpha, A, lda,
                           x, beta, y);
    else if (rows == 16)
        hipLaunchKernelGGL((sgemv_row_kernel<16, 1>), grid, block, 0, s, M, N, alpha, A, lda, x,
                           beta, y);
    else if (rows == 4)
        hipLaunchKernelGGL((sgemv_row_kernel<4, 4>), grid, block, 0, s, M, N, alpha, A, lda, x,
                           beta, y);
    else
        hipLaunchKernelGGL((sgemv_row_kernel<1, 8>), grid, block, 0, s, M, N, alpha, A, lda, x,
                           beta, y);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

omb_status launch_sgemv_col(int64_t M, int64_t N, float alpha, const float* A, int64_t lda,
                            const float* x, float beta, float* y, hipStream_t s) {
    if (M <= 0) return OMB_OK;
    const int64_t cus = device_cu_count();
    if (M % 64 == 0 && M / 64 >= cus && N > 0 && N % 256 == 0 && M / 64 <= 0x7fffffffLL) {
        hipLaunchKernelGGL(sgemv_col_rows64_kernel, dim3((unsigned)(M / 64)), dim3(NT), 0, s, M,
                           N, alpha, A, lda, x, beta, y);
        OMB_HIP_TRY(hipGetLastError());
        return OMB_OK;
    }
    const int64_t nrb = ceil_div(M, COL_ROWS);
    // About one workgroup per CU; every chunk at least 64 columns (one
    // unrolled step of the 4 wavefronts).
    int64_t nchunks = nrb >= cus ? 1 : cus / nrb;
    nchunks = std::max<int64_t>(1, std::min<int64_t>(nchunks, ceil_div(N, 4 * COL_UNROLL)));
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
    dim3 grid((unsigned)nrb, (unsigned)nchunks);
    float* part = reinterpret_cast<float*>(ws->buf);
    hipLaunchKernelGGL(sgemv_col_kernel, grid, dim3(NT), 0, s, M, N, alpha, A, lda, x, beta, y,
                       part, mpad, cpc, (int)nchunks);
    if (nchunks > 1)
        hipLaunchKernelGGL(sgemv_col_combine_kernel, dim3((unsigned)ceil_div(M, NT)),
                           dim3(NT), 0, s, M, alpha, part, mpad, (int)nchunks, beta, y);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

}  // namespace omb
