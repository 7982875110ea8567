// Outer product (rank-1 update) and transpose for gfx950: the bandwidth-bound
// BLAS routines beside SGEMV.
//
// SGER, C += alpha * x * y^T (counterpart of amt::outer_prod_helper,
// include/outer.hpp:15-48, whose simd_loop<OUTER> does c[i] += a[i] * b[j]
// line by line, include/simd_loop.hpp:38-50).  Every element is one fused
// multiply-add, c = fma(alpha * x[i], y[j], c), in whichever storage order C
// has: a workgroup walks whole lines of C along its unit-stride index with
// 128-bit read-modify-writes.
//
// Transpose, C = A^T (counterpart of amt::transpose_helper,
// include/trans.hpp:23-80, out-of-place and square in-place).  When the
// unit-stride index of A is the same matrix index as that of C the transpose
// is a line copy; otherwise 64x64 tiles go through LDS so that both the
// reads and the writes are coalesced.  Square in-place transposes swap tile
// pairs (i, j) and (j, i) through LDS in one workgroup.  Writes are
// non-temporal.
#include <algorithm>

#include "common.hpp"

namespace omb {
namespace {

constexpr int NT = 256;
constexpr int GER_UNROLL = 4;  // 128-bit groups per thread per step

// 128-bit streaming store of 4 consecutive floats at any 4-byte aligned
// address (non-temporal: measured 5-8% faster than plain stores for these
// read-modify-write and copy streams).
__device__ inline void st4u(float* p, float4 v) {
    f32x4_u t = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(t, reinterpret_cast<f32x4_u*>(p));
}

// Rank-1 update of a row-major view D (R x Cc, rows ld apart): D[r][c] +=
// a(r, c) * b(r, c) with a = alpha * (COLSCALE ? v[c] : u[r]) and b the other
// vector's element, so that the x factor of C += alpha x y^T is the scaled
// one in both storage orders.  Workgroups take rows in a grid stride; the
// threads of a workgroup cover a row with 128-bit groups.
template <bool COLSCALE>
__global__ __launch_bounds__(NT) void ger_kernel(int64_t R, int64_t Cc, float alpha,
                                                 const float* __restrict__ u,
                                                 const float* __restrict__ v, float* D,
                                                 int64_t ld) {
    const int64_t nq = Cc >> 2;
    for (int64_t r = blockIdx.x; r < R; r += gridDim.x) {
        float* row = D + r * ld;
        const float ur = COLSCALE ? u[r] : alpha * u[r];
        int64_t q = threadIdx.x;
        for (; q + NT * (GER_UNROLL - 1) < nq; q += NT * GER_UNROLL) {
            float4 c[GER_UNROLL], w[GER_UNROLL];
#pragma unroll
            for (int k = 0; k < GER_UNROLL; ++k) {
                c[k] = ld4u_nt(row + 4 * (q + NT * k));
                w[k] = ld4u(v + 4 * (q + NT * k));
            }
#pragma unroll
            for (int k = 0; k < GER_UNROLL; ++k) {
                float4 o;
                if (COLSCALE) {
                    o.x = fmaf(alpha * w[k].x, ur, c[k].x);
                    o.y = fmaf(alpha * w[k].y, ur, c[k].y);
                    o.z = fmaf(alpha * w[k].z, ur, c[k].z);
                    o.w = fmaf(alpha * w[k].w, ur, c[k].w);
                } else {
                    o.x = fmaf(ur, w[k].x, c[k].x);
                    o.y = fmaf(ur, w[k].y, c[k].y);
                    o.z = fmaf(ur, w[k].z, c[k].z);
                    o.w = fmaf(ur, w[k].w, c[k].w);
                }
                st4u(row + 4 * (q + NT * k), o);
            }
        }
        for (; q < nq; q += NT) {
            const float4 c = ld4u(row + 4 * q), w = ld4u(v + 4 * q);
            float4 o;
            if (COLSCALE) {
                o = make_float4(fmaf(alpha * w.x, ur, c.x), fmaf(alpha * w.y, ur, c.y),
                                fmaf(alpha * w.z, ur, c.z), fmaf(alpha * w.w, ur, c.w));
            } else {
                o = make_float4(fmaf(ur, w.x, c.x), fmaf(ur, w.y, c.y), fmaf(ur, w.z, c.z),
                                fmaf(ur, w.w, c.w));
            }
            st4u(row + 4 * q, o);
        }
        const int64_t t = (nq << 2) + threadIdx.x;
        if (t < Cc) row[t] = COLSCALE ? fmaf(alpha * v[t], ur, row[t]) : fmaf(ur, v[t], row[t]);
    }
}

// D (R x Cc, ldd) = S (R x Cc, lds), both row-major: a line copy.
__global__ __launch_bounds__(NT) void copy_lines_kernel(int64_t R, int64_t Cc,
                                                        const float* __restrict__ S, int64_t lds,
                                                        float* __restrict__ D, int64_t ldd) {
    const int64_t nq = Cc >> 2;
    for (int64_t r = blockIdx.x; r < R; r += gridDim.x) {
        const float* s = S + r * lds;
        float* d = D + r * ldd;
        int64_t q = threadIdx.x;
        for (; q + NT * (GER_UNROLL - 1) < nq; q += NT * GER_UNROLL) {
            float4 t[GER_UNROLL];
#pragma unroll
            for (int k = 0; k < GER_UNROLL; ++k) t[k] = ld4u_nt(s + 4 * (q + NT * k));
#pragma unroll
            for (int k = 0; k < GER_UNROLL; ++k) st4u(d + 4 * (q + NT * k), t[k]);
        }
        for (; q < nq; q += NT) st4u(d + 4 * q, ld4u(s + 4 * q));
        const int64_t t = (nq << 2) + threadIdx.x;
        if (t < Cc) d[t] = s[t];
    }
}

constexpr int TT = 64;       // transpose tile edge
constexpr int TPAD = TT + 1;  // LDS row pitch (conflict-free column reads)

// Read a TT x TT tile of S (row-major, lds) at (r0, c0) into LDS: thread t
// takes the 4-float group t % 16 of rows t / 16 + 16k.
__device__ inline void tile_load(float (*tile)[TPAD], const float* S, int64_t lds, int64_t R,
                                 int64_t Cc, int64_t r0, int64_t c0) {
    const int g = threadIdx.x & 15, rr = threadIdx.x >> 4;
#pragma unroll
    for (int k = 0; k < TT / 16; ++k) {
        const int lr = rr + 16 * k;
        const int64_t r = r0 + lr, c = c0 + 4 * g;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (r < R) {
            if (c + 3 < Cc) {
                v = ld4u_nt(S + r * lds + c);
            } else {
                const float* p = S + r * lds;
                if (c < Cc) v.x = p[c];
                if (c + 1 < Cc) v.y = p[c + 1];
                if (c + 2 < Cc) v.z = p[c + 2];
            }
        }
        tile[lr][4 * g + 0] = v.x;
        tile[lr][4 * g + 1] = v.y;
        tile[lr][4 * g + 2] = v.z;
        tile[lr][4 * g + 3] = v.w;
    }
}

// Write the transpose of an LDS tile (rows of S at r0.., columns c0..) into
// D (row-major, ldd, Cc x R): D[c][r] = tile[r - r0][c - c0].
__device__ inline void tile_store_t(const float (*tile)[TPAD], float* D, int64_t ldd, int64_t R,
                                    int64_t Cc, int64_t r0, int64_t c0) {
    const int g = threadIdx.x & 15, cc = threadIdx.x >> 4;
#pragma unroll
    for (int k = 0; k < TT / 16; ++k) {
        const int lc = cc + 16 * k;
        const int64_t c = c0 + lc, r = r0 + 4 * g;
        if (c >= Cc) continue;
        const float4 v = make_float4(tile[4 * g + 0][lc], tile[4 * g + 1][lc],
                                     tile[4 * g + 2][lc], tile[4 * g + 3][lc]);
        float* p = D + c * ldd;
        if (r + 3 < R) {
            st4u(p + r, v);
        } else {
            if (r < R) p[r] = v.x;
            if (r + 1 < R) p[r + 1] = v.y;
            if (r + 2 < R) p[r + 2] = v.z;
        }
    }
}

// D (Cc x R, ldd) = S^T (S: R x Cc, lds), both row-major; 64x64 tiles in a
// grid stride over a fixed grid (4 workgroups per CU measured 62% of HBM
// against 56% for one workgroup per tile).
__global__ __launch_bounds__(NT) void transpose_tiles_kernel(int64_t R, int64_t Cc,
                                                             const float* __restrict__ S,
                                                             int64_t lds, float* __restrict__ D,
                                                             int64_t ldd) {
    __shared__ float tile[TT][TPAD];
    const int64_t tc = (Cc + TT - 1) / TT, ntiles = tc * ((R + TT - 1) / TT);
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t c0 = (t % tc) * TT, r0 = (t / tc) * TT;
        __syncthreads();  // the previous tile has been written out
        tile_load(tile, S, lds, R, Cc, r0, c0);
        __syncthreads();
        tile_store_t(tile, D, ldd, R, Cc, r0, c0);
    }
}

// Square in-place transpose of A (n x n, row-major view with ld): workgroup
// (bj, bi) with bi <= bj swaps tile (bi, bj) with tile (bj, bi), each
// transposed; bi == bj transposes the diagonal tile.  Workgroups below the
// diagonal have nothing to do.
__global__ __launch_bounds__(NT) void transpose_inplace_kernel(int64_t n, float* A, int64_t ld) {
    __shared__ float t0[TT][TPAD];
    __shared__ float t1[TT][TPAD];
    const int64_t bj = blockIdx.x, bi = blockIdx.y;
    if (bi > bj) return;
    const int64_t r0 = bi * TT, c0 = bj * TT;
    tile_load(t0, A, ld, n, n, r0, c0);
    if (bi != bj) tile_load(t1, A, ld, n, n, c0, r0);
    __syncthreads();
    tile_store_t(t0, A, ld, n, n, r0, c0);  // (bi, bj)^T -> (bj, bi)
    if (bi != bj) tile_store_t(t1, A, ld, n, n, c0, r0);
}

int64_t row_grid(int64_t rows) {
    return std::max<int64_t>(1, std::min<int64_t>(rows, 4 * (int64_t)device_cu_count()));
}

bool overlaps(const float* a, int64_t na, const float* b, int64_t nb) {
    return na > 0 && nb > 0 && a < b + nb && b < a + na;
}

int64_t span_of(int64_t rows, int64_t cols, int64_t rs, int64_t cs) {
    return (rows <= 0 || cols <= 0) ? 0 : (rows - 1) * rs + (cols - 1) * cs + 1;
}

}  // namespace
}  // namespace omb

using namespace omb;

extern "C" omb_status omb_sger(int64_t M, int64_t N, float alpha, const float* x, const float* y,
                               float* C, int64_t rsC, int64_t csC, omb_stream_t stream) {
    clear_error();
    if (M < 0 || N < 0) {
        set_error("omb_sger: negative size");
        return OMB_EINVAL;
    }
    if (M == 0 || N == 0) return OMB_OK;
    if (!x || !y || !C) {
        set_error("omb_sger: null pointer");
        return OMB_EINVAL;
    }
    bool row = false;
    int64_t ld = 0;
    OMB_TRY(resolve_layout("omb_sger C", M, N, rsC, csC, &row, &ld));
    hipStream_t s = (hipStream_t)stream;
    if (row)
        hipLaunchKernelGGL(ger_kernel<false>, dim3((unsigned)row_grid(M)), dim3(NT), 0, s, M, N,
                           alpha, x, y, C, ld);
    else  // column-major C: its lines are the columns, rows of the view are j
        hipLaunchKernelGGL(ger_kernel<true>, dim3((unsigned)row_grid(N)), dim3(NT), 0, s, N, M,
                           alpha, y, x, C, ld);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

extern "C" omb_status omb_stranspose(int64_t M, int64_t N, const float* A, int64_t rsA,
                                     int64_t csA, float* C, int64_t rsC, int64_t csC,
                                     omb_stream_t stream) {
    clear_error();
    if (M < 0 || N < 0) {
        set_error("omb_stranspose: negative size");
        return OMB_EINVAL;
    }
    if (M == 0 || N == 0) return OMB_OK;
    if (!A || !C) {
        set_error("omb_stranspose: null pointer");
        return OMB_EINVAL;
    }
    bool a_row = false, c_row = false;
    int64_t lda = 0, ldc = 0;
    OMB_TRY(resolve_layout("omb_stranspose A", M, N, rsA, csA, &a_row, &lda));
    OMB_TRY(resolve_layout("omb_stranspose C", N, M, rsC, csC, &c_row, &ldc));
    if (overlaps(A, span_of(M, N, rsA, csA), C, span_of(N, M, rsC, csC))) {
        set_error("omb_stranspose: A and C overlap (use omb_stranspose_inplace)");
        return OMB_EINVAL;
    }
    // Row-major views: S is A as stored (R x Cc), D is C as stored.
    //   A row (M x N), C row (N x M): D = S^T           -> tiles
    //   A row,         C col: C as stored is M x N = S  -> copy
    //   A col (N x M as stored), C row (N x M) = S      -> copy
    //   A col,         C col (M x N as stored) = S^T    -> tiles
    const int64_t R = a_row ? M : N, Cc = a_row ? N : M;
    hipStream_t s = (hipStream_t)stream;
    if (a_row == c_row) {
        const int64_t ntiles = ceil_div(Cc, TT) * ceil_div(R, TT);
        const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(ntiles, 4 * device_cu_count()));
        hipLaunchKernelGGL(transpose_tiles_kernel, dim3((unsigned)grid), dim3(NT), 0, s, R, Cc, A,
                           lda, C, ldc);
    } else {
        hipLaunchKernelGGL(copy_lines_kernel, dim3((unsigned)row_grid(R)), dim3(NT), 0, s, R, Cc,
                           A, lda, C, ldc);
    }
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

extern "C" omb_status omb_stranspose_inplace(int64_t n, float* A, int64_t ld,
                                             omb_stream_t stream) {
    clear_error();
    if (n < 0 || (n > 0 && ld < n)) {
        set_error("omb_stranspose_inplace: need n >= 0 and ld >= n");
        return OMB_EINVAL;
    }
    if (n == 0) return OMB_OK;
    if (!A) {
        set_error("omb_stranspose_inplace: null pointer");
        return OMB_EINVAL;
    }
    const int64_t t = ceil_div(n, TT);
    if (t > 65535) {
        set_error("omb_stranspose_inplace: matrix too large for one launch");
        return OMB_EUNSUPPORTED;
    }
    hipLaunchKernelGGL(transpose_inplace_kernel, dim3((unsigned)t, (unsigned)t), dim3(NT), 0,
                       (hipStream_t)stream, n, A, ld);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}
