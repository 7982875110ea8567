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

// The 4 rows of a TT x TT tile of S (row-major, lds) at (r0, c0) that
// thread t reads: 4-float group t % 16 of rows t / 16 + 16k, zero past the
// edges; kept in registers so that several tiles' loads are in flight.
__device__ inline void tile_regs(float4 (&v)[TT / 16], const float* S, int64_t lds, int64_t R,
                                 int64_t Cc, int64_t r0, int64_t c0) {
    const int g = threadIdx.x & 15, rr = threadIdx.x >> 4;
#pragma unroll
    for (int k = 0; k < TT / 16; ++k) {
        const int64_t r = r0 + rr + 16 * k, c = c0 + 4 * g;
        v[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (r < R) {
            if (c + 3 < Cc) {
                v[k] = ld4u_nt(S + r * lds + c);
            } else {
                const float* p = S + r * lds;
                if (c < Cc) v[k].x = p[c];
                if (c + 1 < Cc) v[k].y = p[c + 1];
                if (c + 2 < Cc) v[k].z = p[c + 2];
            }
        }
    }
}

__device__ inline void tile_put(float (*tile)[TPAD], const float4 (&v)[TT / 16]) {
    const int g = threadIdx.x & 15, rr = threadIdx.x >> 4;
#pragma unroll
    for (int k = 0; k < TT / 16; ++k) {
        const int lr = rr + 16 * k;
        tile[lr][4 * g + 0] = v[k].x;
        tile[lr][4 * g + 1] = v[k].y;
        tile[lr][4 * g + 2] = v[k].z;
        tile[lr][4 * g + 3] = v[k].w;
    }
}

// D (Cc x R, ldd) = S^T (S: R x Cc, lds), both row-major, over the whole
// 64x64 tiles (the first TR x TC tiles): each step of a workgroup reads TPB
// consecutive tiles into registers (all loads in flight at once), parks
// them in LDS and writes their transposes; grid stride over a fixed grid.
// No edge guards: guarded loads and stores in this loop measured slower
// (0.55 of HBM peak at 16384^2 against 0.65), so partial tiles go to
// transpose_edge_kernel.  One tile per step measured 0.57 (a plain copy,
// read and write, reaches 0.75 there).
constexpr int TPB = 4;
__global__ __launch_bounds__(NT) void transpose_tiles_kernel(int64_t TR, int64_t TC,
                                                             const float* __restrict__ S,
                                                             int64_t lds, float* __restrict__ D,
                                                             int64_t ldd) {
    __shared__ float tile[TPB][TT][TPAD];
    const int64_t ntiles = TR * TC;
    const int g = threadIdx.x & 15, rr = threadIdx.x >> 4;
    for (int64_t t0 = (int64_t)blockIdx.x * TPB; t0 < ntiles; t0 += (int64_t)gridDim.x * TPB) {
        float4 v[TPB][TT / 16];
#pragma unroll
        for (int b = 0; b < TPB; ++b) {
            const int64_t t = t0 + b < ntiles ? t0 + b : ntiles - 1;
            const int64_t r0 = (t / TC) * TT, c0 = (t % TC) * TT;
#pragma unroll
            for (int k = 0; k < TT / 16; ++k)
                v[b][k] = ld4u_nt(S + (r0 + rr + 16 * k) * lds + c0 + 4 * g);
        }
        __syncthreads();  // the previous step's tiles have been written out
#pragma unroll
        for (int b = 0; b < TPB; ++b) tile_put(tile[b], v[b]);
        __syncthreads();
#pragma unroll
        for (int b = 0; b < TPB; ++b) {
            const int64_t t = t0 + b;
            if (t >= ntiles) break;
            const int64_t r0 = (t / TC) * TT, c0 = (t % TC) * TT;
#pragma unroll
            for (int k = 0; k < TT / 16; ++k) {
                const int lc = rr + 16 * k;
                st4u(D + (c0 + lc) * ldd + r0 + 4 * g,
                     make_float4(tile[b][4 * g][lc], tile[b][4 * g + 1][lc], tile[b][4 * g + 2][lc],
                                 tile[b][4 * g + 3][lc]));
            }
        }
    }
}

// The partial tiles of the same transpose: the last tile column (when Cc %
// TT != 0) for the TR whole tile rows, then the last tile row (when R % TT
// != 0) across all tile columns; one tile per workgroup, guarded.
__global__ __launch_bounds__(NT) void transpose_edge_kernel(int64_t R, int64_t Cc,
                                                            const float* __restrict__ S,
                                                            int64_t lds, float* __restrict__ D,
                                                            int64_t ldd) {
    __shared__ float tile[TT][TPAD];
    const int64_t TR = R / TT, TC = Cc / TT, tcol = (Cc + TT - 1) / TT;
    const int64_t ncol = Cc % TT ? TR : 0;  // right strip
    const int64_t e = blockIdx.x;
    int64_t tr, tc;
    if (e < ncol) {
        tr = e;
        tc = TC;
    } else {
        tr = TR;
        tc = e - ncol;
        if (tc >= tcol) return;
    }
    float4 v[TT / 16];
    tile_regs(v, S, lds, R, Cc, tr * TT, tc * TT);
    tile_put(tile, v);
    __syncthreads();
    tile_store_t(tile, D, ldd, R, Cc, tr * TT, tc * TT);
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
    float4 v0[TT / 16], v1[TT / 16];  // both tiles' loads in flight at once
    tile_regs(v0, A, ld, n, n, r0, c0);
    if (bi != bj) tile_regs(v1, A, ld, n, n, c0, r0);
    tile_put(t0, v0);
    if (bi != bj) tile_put(t1, v1);
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
        const int64_t TR = R / TT, TC = Cc / TT, ntiles = TR * TC;
        if (ntiles > 0) {
            const int64_t grid = std::max<int64_t>(
                1, std::min<int64_t>(ceil_div(ntiles, TPB), 2 * device_cu_count()));
            hipLaunchKernelGGL(transpose_tiles_kernel, dim3((unsigned)grid), dim3(NT), 0, s, TR,
                               TC, A, lda, C, ldc);
        }
        const int64_t nedge = (Cc % TT ? TR : 0) + (R % TT ? ceil_div(Cc, TT) : 0);
        if (nedge > 0x7fffffffLL) {
            set_error("omb_stranspose: too many edge tiles for one launch");
            return OMB_EUNSUPPORTED;
        }
        if (nedge > 0)
            hipLaunchKernelGGL(transpose_edge_kernel, dim3((unsigned)nedge), dim3(NT), 0, s, R, Cc,
                               A, lda, C, ldc);
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
