// One wavefront's share of a row-major SGEMV: y[r] = alpha * A[r,:] . x +
// beta * y[r] for up to ROWS consecutive rows r0 .. r0+nr-1.  Shared by the
// standalone row kernel and the grouped persistent kernel so that both give
// bitwise the same y: lane j accumulates the 4-float column groups j, j+64,
// ... in order (128-bit loads at any alignment), then the N % 4 trailing
// columns (lanes 0..2), and a fixed butterfly adds the lanes.  ROWS and
// UNROLL change only how many loads are in flight, never a row's order of
// accumulation, and neither does the alignment of A, lda or x.
#pragma once

#include "common.hpp"

namespace omb {

// (PA / PX / PY: generic or global-address-space pointers.  STREAM: rows of A
// are loaded with the non-temporal hint; the grouped kernel, whose items
// are short and share cache lines across item boundaries, loads them
// normally -- measured 3.5% faster there.)
template <int ROWS, int UNROLL, bool STREAM, typename PA, typename PX, typename PY>
__device__ __forceinline__ void gemv_rows_wave(int64_t r0, int nr, int64_t N, float alpha,
                                               PA __restrict__ A, int64_t lda,
                                               PX __restrict__ x, float beta,
                                               PY __restrict__ y, int lane) {
    float acc[ROWS];
    PA ar[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        acc[r] = 0.f;
        ar[r] = A + (r0 + (r < nr ? r : nr - 1)) * lda;  // rows past nr repeat the last one
    }
    const int64_t n4 = N >> 2;
    int64_t j = lane;
    for (; j + 64 * (UNROLL - 1) < n4; j += 64 * UNROLL) {
        float4 xv[UNROLL];
        float4 av[ROWS][UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) xv[u] = ld4u(x + 4 * (j + 64 * u));
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                const auto p = ar[r] + 4 * (j + 64 * u);
                av[r][u] = STREAM ? ld4u_nt(p) : ld4u(p);
            }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u)
#pragma unroll
            for (int r = 0; r < ROWS; ++r) acc[r] = dot4(av[r][u], xv[u], acc[r]);
    }
    for (; j < n4; j += 64) {
        const float4 xv = ld4u(x + 4 * j);
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const auto p = ar[r] + 4 * j;
            acc[r] = dot4(STREAM ? ld4u_nt(p) : ld4u(p), xv, acc[r]);
        }
    }
    const int64_t t = (n4 << 2) + lane;
    if (t < N) {
        const float xv = x[t];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) acc[r] = fmaf(ar[r][t], xv, acc[r]);
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) acc[r] = wave_sum(acc[r]);
    if (lane < nr) {
        float s = acc[0];
#pragma unroll
        for (int r = 1; r < ROWS; ++r)
            if (lane == r) s = acc[r];
        float out = alpha * s;
        if (beta != 0.f) out += beta * y[r0 + lane];
        y[r0 + lane] = out;
    }
}

}  // namespace omb
