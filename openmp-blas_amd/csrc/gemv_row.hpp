// One wavefront's share of a row-major SGEMV: y[r] = alpha * A[r,:] . x +
// beta * y[r] for up to 4 consecutive rows r0 .. r0+nr-1.  Shared by the
// standalone row kernel and the grouped persistent kernel so that both give
// bitwise the same y (lane j accumulates the float4 columns j, j+64, ... in
// order; a fixed butterfly adds the lanes).
#pragma once

#include "common.hpp"

namespace omb {

constexpr int kGemvRows = 4;  // rows per wavefront pass

// UNROLL = float4 loads in flight per row per lane; it changes only the
// instruction schedule, not any lane's order of accumulation.
template <bool VEC, int UNROLL = 4>
__device__ __forceinline__ void gemv_rows_wave(int64_t r0, int nr, int64_t N, float alpha,
                                               const float* __restrict__ A, int64_t lda,
                                               const float* __restrict__ x, float beta,
                                               float* __restrict__ y, int lane) {
    float acc[kGemvRows];
#pragma unroll
    for (int r = 0; r < kGemvRows; ++r) acc[r] = 0.f;
    int64_t done = 0;
    if (VEC) {
        const int64_t n4 = N >> 2;
        const float4* x4 = reinterpret_cast<const float4*>(x);
        const float4* a4[kGemvRows];
#pragma unroll
        for (int r = 0; r < kGemvRows; ++r)
            a4[r] = reinterpret_cast<const float4*>(A + (r0 + (r < nr ? r : nr - 1)) * lda);
        int64_t j = lane;
        for (; j + 64 * (UNROLL - 1) < n4; j += 64 * UNROLL) {
            float4 xv[UNROLL];
            float4 av[kGemvRows][UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) xv[u] = x4[j + 64 * u];
#pragma unroll
            for (int r = 0; r < kGemvRows; ++r)
#pragma unroll
                for (int u = 0; u < UNROLL; ++u) av[r][u] = nt_load4(&a4[r][j + 64 * u]);
#pragma unroll
            for (int u = 0; u < UNROLL; ++u)
#pragma unroll
                for (int r = 0; r < kGemvRows; ++r) {
                    acc[r] = fmaf(av[r][u].x, xv[u].x, acc[r]);
                    acc[r] = fmaf(av[r][u].y, xv[u].y, acc[r]);
                    acc[r] = fmaf(av[r][u].z, xv[u].z, acc[r]);
                    acc[r] = fmaf(av[r][u].w, xv[u].w, acc[r]);
                }
        }
        for (; j < n4; j += 64) {
            const float4 xv = x4[j];
#pragma unroll
            for (int r = 0; r < kGemvRows; ++r) {
                const float4 av = nt_load4(&a4[r][j]);
                acc[r] = fmaf(av.x, xv.x, acc[r]);
                acc[r] = fmaf(av.y, xv.y, acc[r]);
                acc[r] = fmaf(av.z, xv.z, acc[r]);
                acc[r] = fmaf(av.w, xv.w, acc[r]);
            }
        }
        done = n4 << 2;
    }
    for (int64_t j = done + lane; j < N; j += 64) {
        const float xv = x[j];
#pragma unroll
        for (int r = 0; r < kGemvRows; ++r)
            if (r < nr) acc[r] = fmaf(A[(r0 + r) * lda + j], xv, acc[r]);
    }
#pragma unroll
    for (int r = 0; r < kGemvRows; ++r) acc[r] = wave_sum(acc[r]);
    if (lane < nr) {
        float s = acc[0];
#pragma unroll
        for (int r = 1; r < kGemvRows; ++r)
            if (lane == r) s = acc[r];
        float out = alpha * s;
        if (beta != 0.f) out += beta * y[r0 + lane];
        y[r0 + lane] = out;
    }
}

// 128-bit path allowed for this operand set.
inline bool gemv_row_vec(const float* A, int64_t lda, const float* x) {
    return ((reinterpret_cast<uintptr_t>(A) & 15) == 0) &&
           ((reinterpret_cast<uintptr_t>(x) & 15) == 0) && (lda % 4 == 0);
}

}  // namespace omb
