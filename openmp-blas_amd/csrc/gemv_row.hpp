// One wavefront's share of a row-major SGEMV: y[r] = alpha * A[r,:] . x +
// beta * y[r] for up to ROWS consecutive rows r0 .. r0+nr-1.  Shared by the
// standalone row kernel and the grouped persistent kernel so that both give
// bitwise the same y: lane j accumulates the 4-float column groups j, j+64,
// ... in order (128-bit loads at any alignment), then the N % 4 trailing
// columns (lanes 0..2), and a fixed butterfly adds the lanes.  ROWS and
// UNROLL change only how many loads are in flight and OFF32 only how their
// addresses are formed, never a row's order of accumulation, and neither
// does the alignment of A, lda or x.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace omb {

// A 64-bit pointer that every lane of the wavefront holds: made scalar, so a
// load through it adds one 32-bit lane offset to a scalar base.
template <typename P>
__device__ inline P wave_uniform_ptr(P p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (P)(((uint64_t)hi << 32) | lo);
}

// Rows of fewer columns than this take OFF32 addressing (the largest byte
// offset within a row, 4 N plus one unrolled step, stays below 2^32).
constexpr int64_t kGemvOff32Cols = (int64_t)1 << 29;

// (PA / PX / PY: generic or global-address-space pointers.  STREAM: rows of A
// are loaded with the non-temporal hint; the grouped kernel, whose items
// are short and share cache lines across item boundaries, loads them
// normally -- measured 3.5% faster there.  OFF32, for N < kGemvOff32Cols and
// a wavefront-uniform r0: the row bases are scalar and every load adds one
// 32-bit lane offset, which frees about 90 VGPRs of per-row addresses and
// measured 1.5% more bandwidth in the standalone row kernel, 0.8% in the
// grouped kernel's GEMV items.)
template <int ROWS, int UNROLL, bool STREAM, bool OFF32 = false, typename PA, typename PX,
          typename PY>
__device__ __forceinline__ void gemv_rows_wave(int64_t r0, int nr, int64_t N, float alpha,
                                               PA __restrict__ A, int64_t lda,
                                               PX __restrict__ x, float beta,
                                               PY __restrict__ y, int lane) {
    typedef typename std::conditional<OFF32, int, int64_t>::type Idx;
    // float4 group j of row pointer p
    auto at = [](auto p, Idx j) {
        if constexpr (OFF32) return (decltype(p))((uint64_t)p + 16u * (uint32_t)j);
        else return p + 4 * j;
    };
    float acc[ROWS];
    PA ar[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        acc[r] = 0.f;
        ar[r] = A + (r0 + (r < nr ? r : nr - 1)) * lda;  // rows past nr repeat the last one
        if constexpr (OFF32) ar[r] = wave_uniform_ptr(ar[r]);
    }
    if constexpr (OFF32) x = wave_uniform_ptr(x);
    const Idx n4 = (Idx)(N >> 2);
    Idx j = lane;
    for (; j + 64 * (UNROLL - 1) < n4; j += 64 * UNROLL) {
        float4 xv[UNROLL];
        float4 av[ROWS][UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) xv[u] = ld4u(at(x, j + 64 * u));
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                const auto p = at(ar[r], j + 64 * u);
                av[r][u] = STREAM ? ld4u_nt(p) : ld4u(p);
            }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u)
#pragma unroll
            for (int r = 0; r < ROWS; ++r) acc[r] = dot4(av[r][u], xv[u], acc[r]);
    }
    for (; j < n4; j += 64) {
        const float4 xv = ld4u(at(x, j));
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const auto p = at(ar[r], j);
            acc[r] = dot4(STREAM ? ld4u_nt(p) : ld4u(p), xv, acc[r]);
        }
    }
    const Idx t = (n4 << 2) + lane;
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
