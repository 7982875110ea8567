// Shared pieces of the split-BF16 GEMM kernels: tile sizes, the MFMA, the
// swizzled LDS slice layout and the exact 3-way FP32 -> BF16 split.
#pragma once

#include "common.hpp"

namespace omb {
namespace split {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int TM = 128, TN = 128, TK = 16;  // C tile and k slice of one workgroup
constexpr int NT = 256;                      // threads per workgroup (4 wavefronts)
constexpr int kGroupM = 8;                   // tile rows per L2 group
constexpr int SLICE_A = 3 * TM * TK;         // bf16 elements per A slice (3 planes)
constexpr int SLICE_B = 3 * TN * TK;
constexpr int kFoldSlices = 16;              // inner MFMA accumulation spans 256 k

__device__ inline f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ inline f32x16 mfma(f16x8 a, f16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

// LDS position (in 16-byte chunks, within one plane) of chunk h (0/1) of
// slice row r: the chunk index is flipped by bits 2 and 3 of the row so that
// the 8 or 16 lanes a ds_read_b128 serves per cycle (rows r..r+15, same h)
// hit distinct banks.
__device__ inline int lds_chunk(int r, int h) {
    return (r << 1) | (h ^ ((r >> 2) & 1) ^ ((r >> 3) & 1));
}

// x = h0 + h1 + h2 exactly for |x| >= 2^-110 and zero (each subtraction is
// exact in FP32); below 2^-110 the last piece rounds to BF16's subnormal
// step, within 2^-134.  A finite x above BF16's largest finite value
// (|x| > 0x1.fep127 rounds to infinity) takes h0 truncated instead, so the
// split stays finite and exact.  An infinite or NaN x gives NaN pieces:
// every output it feeds is NaN.
__device__ inline void split3(float x, __bf16& h0, __bf16& h1, __bf16& h2) {
    h0 = (__bf16)x;
    if (__builtin_isinf((float)h0) && !__builtin_isinf(x))
        h0 = __builtin_bit_cast(__bf16, (unsigned short)(__float_as_uint(x) >> 16));
    const float r1 = x - (float)h0;
    h1 = (__bf16)r1;
    h2 = (__bf16)(r1 - (float)h1);
}

// Scaled FP16 split (two planes per packed slice): the power-of-two exponent
// that brings a line whose largest magnitude is amax into [2^14, 2^15) (0
// for a zero, infinite or NaN amax), and x * 2^e = h0 + h1 + r with
// h0 = fp16(x 2^e) and h1 = fp16(x 2^e - h0) (the subtraction is exact),
// |r| <= max(2^-22 |x 2^e|, 2^-25): relative to x, or FP16's subnormal step.
__device__ inline int line_exp(float amax) {
    if (!(amax > 0.f) || __builtin_isinf(amax)) return 0;
    return 14 - ilogbf(amax);
}
__device__ inline void split2_f16(float xs, _Float16& h0, _Float16& h1) {
    h0 = (_Float16)xs;
    h1 = (_Float16)(xs - (float)h0);
}

// Split 8 consecutive-k values into the three 16-byte plane chunks.
__device__ inline void split8(float4 lo, float4 hi, uint4& p0, uint4& p1, uint4& p2) {
    const float v[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
    bf16x8 a, b, c;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        __bf16 t0, t1, t2;
        split3(v[j], t0, t1, t2);
        a[j] = t0;
        b[j] = t1;
        c[j] = t2;
    }
    p0 = *reinterpret_cast<const uint4*>(&a);
    p1 = *reinterpret_cast<const uint4*>(&b);
    p2 = *reinterpret_cast<const uint4*>(&c);
}

}  // namespace split
}  // namespace omb

namespace omb {
namespace split {

// Fragments of one k slice for one wavefront owning BI x BJ blocks of 32x32
// of C (rows wrow0.., columns wcol0.. of the slice), from the swizzled LDS
// slice (PA rows per A plane, PB per B plane; PLA planes in the A part).
// V is the 8-element operand vector of the MFMA: bf16x8 (BF16 pieces) or
// f16x8 (scaled FP16 pieces).
template <int NPLANES, int BI, int BJ, int PA, int PB, typename V, int PLA>
__device__ __forceinline__ void read_frags(const uint4* slice, int wrow0, int wcol0, int frow,
                                           int fh, V (&a)[NPLANES][BI], V (&b)[NPLANES][BJ]) {
    static_assert(NPLANES <= PLA, "the slice holds every plane the products use");
    const uint4* la = slice;
    const uint4* lb = slice + PLA * 2 * PA;
#pragma unroll
    for (int pl = 0; pl < NPLANES; ++pl) {
#pragma unroll
        for (int i = 0; i < BI; ++i) {
            uint4 va = la[pl * 2 * PA + lds_chunk(wrow0 + i * 32 + frow, fh)];
            a[pl][i] = *reinterpret_cast<const V*>(&va);
        }
#pragma unroll
        for (int j = 0; j < BJ; ++j) {
            uint4 vb = lb[pl * 2 * PB + lds_chunk(wcol0 + j * 32 + frow, fh)];
            b[pl][j] = *reinterpret_cast<const V*>(&vb);
        }
    }
}

// The slice's MFMAs from its fragments, smallest products first, all into
// the inner accumulators (NPROD = 6: all pieces of 3 planes; 3: the three
// largest products of 2 planes).
template <int NPROD, int NPLANES, int BI, int BJ, typename V>
__device__ __forceinline__ void mfma_frags(const V (&a)[NPLANES][BI], const V (&b)[NPLANES][BJ],
                                           f32x16 (&inner)[BI][BJ]) {
    static_assert(NPLANES == (NPROD == 6 ? 3 : 2), "planes of the products");
#pragma unroll
    for (int i = 0; i < BI; ++i)
#pragma unroll
        for (int j = 0; j < BJ; ++j) {
            if constexpr (NPROD == 6) {
                inner[i][j] = mfma(a[0][i], b[2][j], inner[i][j]);
                inner[i][j] = mfma(a[1][i], b[1][j], inner[i][j]);
                inner[i][j] = mfma(a[2][i], b[0][j], inner[i][j]);
            }
            inner[i][j] = mfma(a[0][i], b[1][j], inner[i][j]);
            inner[i][j] = mfma(a[1][i], b[0][j], inner[i][j]);
            inner[i][j] = mfma(a[0][i], b[0][j], inner[i][j]);
        }
}

// MFMAs of one k slice: its fragment reads, then its MFMAs.
// PLA: planes of the A part of the LDS slice (3, or 2 for the two-plane forms).
template <int NPROD, int BI = 2, int BJ = 2, int PA = TM, int PB = TN, typename V = bf16x8,
          int PLA = 3>
__device__ __forceinline__ void mfma_slice(const uint4* slice, int wrow0, int wcol0, int frow,
                                           int fh, f32x16 (&inner)[BI][BJ]) {
    constexpr int NPLANES = NPROD == 6 ? 3 : 2;
    V a[NPLANES][BI], b[NPLANES][BJ];
    read_frags<NPLANES, BI, BJ, PA, PB, V, PLA>(slice, wrow0, wcol0, frow, fh, a, b);
    mfma_frags<NPROD, NPLANES, BI, BJ, V>(a, b, inner);
}

// outer += inner; inner = 0 (FP32 adds: the running sum over K is rounded
// by the vector ALUs, not inside the MFMA accumulation).
template <int BI, int BJ>
__device__ __forceinline__ void fold(f32x16 (&inner)[BI][BJ], f32x16 (&outer)[BI][BJ]) {
#pragma unroll
    for (int i = 0; i < BI; ++i)
#pragma unroll
        for (int j = 0; j < BJ; ++j) {
            outer[i][j] += inner[i][j];
            inner[i][j] = f32x16{};
        }
}

// C = alpha * acc + beta * C for the wavefront's BI x BJ blocks (row-major C).
// 32x32 accumulator layout: lane l holds column l % 32, value r holds row
// 8*(r/4) + 4*(l/32) + r%4.  With row and column exponents (scaled FP16
// operands: row m of A was scaled by 2^ea[m], column n of B by 2^eb[n]),
// acc is first scaled back by 2^-(ea[m] + eb[n]), exactly unless the result
// leaves FP32's normal range (as the FP32 product itself would).
template <int BI, int BJ, typename PC, typename PE = const int*>
__device__ __forceinline__ void store_tile(const f32x16 (&acc)[BI][BJ], int64_t m_base,
                                           int64_t n_base, int lane, PC C, int64_t ldc,
                                           int64_t M, int64_t N, float alpha, float beta,
                                           PE ea = nullptr, PE eb = nullptr) {
#pragma unroll
    for (int i = 0; i < BI; ++i)
#pragma unroll
        for (int j = 0; j < BJ; ++j) {
            const int64_t n = n_base + j * 32 + (lane & 31);
            if (n >= N) continue;
            const int en = ea ? eb[n] : 0;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t m = m_base + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (m < M) {
                    const PC c = C + m * ldc + n;
                    const float a = ea ? ldexpf(acc[i][j][r], -(ea[m] + en)) : acc[i][j][r];
                    const float v = alpha * a;
                    *c = beta != 0.f ? v + beta * *c : v;
                }
            }
        }
}

}  // namespace split
}  // namespace omb
