// Split-BF16 SGEMM for gfx950: an FP32 GEMM computed on the BF16 MFMA cores.
//
// Every FP32 operand element x is split into three BF16 terms
//   x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1)
// (each subtraction is exact in FP32; 3 x 8 significand bits cover FP32's 24,
// so the split is exact for |x| >= 2^-110, and within 2^-134 below),
// and  a*b ~ a0b0 + (a0b1 + a1b0) + (a0b2 + a1b1 + a2b0)  -- 6 products, the
// dropped terms are below 2^-24 relative.  The 3-product variant keeps only
// a0b0 + a0b1 + a1b0.
//
// Pass 1 (split_pack_kernel) converts an operand, in either storage order,
// into "packed" K-contiguous tiles: for every BR-row block and 16-wide k
// slice, the three planes are stored back to back, so one workgroup's slice
// of an operand is a single contiguous run of 3*BR*16 BF16 values (this is the
// GPU form of the reference's panel packing, include/utils.hpp:99-141).
// Padding rows/columns are zero.
//
// Pass 2 (sgemm_bf16x3_kernel): a 256-thread workgroup owns a 128x128 tile of
// C; four wavefronts each own 64x64 (2x2 blocks of v_mfma_f32_32x32x16_bf16).
// Slices are prefetched two ahead into registers and staged through LDS
// (double buffered, bank-swizzled).  All products of 16 slices (256 k) are
// accumulated by the MFMAs into an inner register tile, smallest first, which
// is then added into an outer tile with FP32 adds: the running sum over K
// never absorbs MFMA rounding at its own magnitude.
#include <algorithm>

#include "split_common.hpp"

namespace omb {
namespace {

using namespace split;

// ----------------------------------------------------------------- pass 1 --
// One workgroup converts a 128-row x 32-k tile (2 slices) of an operand:
// coalesced FP32 reads along whichever index is unit-stride, the 3-way split
// in registers, the packed tile assembled in LDS, then 24 KB of contiguous
// coalesced stores (a row tile's slices are adjacent in the packed form).
// Two slices (24 KB of LDS, 6 workgroups per CU) measured 7-18% faster than
// four at 8192^2.
constexpr int SPLIT_SL = 2;  // slices per workgroup

// Store 4 consecutive k (kk4 .. kk4+3 within the 64-k tile) of one row: the
// BF16 pieces (PLN = 3 planes, or 2 for the 3-product form, which never
// reads the third), or (F16, PLN = 2) the two FP16 pieces of the
// row-scaled values v * 2^e.
template <bool F16, int PLN>
__device__ inline void put4(__bf16* lds, int row, int kk4, const float (&v)[4], int e) {
    typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
    u16x4 p0, p1, p2 = {};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (F16) {
            _Float16 a, b;
            split2_f16(ldexpf(v[i], e), a, b);
            p0[i] = __builtin_bit_cast(unsigned short, a);
            p1[i] = __builtin_bit_cast(unsigned short, b);
        } else {
            __bf16 a, b, c;
            split3(v[i], a, b, c);
            p0[i] = __builtin_bit_cast(unsigned short, a);
            p1[i] = __builtin_bit_cast(unsigned short, b);
            p2[i] = __builtin_bit_cast(unsigned short, c);
        }
    }
    const int sl = kk4 >> 4, k16 = kk4 & 15;
    __bf16* base = lds + ((sl * PLN) * TM + row) * TK + k16;
    *reinterpret_cast<u16x4*>(base) = p0;
    *reinterpret_cast<u16x4*>(base + TM * TK) = p1;
    if (PLN == 3) *reinterpret_cast<u16x4*>(base + 2 * TM * TK) = p2;
}

// Row r's scale exponent for the FP16 split (0 for the BF16 split).
template <bool F16>
__device__ inline int row_exp(const unsigned* amax, int64_t r, int64_t R) {
    if (!F16 || r >= R) return 0;
    return line_exp(__uint_as_float(amax[r]));
}

__device__ inline float ld(const float* X, int64_t R, int64_t K, int64_t rs, int64_t cs,
                           int64_t r, int64_t k) {
    return (r < R && k < K) ? X[r * rs + k * cs] : 0.f;
}

// F16: rows are scaled by 2^e (e from the row's largest magnitude, `amax`
// as float bits) and split into two FP16 pieces; the workgroups of the
// first k group also write each row's e to `exps`.
// Row tile rt, k group kg (slices kg*SPLIT_SL ..) of one operand.
template <bool F16, int PLN>
__device__ __forceinline__ void pack_tile(const float* __restrict__ X, int64_t R, int64_t K,
                                          int64_t rs, int64_t cs, int64_t nkt, int k_major,
                                          __bf16* __restrict__ out,
                                          const unsigned* __restrict__ amax,
                                          int* __restrict__ exps, int64_t rt, int64_t kg) {
    __shared__ __attribute__((aligned(16))) __bf16 lds[SPLIT_SL * 3 * TM * TK];  // 24 KB
    const int tid = threadIdx.x;
    const int64_t kt0 = kg * SPLIT_SL;
    const int64_t r0 = rt * TM, k0 = kt0 * TK;
    if (F16 && kg == 0 && tid < TM && r0 + tid < R)
        exps[r0 + tid] = row_exp<true>(amax, r0 + tid, R);
    // 128-bit loads (any alignment) need a unit stride along the loaded
    // index; other strides take the element path.  The values are the same.
    const bool vec = k_major ? cs == 1 : rs == 1;
    if (k_major) {
        // 128 rows x (4 * SPLIT_SL) float4 along k: thread -> (row, k4)
#pragma unroll
        for (int j = 0; j < 2 * SPLIT_SL; ++j) {
            const int f = tid + NT * j, row = f / (4 * SPLIT_SL), kk4 = (f % (4 * SPLIT_SL)) * 4;
            const int64_t r = r0 + row, k = k0 + kk4;
            float v[4];
            if (vec && r < R && k + 3 < K) {
                const float4 t = ld4u_nt(X + r * rs + k);
                v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = ld(X, R, K, rs, cs, r, k + e);
            }
            put4<F16, PLN>(lds, row, kk4, v, row_exp<F16>(amax, r, R));
        }
    } else {
        // (16 * SPLIT_SL) k-columns x 32 float4 along rows: thread -> (4 columns, 4 rows)
        static_assert(SPLIT_SL * 128 % NT == 0, "whole passes of the workgroup");
#pragma unroll
        for (int j = 0; j < SPLIT_SL * 128 / NT; ++j) {
            const int f = tid + NT * j, kq = f >> 5, row4 = (f & 31) * 4;
            const int64_t r = r0 + row4;
            float m[4][4];  // [k][row]
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int64_t k = k0 + kq * 4 + c;
                if (vec && k < K && r + 3 < R) {
                    const float4 t = ld4u_nt(X + k * cs + r);
                    m[c][0] = t.x; m[c][1] = t.y; m[c][2] = t.z; m[c][3] = t.w;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) m[c][e] = ld(X, R, K, rs, cs, r + e, k);
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float v[4] = {m[0][e], m[1][e], m[2][e], m[3][e]};
                put4<F16, PLN>(lds, row4 + e, kq * 4, v, row_exp<F16>(amax, r + e, R));
            }
        }
    }
    __syncthreads();
    const int nsl = (int)(nkt - kt0 < SPLIT_SL ? nkt - kt0 : SPLIT_SL);
    static_assert(PLN == 3 || PLN == 2, "3 planes, or the 2 a 3-product form reads");
    static_assert(!F16 || PLN == 2, "the FP16 split has two pieces");
    constexpr int slice = PLN * TM * TK;  // elements per packed slice
    const uint4* src = reinterpret_cast<const uint4*>(lds);
    uint4* dst = reinterpret_cast<uint4*>(out + (rt * nkt + kt0) * (int64_t)slice);
    const int n16 = nsl * slice / 8;
    for (int i = tid; i < n16; i += NT) dst[i] = src[i];
}

template <bool F16, int PLN>
__global__ __launch_bounds__(NT) void split_pack_kernel(const float* __restrict__ X, int64_t R,
                                                        int64_t K, int64_t rs, int64_t cs,
                                                        int64_t nkt, int k_major,
                                                        __bf16* __restrict__ out,
                                                        const unsigned* __restrict__ amax,
                                                        int* __restrict__ exps) {
    pack_tile<F16, PLN>(X, R, K, rs, cs, nkt, k_major, out, amax, exps, blockIdx.x, blockIdx.y);
}

// Both operands of one product in one launch (same K): row tiles of the
// first, then of the second, along blockIdx.x.
struct PackOperand {
    const float* X;
    int64_t R, rs, cs;
    int k_major;
    __bf16* out;
    int64_t rtiles;
};
template <int PLN>
__global__ __launch_bounds__(NT) void split_pack_pair_kernel(PackOperand a, PackOperand b,
                                                             int64_t K, int64_t nkt) {
    const bool first = (int64_t)blockIdx.x < a.rtiles;
    const PackOperand& o = first ? a : b;
    const int64_t rt = first ? (int64_t)blockIdx.x : (int64_t)blockIdx.x - a.rtiles;
    pack_tile<false, PLN>(o.X, o.R, K, o.rs, o.cs, nkt, o.k_major, o.out, nullptr, nullptr, rt,
                          blockIdx.y);
}

// ----------------------------------------------------------------- pass 2 --
struct MArgs {
    const __bf16* Ap;
    const __bf16* Bp;
    int64_t nkt;
    float* C;
    int64_t ldc;
    int64_t M, N;
    float alpha, beta;
    int64_t tiles_m, tiles_n;
    int64_t ktc;  // slices per split (blockIdx.y), a multiple of kFoldSlices
    float* part;  // split partial sums [split][M][N], or null (one split)
    const int* ea;  // scaled FP16 operands: row exponents of A (else null)
    const int* eb;  // and column exponents of B
};

// One slice in flight is PL uint4 of A and PL of B per thread (PL = planes
// per slice, 3 or 2: one 16-byte load per plane), kept in named registers
// (arrays of them were demoted to scratch by the compiler).
#define OMB_GLOAD(kt, r0, r1, r2, r3, r4, r5)                            \
    do {                                                                  \
        const uint4* a_ = ga + (kt) * (SA / 8) + tid;                     \
        const uint4* b_ = gb + (kt) * (SB / 8) + tid;                     \
        r0 = a_[0];                                                       \
        r1 = a_[NT];                                                      \
        if constexpr (PL == 3) r2 = a_[2 * NT];                           \
        r3 = b_[0];                                                       \
        r4 = b_[NT];                                                      \
        if constexpr (PL == 3) r5 = b_[2 * NT];                           \
    } while (0)
#define OMB_LSTORE(l, r0, r1, r2, r3, r4, r5)                             \
    do {                                                                  \
        (l)[q0] = r0;                                                     \
        (l)[q1] = r1;                                                     \
        if constexpr (PL == 3) (l)[q2] = r2;                              \
        (l)[q3] = r3;                                                     \
        (l)[q4] = r4;                                                     \
        if constexpr (PL == 3) (l)[q5] = r5;                              \
    } while (0)

// Slice chunk q of A (0 .. PL*TM*2-1): plane q / (2*TM), row (q / 2) % TM,
// half q % 2 (same for B after the A part) -> its swizzled LDS position.
__device__ __forceinline__ int lds_index(int q) {
    const int pl = q / (2 * TM), rem = q - pl * 2 * TM;
    return pl * 2 * TM + lds_chunk(rem >> 1, rem & 1);
}

// V: bf16x8 for the BF16 pieces (PL = 3 planes per packed slice), f16x8
// for the scaled FP16 pieces (PL = 2; the sums are scaled back by the row
// and column exponents when stored).
template <int NPROD, typename V, int PL>
__global__ __launch_bounds__(NT, 2) void sgemm_bf16x3_kernel(MArgs p) {
    constexpr int SA = PL * TM * TK, SB = PL * TN * TK;  // elements per packed slice
    static_assert(SA / 8 / NT == PL && SB / 8 / NT == PL, "one 16-byte load per plane");
    __shared__ __attribute__((aligned(16))) uint4 lds[2][(SA + SB) / 8];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t ntiles = p.tiles_m * p.tiles_n;
    int64_t t = xcd_remap(blockIdx.x, ntiles);
    int64_t tm, tn;
    group_tile(t, p.tiles_m, p.tiles_n, kGroupM, &tm, &tn);

    // Split blockIdx.y takes slices [y * ktc, min(nkt, (y + 1) * ktc)) of the
    // tile's row and column panels; ktc is a multiple of kFoldSlices, so the
    // accumulation blocks are the same as without the split.
    const int64_t kt0 = (int64_t)blockIdx.y * p.ktc;
    const uint4* ga = reinterpret_cast<const uint4*>(p.Ap + (tm * p.nkt + kt0) * SA);
    const uint4* gb = reinterpret_cast<const uint4*>(p.Bp + (tn * p.nkt + kt0) * SB);
    // LDS slots of this thread's chunks (loop invariant).
    const int q0 = lds_index(tid), q1 = lds_index(tid + NT), q2 = lds_index(tid + 2 * NT);
    const int q3 = SA / 8 + q0, q4 = SA / 8 + q1, q5 = SA / 8 + q2;

    f32x16 inner[2][2], outer[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) inner[i][j] = outer[i][j] = f32x16{};
    const int frow = lane & 31, fh = lane >> 5;
    auto compute = [&](int buf, int64_t kt) {
        mfma_slice<NPROD, 2, 2, TM, TN, V, PL>(lds[buf], wm * 64, wn * 64, frow, fh, inner);
    };
    const int64_t nkt = p.nkt - kt0 < p.ktc ? p.nkt - kt0 : p.ktc;
    auto fold_at = [&](int64_t kt) {
        if ((kt + 1) % kFoldSlices == 0 || kt + 1 == nkt) fold(inner, outer);
    };
    uint4 x0, x1, x2, x3, x4, x5;  // even slices in flight
    uint4 y0, y1, y2, y3, y4, y5;  // odd slices in flight
    if constexpr (NPROD == 3 && PL == 2) {
        // Two-plane forms: the fragments are double-buffered in registers.
        // After each barrier a wave issues the next slice's loads and LDS
        // reads, then the MFMAs of the slice whose fragments it holds, so
        // the LDS reads overlap the MFMAs (the scheduling barriers keep that
        // order).  The loop runs whole slice pairs, with an odd last slice
        // after it, and loads and stores are unconditional with the slice
        // index clamped to the last one (past the end a store writes a copy
        // of the last slice into a buffer that is not read again): with one
        // path through the loop the compiler's wait counts are exact, and a
        // store waits only for its own loads.  The MFMA and fold order is
        // that of the loop below, so the results are the same.
        V fa0[2][2], fb0[2][2];  // fragments of an even slice
        V fa1[2][2], fb1[2][2];  // of an odd slice
        const int64_t klast = nkt - 1;  // >= 0: every split holds a slice
        auto kc = [&](int64_t k) { return k < klast ? k : klast; };
        OMB_GLOAD(0, x0, x1, x2, x3, x4, x5);
        OMB_GLOAD(kc(1), y0, y1, y2, y3, y4, y5);
        OMB_LSTORE(lds[0], x0, x1, x2, x3, x4, x5);
        OMB_LSTORE(lds[1], y0, y1, y2, y3, y4, y5);
        OMB_GLOAD(kc(2), x0, x1, x2, x3, x4, x5);
        __syncthreads();
        read_frags<2, 2, 2, TM, TN, V, PL>(lds[0], wm * 64, wn * 64, frow, fh, fa0, fb0);
        __syncthreads();  // every wave has read slice 0 before slice 2 replaces it
        int64_t kt = 0;
        for (; kt + 1 < nkt; kt += 2) {
            // fa0/fb0: slice kt; LDS 1: slice kt+1; x*: slice kt+2
            read_frags<2, 2, 2, TM, TN, V, PL>(lds[1], wm * 64, wn * 64, frow, fh, fa1, fb1);
            OMB_GLOAD(kc(kt + 3), y0, y1, y2, y3, y4, y5);
            __builtin_amdgcn_sched_barrier(0);
            mfma_frags<NPROD, 2, 2, 2, V>(fa0, fb0, inner);
            __builtin_amdgcn_sched_barrier(0);
            fold_at(kt);
            OMB_LSTORE(lds[0], x0, x1, x2, x3, x4, x5);
            __syncthreads();
            // fa1/fb1: slice kt+1; LDS 0: slice kt+2; y*: slice kt+3
            read_frags<2, 2, 2, TM, TN, V, PL>(lds[0], wm * 64, wn * 64, frow, fh, fa0, fb0);
            OMB_GLOAD(kc(kt + 4), x0, x1, x2, x3, x4, x5);
            __builtin_amdgcn_sched_barrier(0);
            mfma_frags<NPROD, 2, 2, 2, V>(fa1, fb1, inner);
            __builtin_amdgcn_sched_barrier(0);
            fold_at(kt + 1);
            OMB_LSTORE(lds[1], y0, y1, y2, y3, y4, y5);
            __syncthreads();
        }
        if (kt < nkt) {  // odd slice count: fa0/fb0 hold the last slice
            mfma_frags<NPROD, 2, 2, 2, V>(fa0, fb0, inner);
            fold_at(kt);
        }
    } else {
        OMB_GLOAD(0, x0, x1, x2, x3, x4, x5);
        if (nkt > 1) OMB_GLOAD(1, y0, y1, y2, y3, y4, y5);
        OMB_LSTORE(lds[0], x0, x1, x2, x3, x4, x5);
        __syncthreads();
        for (int64_t kt = 0; kt < nkt; kt += 2) {
            // even slice kt from LDS 0; y* hold kt+1; load kt+2 into x*
            if (kt + 2 < nkt) OMB_GLOAD(kt + 2, x0, x1, x2, x3, x4, x5);
            compute(0, kt);
            fold_at(kt);
            if (kt + 1 < nkt) OMB_LSTORE(lds[1], y0, y1, y2, y3, y4, y5);
            __syncthreads();
            if (kt + 1 >= nkt) break;
            // odd slice kt+1 from LDS 1; x* hold kt+2; load kt+3 into y*
            if (kt + 3 < nkt) OMB_GLOAD(kt + 3, y0, y1, y2, y3, y4, y5);
            compute(1, kt + 1);
            fold_at(kt + 1);
            if (kt + 2 < nkt) OMB_LSTORE(lds[0], x0, x1, x2, x3, x4, x5);
            __syncthreads();
        }
    }

    // (The exponents scale a split's partial sums before they are stored:
    // a power-of-two scale commutes with the in-order sum of the partials.)
    if (p.part)  // partial sums of this split (alpha 1, beta 0 store them unchanged)
        store_tile(outer, tm * TM + wm * 64, tn * TN + wn * 64, lane,
                   p.part + (int64_t)blockIdx.y * p.M * p.N, p.N, p.M, p.N, 1.f, 0.f, p.ea,
                   p.eb);
    else
        store_tile(outer, tm * TM + wm * 64, tn * TN + wn * 64, lane, p.C, p.ldc, p.M, p.N,
                   p.alpha, p.beta, p.ea, p.eb);
}

// ------------------------------------------------- scaled FP16 split --
// Largest magnitude of every row of an R x K operand (element (r, k) at
// X[r*rs + k*cs]), as the bits of |x| (for non-negative floats the unsigned
// order is the numeric order, and NaN's bits exceed infinity's, so a NaN
// row max is NaN).  The order of a max does not matter: deterministic.
// k-contiguous rows (cs == 1): one wavefront per row, written directly.
__global__ __launch_bounds__(NT) void amax_rows_kc_kernel(const float* __restrict__ X,
                                                          int64_t R, int64_t K, int64_t rs,
                                                          unsigned* __restrict__ amax) {
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    if (r >= R) return;
    const float* row = X + r * rs;
    unsigned m = 0;
    const int64_t n4 = K >> 2;
    for (int64_t j = lane; j < n4; j += 64) {
        const float4 v = ld4u_nt(row + 4 * j);
        m = max(m, max(max(__float_as_uint(fabsf(v.x)), __float_as_uint(fabsf(v.y))),
                       max(__float_as_uint(fabsf(v.z)), __float_as_uint(fabsf(v.w)))));
    }
    const int64_t t = (n4 << 2) + lane;
    if (t < K) m = max(m, __float_as_uint(fabsf(row[t])));
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off, 64));
    if (lane == 0) amax[r] = m;
}

// Other layouts, over 256-k chunks combined with atomicMax into a zeroed
// `amax`.  rs == 1 (consecutive rows adjacent in memory): one thread per 4
// rows, one 128-bit load per k (1 KB per wavefront), 8 k in flight.
constexpr int kAmaxChunk = 256;
__global__ __launch_bounds__(NT) void amax_rows4_kernel(const float* __restrict__ X, int64_t R,
                                                        int64_t K, int64_t cs,
                                                        unsigned* __restrict__ amax) {
    const int64_t r = ((int64_t)blockIdx.x * NT + threadIdx.x) * 4;
    if (r >= R) return;
    const int64_t k0 = (int64_t)blockIdx.y * kAmaxChunk;
    const int64_t k1 = k0 + kAmaxChunk < K ? k0 + kAmaxChunk : K;
    unsigned m[4] = {0u, 0u, 0u, 0u};
    int64_t k = k0;
    if (r + 3 < R) {
        for (; k + 8 <= k1; k += 8) {
            float4 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = ld4u_nt(X + r + (k + u) * cs);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                m[0] = max(m[0], __float_as_uint(fabsf(v[u].x)));
                m[1] = max(m[1], __float_as_uint(fabsf(v[u].y)));
                m[2] = max(m[2], __float_as_uint(fabsf(v[u].z)));
                m[3] = max(m[3], __float_as_uint(fabsf(v[u].w)));
            }
        }
    }
    for (; k < k1; ++k)
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (r + e < R) m[e] = max(m[e], __float_as_uint(fabsf(X[r + e + k * cs])));
#pragma unroll
    for (int e = 0; e < 4; ++e)
        if (r + e < R) atomicMax(&amax[r + e], m[e]);
}

// Any strides: one thread per row.
__global__ __launch_bounds__(NT) void amax_rows_any_kernel(const float* __restrict__ X,
                                                           int64_t R, int64_t K, int64_t rs,
                                                           int64_t cs,
                                                           unsigned* __restrict__ amax) {
    const int64_t r = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (r >= R) return;
    const int64_t k0 = (int64_t)blockIdx.y * kAmaxChunk;
    const int64_t k1 = k0 + kAmaxChunk < K ? k0 + kAmaxChunk : K;
    unsigned m = 0;
    for (int64_t k = k0; k < k1; ++k) m = max(m, __float_as_uint(fabsf(X[r * rs + k * cs])));
    atomicMax(&amax[r], m);
}

__global__ __launch_bounds__(NT) void zero_u32_kernel(unsigned* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
        p[i] = 0u;
}

}  // namespace

int64_t split_packed_elems(int64_t R, int64_t K) {
    if (R <= 0 || K <= 0) return 0;
    return 3 * round_up(R, TM) * round_up(K, TK);
}

// The two-plane form (the scaled FP16 pieces, or the two BF16 pieces the
// 3-product form reads).
int64_t split16_packed_elems(int64_t R, int64_t K) {
    if (R <= 0 || K <= 0) return 0;
    return 2 * round_up(R, TM) * round_up(K, TK);
}

omb_status launch_split_pack(int64_t R, int64_t K, const float* X, int64_t rs, int64_t cs,
                             __bf16* out, hipStream_t s, int planes) {
    if (R <= 0 || K <= 0) return OMB_OK;
    const int64_t rtiles = ceil_div(R, TM), nkt = ceil_div(K, TK);
    const int64_t kgroups = ceil_div(nkt, SPLIT_SL);
    if (rtiles > 0x7fffffffLL || kgroups > 65535) {
        set_error("split pack: operand too large for one launch");
        return OMB_EUNSUPPORTED;
    }
    const int k_major = cs == 1 ? 1 : 0;
    if (planes == 2)
        hipLaunchKernelGGL((split_pack_kernel<false, 2>), dim3((unsigned)rtiles, (unsigned)kgroups),
                           dim3(NT), 0, s, X, R, K, rs, cs, nkt, k_major, out, nullptr, nullptr);
    else
        hipLaunchKernelGGL((split_pack_kernel<false, 3>), dim3((unsigned)rtiles, (unsigned)kgroups),
                           dim3(NT), 0, s, X, R, K, rs, cs, nkt, k_major, out, nullptr, nullptr);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

omb_status launch_split_pack_pair(int64_t M, int64_t N, int64_t K, const float* A, int64_t rsA,
                                  int64_t csA, __bf16* ap, const float* Bt, int64_t rsB,
                                  int64_t csB, __bf16* bp, hipStream_t s, int planes) {
    if (K <= 0 || (M <= 0 && N <= 0)) return OMB_OK;
    const int64_t nkt = ceil_div(K, TK), kgroups = ceil_div(nkt, SPLIT_SL);
    const PackOperand a{A, M, rsA, csA, csA == 1 ? 1 : 0, ap, M > 0 ? ceil_div(M, TM) : 0};
    const PackOperand b{Bt, N, rsB, csB, csB == 1 ? 1 : 0, bp, N > 0 ? ceil_div(N, TM) : 0};
    if (a.rtiles + b.rtiles > 0x7fffffffLL || kgroups > 65535) {
        set_error("split pack: operands too large for one launch");
        return OMB_EUNSUPPORTED;
    }
    const dim3 grid((unsigned)(a.rtiles + b.rtiles), (unsigned)kgroups);
    if (planes == 2)
        hipLaunchKernelGGL(split_pack_pair_kernel<2>, grid, dim3(NT), 0, s, a, b, K, nkt);
    else
        hipLaunchKernelGGL(split_pack_pair_kernel<3>, grid, dim3(NT), 0, s, a, b, K, nkt);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

omb_status launch_split16_pack(int64_t R, int64_t K, const float* X, int64_t rs, int64_t cs,
                               __bf16* out, unsigned* amax, int* exps, hipStream_t s) {
    if (R <= 0 || K <= 0) return OMB_OK;
    const int64_t rtiles = ceil_div(R, TM), nkt = ceil_div(K, TK);
    const int64_t kgroups = ceil_div(nkt, SPLIT_SL);
    const int64_t achunks = ceil_div(K, kAmaxChunk);
    if (rtiles > 0x7fffffffLL || kgroups > 65535 || achunks > 65535) {
        set_error("split pack: operand too large for one launch");
        return OMB_EUNSUPPORTED;
    }
    if (cs == 1) {
        hipLaunchKernelGGL(amax_rows_kc_kernel, dim3((unsigned)ceil_div(R, NT / 64)), dim3(NT), 0,
                           s, X, R, K, rs, amax);
    } else {
        hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(R, NT), 1024)),
                           dim3(NT), 0, s, amax, R);
        if (rs == 1)
            hipLaunchKernelGGL(amax_rows4_kernel,
                               dim3((unsigned)ceil_div(R, 4 * NT), (unsigned)achunks), dim3(NT), 0,
                               s, X, R, K, cs, amax);
        else
            hipLaunchKernelGGL(amax_rows_any_kernel,
                               dim3((unsigned)ceil_div(R, NT), (unsigned)achunks), dim3(NT), 0, s,
                               X, R, K, rs, cs, amax);
    }
    OMB_HIP_TRY(hipGetLastError());
    const int k_major = cs == 1 ? 1 : 0;
    hipLaunchKernelGGL((split_pack_kernel<true, 2>), dim3((unsigned)rtiles, (unsigned)kgroups),
                       dim3(NT), 0, s, X, R, K, rs, cs, nkt, k_major, out, amax, exps);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

// FP32-equivalent flop/s of one 128x128 workgroup of the 6-product kernel,
// measured: two per CU (190 TFLOP/s over 512 at 8192^3) and alone on its CU
// (124 TFLOP/s over 256 at 2048^3).
constexpr double kTileRateSplit = 0.37e12;
constexpr double kTileRateSplitAlone = 0.48e12;

int64_t sgemm_split_splits(int64_t M, int64_t N, int64_t K, int64_t* kchunk) {
    return gemm_k_splits(M, N, K, TK * kFoldSlices, kTileRateSplit, kTileRateSplitAlone, kchunk);
}

omb_status launch_sgemm_packed(int64_t M, int64_t N, int64_t K, float alpha, const __bf16* ap,
                               const __bf16* bp, float beta, float* C, int64_t ldc, int nprod,
                               hipStream_t s, float* part, const int* ea, const int* eb,
                               int planes) {
    const int64_t tiles_m = ceil_div(M, TM), tiles_n = ceil_div(N, TN);
    if (tiles_m * tiles_n > 0x7fffffffLL) {
        set_error("sgemm split: too many tiles for one launch");
        return OMB_EUNSUPPORTED;
    }
    if (K <= 0) {  // the kernel loads at least one slice per split
        set_error("sgemm split: K must be positive");
        return OMB_EINVAL;
    }
    const int64_t nkt = ceil_div(K, TK);
    MArgs a{ap, bp, nkt, C, ldc, M, N, alpha, beta, tiles_m, tiles_n, nkt, nullptr, ea, eb};
    // Split K over blockIdx.y when the tiles alone leave CUs idle; the
    // splits' partials are then added in split order.
    int64_t kchunk = K;
    const int64_t splits = sgemm_split_splits(M, N, K, &kchunk);
    if (splits > 1) {
        if (!part) {
            Workspace* ws = nullptr;
            OMB_TRY(get_workspace(s, (size_t)(splits * M * N) * sizeof(float), &ws));
            part = reinterpret_cast<float*>(ws->buf);
        }
        a.ktc = kchunk / TK;
        a.part = part;
    }
    dim3 grid((unsigned)(tiles_m * tiles_n), (unsigned)splits), block(NT);
    if (ea)
        hipLaunchKernelGGL((sgemm_bf16x3_kernel<3, f16x8, 2>), grid, block, 0, s, a);
    else if (nprod == 3 && planes == 2)
        hipLaunchKernelGGL((sgemm_bf16x3_kernel<3, bf16x8, 2>), grid, block, 0, s, a);
    else if (nprod == 3)
        hipLaunchKernelGGL((sgemm_bf16x3_kernel<3, bf16x8, 3>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((sgemm_bf16x3_kernel<6, bf16x8, 3>), grid, block, 0, s, a);
    OMB_HIP_TRY(hipGetLastError());
    if (splits > 1) return launch_splitk_reduce(M, N, splits, alpha, beta, part, C, ldc, s);
    return OMB_OK;
}

omb_status launch_sgemm_split(int64_t M, int64_t N, int64_t K, float alpha,
                              const float* A, int64_t rsA, int64_t csA,
                              const float* B, int64_t rsB, int64_t csB,
                              float beta, float* C, int64_t ldc, int nprod, hipStream_t s) {
    // the 3-product form packs only the two planes it reads
    const int planes = nprod == 3 ? 2 : 3;
    const size_t a_elems = (size_t)(planes == 2 ? split16_packed_elems(M, K)
                                                : split_packed_elems(M, K));
    const size_t b_elems = (size_t)(planes == 2 ? split16_packed_elems(N, K)
                                                : split_packed_elems(N, K));
    // one workspace request for the packed operands and any split partials
    const size_t packed = round_up((int64_t)((a_elems + b_elems) * sizeof(__bf16)), 256);
    int64_t kchunk = K;
    const int64_t splits = sgemm_split_splits(M, N, K, &kchunk);
    const size_t partial = splits > 1 ? (size_t)(splits * M * N) * sizeof(float) : 0;
    Workspace* ws = nullptr;
    OMB_TRY(get_workspace(s, packed + partial, &ws));
    __bf16* ap = reinterpret_cast<__bf16*>(ws->buf);
    __bf16* bp = ap + a_elems;
    float* part = partial ? reinterpret_cast<float*>(static_cast<char*>(ws->buf) + packed) : nullptr;
    // A as rows m, columns k; B^T as rows n, columns k (element (n,k) = B(k,n)).
    OMB_TRY(launch_split_pack_pair(M, N, K, A, rsA, csA, ap, B, csB, rsB, bp, s, planes));
    return launch_sgemm_packed(M, N, K, alpha, ap, bp, beta, C, ldc, nprod, s, part, nullptr,
                               nullptr, planes);
}

// Scaled FP16 split: the exponents (M + N ints) and row maxima (M + N
// unsigned) follow the packed operands in the workspace.
size_t split16_extra_bytes(int64_t M, int64_t N) {
    return (size_t)round_up((int64_t)((M + N) * (sizeof(int) + sizeof(unsigned))), 256);
}

// Workspace layout of one scaled split-FP16 product: the 2-plane packed A
// and B^T, the exponents and maxima, then any split partials.
size_t split16_ws_bytes(int64_t M, int64_t N, int64_t K) {
    const size_t packed = round_up(
        (split16_packed_elems(M, K) + split16_packed_elems(N, K)) * (int64_t)sizeof(__bf16), 256);
    int64_t kchunk = K;
    const int64_t splits = sgemm_split_splits(M, N, K, &kchunk);
    return packed + split16_extra_bytes(M, N) +
           (splits > 1 ? (size_t)(splits * M * N) * sizeof(float) : 0);
}

omb_status launch_sgemm_split16_ws(int64_t M, int64_t N, int64_t K, float alpha,
                                   const float* A, int64_t rsA, int64_t csA,
                                   const float* B, int64_t rsB, int64_t csB,
                                   float beta, float* C, int64_t ldc, char* ws, hipStream_t s) {
    const size_t a_elems = (size_t)split16_packed_elems(M, K);
    const size_t b_elems = (size_t)split16_packed_elems(N, K);
    const size_t packed = round_up((int64_t)((a_elems + b_elems) * sizeof(__bf16)), 256);
    int64_t kchunk = K;
    const int64_t splits = sgemm_split_splits(M, N, K, &kchunk);
    __bf16* ap = reinterpret_cast<__bf16*>(ws);
    __bf16* bp = ap + a_elems;
    int* ea = reinterpret_cast<int*>(ws + packed);
    int* eb = ea + M;
    unsigned* amax_a = reinterpret_cast<unsigned*>(eb + N);
    unsigned* amax_b = amax_a + M;
    float* part =
        splits > 1 ? reinterpret_cast<float*>(ws + packed + split16_extra_bytes(M, N)) : nullptr;
    // A scaled by rows, B^T (rows n) by rows, i.e. B by columns.
    OMB_TRY(launch_split16_pack(M, K, A, rsA, csA, ap, amax_a, ea, s));
    OMB_TRY(launch_split16_pack(N, K, B, csB, rsB, bp, amax_b, eb, s));
    return launch_sgemm_packed(M, N, K, alpha, ap, bp, beta, C, ldc, 3, s, part, ea, eb, 2);
}

omb_status launch_sgemm_split16(int64_t M, int64_t N, int64_t K, float alpha,
                                const float* A, int64_t rsA, int64_t csA,
                                const float* B, int64_t rsB, int64_t csB,
                                float beta, float* C, int64_t ldc, hipStream_t s) {
    Workspace* ws = nullptr;
    OMB_TRY(get_workspace(s, split16_ws_bytes(M, N, K), &ws));
    return launch_sgemm_split16_ws(M, N, K, alpha, A, rsA, csA, B, rsB, csB, beta, C, ldc,
                                   static_cast<char*>(ws->buf), s);
}

}  // namespace omb
