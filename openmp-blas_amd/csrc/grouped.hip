// Grouped execution: thousands of independent GEMMs and GEMVs in ONE
// persistent kernel launch, captured in a hipGraph.  (The reference has no
// batched API; a caller loops over mtm_helper / mtv_helper.)
//
// Work list: every GEMM is cut into tiles of the class the autotuner chose
// for its (log2 M, log2 N, log2 K) bucket, every GEMV into 64-row chunks;
// items are sorted by estimated cost (longest first).  A fixed grid of
// workgroups (CUs x occupancy) pulls chunks of consecutive items from an
// atomic counter; the last workgroup to finish resets the counters (atomics
// only: no cache-flushing fences, no memset node in the graph).
//
// Every tile class keeps the per-element arithmetic of the standalone kernels
// (plain: sequential k FMAs folded every 128 k; split-BF16: the same MFMA
// slices, product order and 256-k folds; scaled split-FP16: the same line
// exponents, pieces and products; GEMV: the same lane order and butterfly),
// so a grouped result is bitwise the one omb_sgemm / omb_sgemv gives for
// that instance alone -- whatever else is in the batch -- with one
// exception: for K > 512 and few output tiles omb_sgemm may split K into
// coarse blocks of several accumulation blocks each, which rounds
// differently; grouped tiles never split K, so such an instance keeps the
// unsplit fold (same error bound, other bits).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "gemv_row.hpp"
#include "host_util.hpp"
#include "split_common.hpp"

namespace omb {
namespace {

struct DGemm {
    int64_t M, N, K;
    const float* A;
    int64_t lda;
    const float* B;
    int64_t ldb;
    float* C;
    int64_t ldc;
    float alpha, beta;
    int* ea;  // scaled FP16 split: exponent of every row of A, every column of B
    int* eb;
};

// Exponent work of the scaled FP16 split: rows start .. start+31 of A (8
// per wavefront) or columns start .. start+63 of B (a lane per column, the
// k range shared by the 4 wavefronts).
constexpr int kExpRows = 32, kExpCols = 64;
struct ExpItem {
    int inst, col;
    int64_t start;
};

struct DGemv {
    int64_t M, N;
    const float* A;
    int64_t lda;
    const float* x;
    float* y;
    float alpha, beta;
};

constexpr int kTileM[NCLS] = {32, 64, 128, 64, 64};
constexpr int kTileN[NCLS] = {32, 64, 64, 64, 1};

struct Item {
    int cls, inst, tm, tn;
};

struct PArgs {
    const Item* items;
    int nitems;
    const DGemm* gemms;
    const DGemv* gemvs;
    unsigned* counters;  // [0] next dynamic chunk, [1] workgroups done
    int nprod;
    int f16;      // split tiles take the scaled FP16 pieces (3 products)
    int chunk;    // consecutive items per chunk
    int nchunks;  // ceil(nitems / chunk)
};

constexpr int NT = 256;
constexpr int kFoldK = 128;  // k per inner accumulation block (as sgemm_f32)

// threadIdx.x through an empty asm, read at the start of every item: index
// math derived from it cannot be hoisted out of the persistent kernel's item
// loop, where it would hold registers for every item type across all items.
__device__ __forceinline__ int item_tid() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// ------------------------------------------------------------ plain tile --
// BM x BN tile of C = alpha*A*B + beta*C (all row-major); 256 threads on an
// H x W = (BM/TMR) x (BN/TNR) grid.  Thread (ty, tx) owns GR fragments of FR
// consecutive rows, at rows g*(BM/GR) + ty*FR + i, and GC fragments of FC
// consecutive columns (FR = min(TMR, 4), GR = TMR / FR; likewise for
// columns), so every LDS read is one 64- or 128-bit access.  Slices of 16 k
// are loaded with 128-bit global loads along the unit-stride index, staged
// through two LDS buffers, and the next D slices are prefetched into D
// register sets; LDS fragments for k+1 are read while k's FMAs run.  Every
// element sums k in order and folds the running sum every 128 k, as
// sgemm_f32_kernel does, so the results are the same bits.
template <int N>
struct FragT;
template <>
struct FragT<2> {
    typedef float2 type;
};
template <>
struct FragT<4> {
    typedef float4 type;
};

template <typename P>
__device__ __forceinline__ float4 load4_guard(P p, int n) {
    if (n == 4) return ld4u(p);
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (n > 0) v.x = p[0];
    if (n > 1) v.y = p[1];
    if (n > 2) v.z = p[2];
    return v;
}

__device__ __forceinline__ int clamp4i(int64_t r) { return r <= 0 ? 0 : (r >= 4 ? 4 : (int)r); }

template <int BM, int BN, int TMR, int TNR, int D>
__device__ __forceinline__ void tile_f32(const DGemm& g, int tm, int tn, float* smem) {
    constexpr int BK = 16;
    constexpr int H = BM / TMR, W = BN / TNR;
    static_assert(H * W == NT, "thread grid must cover the workgroup");
    constexpr int FR = TMR < 4 ? TMR : 4, GR = TMR / FR;
    constexpr int FC = TNR < 4 ? TNR : 4, GC = TNR / FC;
    constexpr int kFold = kFoldK / BK;
    constexpr int PA = BM + 4, PB = BN + 4;  // padded LDS rows, 16-byte aligned
    constexpr int NA4 = BM * BK / 4, NB4 = BN * BK / 4;  // float4 per slice
    constexpr int LA = (NA4 + NT - 1) / NT, LB = (NB4 + NT - 1) / NT;
    typedef typename FragT<FR>::type FragA;
    typedef typename FragT<FC>::type FragB;
    // Two LDS buffers of [BK][PA] A and [BK][PB] B: slice k+1 is stored into
    // one while slice k is read from the other (one barrier per slice).
    constexpr int SZ = BK * (PA + PB);
    const int tid = item_tid(), tx = tid % W, ty = tid / W;
    const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
    const int64_t M = g.M, N = g.N, K = g.K;
    gcfloat* const gA = as_global(g.A);
    gcfloat* const gB = as_global(g.B);
    gfloat* const gC = as_global(g.C);

    float4 ra[D][LA], rb[D][LB];  // D slices of loads in flight, one set each
    auto load = [&](int set, int64_t k0) {
#pragma unroll
        for (int i = 0; i < LA; ++i) {
            const int e = tid + i * NT;
            if (e < NA4) {  // A[m][k]: float4 along k
                const int row = e >> 2, kq = (e & 3) * 4;
                const int64_t m = m0 + row, k = k0 + kq;
                const int n = m < M ? clamp4i(K - k) : 0;
                ra[set][i] = load4_guard(gA + (n ? m * g.lda + k : 0), n);
            }
        }
#pragma unroll
        for (int i = 0; i < LB; ++i) {
            const int e = tid + i * NT;
            if (e < NB4) {  // B[k][n]: float4 along n
                const int kk = e / (BN / 4), nq = (e % (BN / 4)) * 4;
                const int64_t nn = n0 + nq, k = k0 + kk;
                const int n = k < K ? clamp4i(N - nn) : 0;
                rb[set][i] = load4_guard(gB + (n ? k * g.ldb + nn : 0), n);
            }
        }
    };
    auto store = [&](int set, int buf) {
        float* As = smem + buf * SZ;
        float* Bs = As + BK * PA;
#pragma unroll
        for (int i = 0; i < LA; ++i) {
            const int e = tid + i * NT;
            if (e < NA4) {
                const int row = e >> 2, kq = (e & 3) * 4;
                As[(kq + 0) * PA + row] = ra[set][i].x;
                As[(kq + 1) * PA + row] = ra[set][i].y;
                As[(kq + 2) * PA + row] = ra[set][i].z;
                As[(kq + 3) * PA + row] = ra[set][i].w;
            }
        }
#pragma unroll
        for (int i = 0; i < LB; ++i) {
            const int e = tid + i * NT;
            if (e < NB4) {
                const int kk = e / (BN / 4), nq = (e % (BN / 4)) * 4;
                *reinterpret_cast<float4*>(&Bs[kk * PB + nq]) = rb[set][i];
            }
        }
    };

    float inner[TMR][TNR], outer[TMR][TNR];
#pragma unroll
    for (int i = 0; i < TMR; ++i)
#pragma unroll
        for (int j = 0; j < TNR; ++j) inner[i][j] = outer[i][j] = 0.f;

    const int64_t nk = ceil_div(K, BK);
    auto compute = [&](int buf, int64_t kt) {
        const float* As = smem + buf * SZ;
        const float* Bs = As + BK * PA;
        FragA fa[2][GR];
        FragB fb[2][GC];
        auto fragment = [&](int k, int set) {
#pragma unroll
            for (int q = 0; q < GR; ++q)
                fa[set][q] = *reinterpret_cast<const FragA*>(&As[k * PA + q * (BM / GR) + ty * FR]);
#pragma unroll
            for (int q = 0; q < GC; ++q)
                fb[set][q] = *reinterpret_cast<const FragB*>(&Bs[k * PB + q * (BN / GC) + tx * FC]);
        };
        fragment(0, 0);
#pragma unroll
        for (int k = 0; k < BK; ++k) {
            if (k + 1 < BK) fragment(k + 1, (k + 1) & 1);
            const float* av = reinterpret_cast<const float*>(fa[k & 1]);
            const float* bv = reinterpret_cast<const float*>(fb[k & 1]);
#pragma unroll
            for (int i = 0; i < TMR; ++i)
#pragma unroll
                for (int j = 0; j < TNR; ++j) inner[i][j] = fmaf(av[i], bv[j], inner[i][j]);
        }
        if ((kt + 1) % kFold == 0 || kt + 1 == nk) {
#pragma unroll
            for (int i = 0; i < TMR; ++i)
#pragma unroll
                for (int j = 0; j < TNR; ++j) {
                    outer[i][j] += inner[i][j];
                    inner[i][j] = 0.f;
                }
        }
    };
    // Slices 0 .. D-1 are loaded up front; step kt loads slice kt + D into
    // the set that held slice kt (stored to LDS in step kt - 1), computes
    // slice kt from LDS buffer kt % 2 and stores slice kt + 1 into the other
    // buffer, so each load has D - 1 steps of compute to arrive in.
#pragma unroll
    for (int s = 0; s < D; ++s)
        if (s == 0 || s < nk) load(s, s * BK);  // slice 0 even when K = 0 (zeros)
    __syncthreads();  // the previous item is done with LDS
    store(0, 0);
    __syncthreads();
    for (int64_t kt0 = 0; kt0 < nk; kt0 += D) {
#pragma unroll
        for (int s = 0; s < D; ++s) {
            const int64_t kt = kt0 + s;
            if (kt >= nk) break;
            if (kt + D < nk) load(s, (kt + D) * BK);
            compute((int)(kt & 1), kt);
            // The other buffer was last read in step kt - 1, which every
            // thread finished before the barrier that ended it.
            if (kt + 1 < nk) store((s + 1) % D, (int)((kt + 1) & 1));
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < TMR; ++i) {
        const int64_t m = m0 + (i / FR) * (BM / GR) + ty * FR + (i % FR);
        if (m >= M) continue;
#pragma unroll
        for (int j = 0; j < TNR; ++j) {
            const int64_t n = n0 + (j / FC) * (BN / GC) + tx * FC + (j % FC);
            if (n < N) {
                gfloat* c = gC + m * g.ldc + n;
                const float v = g.alpha * outer[i][j];
                *c = g.beta != 0.f ? v + g.beta * *c : v;
            }
        }
    }
}

// ------------------------------------------------------------ split tile --
// 64x64 tile on the MFMA cores: each 16-k slice of A (row-major) and B^T
// (B row-major) is read as FP32, split into the three BF16 planes and
// written into the swizzled LDS slice; 4 wavefronts own 32x32 each.
// The 4 consecutive-k values of A (row m, from k ka) and of B (column n,
// from k kb) that one loader thread stages per 16-k slice; elements out of
// range read as 0.  Edge loads read a clamped in-range address and select,
// so every load of the slice is in flight at once (A and B are non-null
// whenever a tile has K > 0).
template <typename P>
__device__ __forceinline__ void load_split_slice(P A, int64_t lda, int64_t M,
                                                 P B, int64_t ldb, int64_t N,
                                                 int64_t K, int64_t m, int64_t ka, int64_t n,
                                                 int64_t kb, float (&va)[4], float (&vb)[4]) {
    if (m < M && ka + 3 < K) {
        const float4 t = ld4u(A + m * lda + ka);
        va[0] = t.x; va[1] = t.y; va[2] = t.z; va[3] = t.w;
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const bool ok = m < M && ka + e < K;
            const float t = A[ok ? m * lda + ka + e : 0];
            va[e] = ok ? t : 0.f;
        }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const bool ok = n < N && kb + e < K;
        const float t = B[ok ? (kb + e) * ldb + n : 0];
        vb[e] = ok ? t : 0.f;
    }
}

template <int NPROD>
__device__ __forceinline__ void tile_split(const DGemm& g, int tm, int tn, uint4* smem) {
    using namespace split;
    constexpr int P = 64;  // rows per plane for A and for B^T
    const int tid = item_tid(), lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int frow = lane & 31, fh = lane >> 5;
    const int64_t m0 = (int64_t)tm * P, n0 = (int64_t)tn * P;
    const int64_t M = g.M, N = g.N, K = g.K;
    gcfloat* const A = as_global(g.A);
    gcfloat* const B = as_global(g.B);
    const int64_t lda = g.lda, ldb = g.ldb;
    // loader roles: A row ra, k quarter qa; B^T row rb (= column of B), k quarter qb
    const int ra = tid >> 2, qa = tid & 3;
    const int rb = tid & 63, qb = tid >> 6;
    uint2* l2 = reinterpret_cast<uint2*>(smem);

    f32x16 inner[1][1], outer[1][1];
    inner[0][0] = outer[0][0] = f32x16{};
    const int64_t nkt = ceil_div(K, TK);
    for (int64_t kt = 0; kt < nkt; ++kt) {
        const int64_t k0 = kt * TK;
        float va[4], vb[4];
        load_split_slice(A, lda, M, B, ldb, N, K, m0 + ra, k0 + 4 * qa, n0 + rb, k0 + 4 * qb, va, vb);
        __syncthreads();  // previous slice consumed
        {
            typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
            bf16x4 pa[3], pb[3];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                __bf16 t0, t1, t2;
                split3(va[e], t0, t1, t2);
                pa[0][e] = t0; pa[1][e] = t1; pa[2][e] = t2;
                split3(vb[e], t0, t1, t2);
                pb[0][e] = t0; pb[1][e] = t1; pb[2][e] = t2;
            }
#pragma unroll
            for (int pl = 0; pl < 3; ++pl) {
                l2[2 * (pl * 2 * P + lds_chunk(ra, qa >> 1)) + (qa & 1)] =
                    *reinterpret_cast<const uint2*>(&pa[pl]);
                l2[2 * (3 * 2 * P + pl * 2 * P + lds_chunk(rb, qb >> 1)) + (qb & 1)] =
                    *reinterpret_cast<const uint2*>(&pb[pl]);
            }
        }
        __syncthreads();
        mfma_slice<NPROD, 1, 1, P, P>(smem, wm * 32, wn * 32, frow, fh, inner);
        if ((kt + 1) % kFoldSlices == 0 || kt + 1 == nkt) fold(inner, outer);
    }
    store_tile(outer, m0 + wm * 32, n0 + wn * 32, lane, as_global(g.C), g.ldc, M, N, g.alpha,
               g.beta);
}

// --------------------------------------------------- scaled FP16 split --
// Line exponents e = 14 - ilogb(max |x|) (0 for a zero, infinite or NaN
// line), one per row of A and column of B.  The maximum is taken over the
// bits of |x| (unsigned order is the numeric order for non-negative floats;
// NaN's bits exceed infinity's) and is order-free, so the exponents are the
// ones omb_sgemm's maxima passes give.
__global__ __launch_bounds__(NT) void grouped_exps_kernel(const ExpItem* __restrict__ items,
                                                          const DGemm* __restrict__ gemms) {
    __shared__ unsigned red[NT];
    const ExpItem w = items[blockIdx.x];
    const DGemm& g = gemms[w.inst];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int RW = kExpRows / (NT / 64);  // rows per wavefront
    if (w.col == 0) {
        const int64_t m0 = w.start + wave * RW;
        unsigned mx[RW];
#pragma unroll
        for (int r = 0; r < RW; ++r) mx[r] = 0;
        for (int64_t k = lane; k < g.K; k += 64) {
#pragma unroll
            for (int r = 0; r < RW; ++r)  // RW independent loads in flight
                if (m0 + r < g.M) {
                    const unsigned b = __float_as_uint(g.A[(m0 + r) * g.lda + k]) & 0x7fffffffu;
                    mx[r] = b > mx[r] ? b : mx[r];
                }
        }
#pragma unroll
        for (int r = 0; r < RW; ++r) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const unsigned o = (unsigned)__shfl_xor((int)mx[r], off);
                mx[r] = o > mx[r] ? o : mx[r];
            }
            if (lane == r && m0 + r < g.M) g.ea[m0 + r] = split::line_exp(__uint_as_float(mx[r]));
        }
    } else {
        const int64_t n = w.start + lane;
        unsigned mx = 0;
        if (n < g.N) {
#pragma unroll 8
            for (int64_t k = wave; k < g.K; k += NT / 64) {
                const unsigned b = __float_as_uint(g.B[k * g.ldb + n]) & 0x7fffffffu;
                mx = b > mx ? b : mx;
            }
        }
        red[threadIdx.x] = mx;
        __syncthreads();
        if (wave == 0 && n < g.N) {
#pragma unroll
            for (int q = 1; q < NT / 64; ++q) mx = red[q * 64 + lane] > mx ? red[q * 64 + lane] : mx;
            g.eb[n] = split::line_exp(__uint_as_float(mx));
        }
    }
}

// The split tile with the scaled FP16 pieces: every element is scaled by
// its row (A) or column (B) exponent and split into two FP16 pieces as
// omb_sgemm's packing pass does; the 3 leading products (a0 b1, a1 b0,
// a0 b0) run on the FP16 MFMA, and the sums are scaled back when stored.
__device__ __forceinline__ void tile_split16(const DGemm& g, int tm, int tn, uint4* smem) {
    using namespace split;
    constexpr int P = 64;
    const int tid = item_tid(), lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int frow = lane & 31, fh = lane >> 5;
    const int64_t m0 = (int64_t)tm * P, n0 = (int64_t)tn * P;
    const int64_t M = g.M, N = g.N, K = g.K;
    gcfloat* const A = as_global(g.A);
    gcfloat* const B = as_global(g.B);
    const int64_t lda = g.lda, ldb = g.ldb;
    const int ra = tid >> 2, qa = tid & 3;
    const int rb = tid & 63, qb = tid >> 6;
    uint2* l2 = reinterpret_cast<uint2*>(smem);
    const int ex_a = m0 + ra < M ? g.ea[m0 + ra] : 0;
    const int ex_b = n0 + rb < N ? g.eb[n0 + rb] : 0;

    f32x16 inner[1][1], outer[1][1];
    inner[0][0] = outer[0][0] = f32x16{};
    const int64_t nkt = ceil_div(K, TK);
    for (int64_t kt = 0; kt < nkt; ++kt) {
        const int64_t k0 = kt * TK;
        float va[4], vb[4];
        load_split_slice(A, lda, M, B, ldb, N, K, m0 + ra, k0 + 4 * qa, n0 + rb, k0 + 4 * qb, va, vb);
        __syncthreads();  // previous slice consumed
        {
            typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
            u16x4 pa[2], pb[2];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                _Float16 h0, h1;
                split2_f16(ldexpf(va[e], ex_a), h0, h1);
                pa[0][e] = __builtin_bit_cast(unsigned short, h0);
                pa[1][e] = __builtin_bit_cast(unsigned short, h1);
                split2_f16(ldexpf(vb[e], ex_b), h0, h1);
                pb[0][e] = __builtin_bit_cast(unsigned short, h0);
                pb[1][e] = __builtin_bit_cast(unsigned short, h1);
            }
#pragma unroll
            for (int pl = 0; pl < 2; ++pl) {
                l2[2 * (pl * 2 * P + lds_chunk(ra, qa >> 1)) + (qa & 1)] =
                    __builtin_bit_cast(uint2, pa[pl]);
                l2[2 * (2 * 2 * P + pl * 2 * P + lds_chunk(rb, qb >> 1)) + (qb & 1)] =
                    __builtin_bit_cast(uint2, pb[pl]);
            }
        }
        __syncthreads();
        mfma_slice<3, 1, 1, P, P, f16x8, 2>(smem, wm * 32, wn * 32, frow, fh, inner);
        if ((kt + 1) % kFoldSlices == 0 || kt + 1 == nkt) fold(inner, outer);
    }
    store_tile(outer, m0 + wm * 32, n0 + wn * 32, lane, as_global(g.C), g.ldc, M, N, g.alpha,
               g.beta, as_global(static_cast<const int*>(g.ea)),
               as_global(static_cast<const int*>(g.eb)));
}

// ------------------------------------------------------------- GEMV item --
__device__ __forceinline__ void gemv_item(const DGemv& v, int chunk) {
    const int tid = item_tid();
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // wavefront w takes rows chunk*64 + 16w .. +15 in one pass (16 rows of
    // 128-bit loads in flight per lane); rows under 2^29 columns use scalar
    // row bases and 32-bit lane offsets (GEMV-only C4 2.63 -> 2.61 ms)
    const int64_t r0 = (int64_t)chunk * 64 + wave * 16;
    if (r0 >= v.M) return;
    const int nr = (int)(v.M - r0 < 16 ? v.M - r0 : 16);
    if (v.N < kGemvOff32Cols)
        gemv_rows_wave<16, 1, false, true>(r0, nr, v.N, v.alpha, as_global(v.A), v.lda,
                                           as_global(v.x), v.beta, as_global(v.y), lane);
    else
        gemv_rows_wave<16, 1, false>(r0, nr, v.N, v.alpha, as_global(v.A), v.lda, as_global(v.x), v.beta,
                                     as_global(v.y), lane);
}

// ------------------------------------------------------- persistent kernel --
// LDS: the largest plain class footprint (128x64: two buffers of
// 16 x (132 + 68) floats); the split tile uses 12 KB.
constexpr int kSmemFloats = 2 * (16 * (128 + 4) + 16 * (64 + 4)) + 16;

// Slices of global loads in flight for the 32x32 and 64x64 plain tiles (8
// VGPRs each).
constexpr int kPrefetch = 4;

// The lite build is bounded to 4 workgroups per CU (at most 128 VGPRs): the
// C4 plain plan measured 3.48 ms so, against 3.55-3.63 ms for builds of 129
// and 165 VGPRs (3 per CU) that spent the registers on deeper split-tile
// prefetch.  The big build (128x64 tiles) is bounded to 2.
template <bool BIG>
__global__ __launch_bounds__(NT, BIG ? 2 : 4) void grouped_kernel(PArgs p) {
    __shared__ __attribute__((aligned(16))) float smem[kSmemFloats];
    __shared__ int s_chunk;
    // Chunks of consecutive (cost-sorted) items: workgroup b starts with
    // chunk b, then takes chunks gridDim.x, gridDim.x + 1, ... from one
    // device-scope counter.  Such atomics on one address serialize (about
    // 70 ns each measured here), so one ticket per chunk instead of one per
    // item keeps the counter off the critical path.
    int c = blockIdx.x;
    for (;;) {
        if (c >= p.nchunks) break;
        const int beg = c * p.chunk, end = min(beg + p.chunk, p.nitems);
        for (int it = beg; it < end; ++it) {
            __syncthreads();  // the previous item is done with LDS
            Item w = p.items[it];
            w.cls = __builtin_amdgcn_readfirstlane(w.cls);
            w.inst = __builtin_amdgcn_readfirstlane(w.inst);
            w.tm = __builtin_amdgcn_readfirstlane(w.tm);
            w.tn = __builtin_amdgcn_readfirstlane(w.tn);
            switch (w.cls) {
                case C32:
                    tile_f32<32, 32, 2, 2, kPrefetch>(p.gemms[w.inst], w.tm, w.tn, smem);
                    break;
                case C64:
                    tile_f32<64, 64, 4, 4, kPrefetch>(p.gemms[w.inst], w.tm, w.tn, smem);
                    break;
                case C128x64:
                    if (BIG) tile_f32<128, 64, 8, 4, 1>(p.gemms[w.inst], w.tm, w.tn, smem);
                    break;
                case CSPLIT:
                    if (p.f16)
                        tile_split16(p.gemms[w.inst], w.tm, w.tn, reinterpret_cast<uint4*>(smem));
                    else if (p.nprod == 3)
                        tile_split<3>(p.gemms[w.inst], w.tm, w.tn, reinterpret_cast<uint4*>(smem));
                    else
                        tile_split<6>(p.gemms[w.inst], w.tm, w.tn, reinterpret_cast<uint4*>(smem));
                    break;
                case CGEMV: gemv_item(p.gemvs[w.inst], w.tm); break;
                default: break;
            }
        }
        __syncthreads();  // every thread has read s_chunk for this chunk
        if (threadIdx.x == 0) s_chunk = (int)gridDim.x + (int)atomicAdd(&p.counters[0], 1u);
        __syncthreads();
        // uniform across the workgroup: keep it (and the descriptors) scalar
        c = __builtin_amdgcn_readfirstlane(s_chunk);
    }
    if (threadIdx.x == 0 && atomicAdd(&p.counters[1], 1u) == gridDim.x - 1) {
        atomicExch(&p.counters[0], 0u);
        atomicExch(&p.counters[1], 0u);
    }
}

// ---------------------------------------------------------- tuning table --
using Key = TuneKey;
std::mutex g_tune_mu;
std::map<Key, int>& tune_table() {
    static auto* t = new std::map<Key, int>();
    return *t;
}

int lg(int64_t x) {
    int b = 0;
    while (x > 1) {
        x >>= 1;
        ++b;
    }
    return b;
}

Key key_of(int mode, const omb_gemm_desc& d) { return Key{mode, lg(d.M), lg(d.N), lg(d.K)}; }

// Class without a tuning entry: split mode has one class; plain takes 32x32
// tiles when the smaller output dimension is under 48, else 64x64.  (The
// 128x64 class needs the larger-register kernel, 2 workgroups per CU instead
// of 3, which slows every other item of a mixed plan; at C4's shapes 64x64
// gives the same total rate, so 128x64 is used only when a loaded tuning
// table names it.)
int default_class(int mode, const omb_gemm_desc& d) {
    if (mode != OMB_MODE_PLAIN_F32) return CSPLIT;
    return std::min(d.M, d.N) < 48 ? C32 : C64;
}

// The table's class for the bucket, if it computes this mode (the loader
// accepts only such entries), else the default.
int choose_class(int mode, const omb_gemm_desc& d) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = tune_table().find(key_of(mode, d));
    return it != tune_table().end() && class_serves_mode(it->second, mode)
               ? it->second
               : default_class(mode, d);
}

}  // namespace
}  // namespace omb

// ------------------------------------------------------------------ plans --
struct omb_group_plan {
    int dev = 0;
    int mode = 0;
    bool big = false;
    int grid = 0;
    int nitems = 0;
    int chunk = 1;
    omb::DGemm* d_gemms = nullptr;
    omb::DGemv* d_gemvs = nullptr;
    omb::Item* d_items = nullptr;
    unsigned* d_counters = nullptr;
    int nexp = 0;                      // scaled FP16 split: exponent work items
    omb::ExpItem* d_exp_items = nullptr;
    int* d_exps = nullptr;             // every GEMM's row and column exponents
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::string text;
};

namespace omb {
namespace {

omb_status validate(const omb_gemm_desc* gemms, int32_t n_gemm, const omb_gemv_desc* gemvs,
                    int32_t n_gemv) {
    if (n_gemm < 0 || n_gemv < 0 || (n_gemm > 0 && !gemms) || (n_gemv > 0 && !gemvs)) {
        set_error("group plan: bad descriptor array");
        return OMB_EINVAL;
    }
    for (int i = 0; i < n_gemm; ++i) {
        const omb_gemm_desc& d = gemms[i];
        const bool empty = d.M == 0 || d.N == 0;
        if (d.M < 0 || d.N < 0 || d.K < 0 || (!empty && (!d.C || d.ldc < d.N)) ||
            (!empty && d.K > 0 && (!d.A || !d.B || d.lda < d.K || d.ldb < d.N))) {
            set_error("group plan: invalid GEMM descriptor " + std::to_string(i));
            return OMB_EINVAL;
        }
    }
    for (int i = 0; i < n_gemv; ++i) {
        const omb_gemv_desc& d = gemvs[i];
        if (d.M < 0 || d.N < 0 || (d.M > 0 && !d.y) ||
            (d.M > 0 && d.N > 0 && (!d.A || !d.x || d.lda < d.N))) {
            set_error("group plan: invalid GEMV descriptor " + std::to_string(i));
            return OMB_EINVAL;
        }
    }
    return OMB_OK;
}

// Estimated seconds of one workgroup on one CU (for longest-first order).
// Estimated seconds one workgroup spends on an item (for the longest-first
// order): per-workgroup rates measured on MI355X with 3 workgroups per CU --
// 32x32 16 GFLOP/s, 64x64 40, 128x64 59, split 64x64 54 (FP32-equivalent),
// GEMV 8 GB/s -- plus about 2 us of fixed cost per item.
double item_cost(const Item& it, const std::vector<DGemm>& g, const std::vector<DGemv>& v) {
    constexpr double kFixed = 2.0e-6;
    if (it.cls == CGEMV) return kFixed + 64.0 * (double)v[it.inst].N * 4.0 / 8.0e9;
    static const double rate[NCLS] = {16e9, 40e9, 59e9, 54e9, 0.0};
    return kFixed + 2.0 * kTileM[it.cls] * kTileN[it.cls] *
                        (double)std::max<int64_t>(g[it.inst].K, 1) / rate[it.cls];
}

template <typename T>
omb_status upload(T** dst, const std::vector<T>& src) {
    *dst = nullptr;
    if (src.empty()) return OMB_OK;
    OMB_HIP_TRY(hipMalloc(reinterpret_cast<void**>(dst), src.size() * sizeof(T)));
    OMB_HIP_TRY(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return OMB_OK;
}

void free_plan(omb_group_plan* p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->d_gemms) (void)hipFree(p->d_gemms);
    if (p->d_gemvs) (void)hipFree(p->d_gemvs);
    if (p->d_items) (void)hipFree(p->d_items);
    if (p->d_counters) (void)hipFree(p->d_counters);
    if (p->d_exp_items) (void)hipFree(p->d_exp_items);
    if (p->d_exps) (void)hipFree(p->d_exps);
    delete p;
}

omb_status launch_plan(omb_group_plan* p, hipStream_t s) {
    if (p->nitems == 0) return OMB_OK;
    if (p->nexp > 0) {  // exponents first (same stream: the tiles read them)
        hipLaunchKernelGGL(grouped_exps_kernel, dim3(p->nexp), dim3(NT), 0, s, p->d_exp_items,
                           p->d_gemms);
        OMB_HIP_TRY(hipGetLastError());
    }
    const bool f16 = p->mode == OMB_MODE_SPLIT_FP16;
    PArgs a{p->d_items, p->nitems, p->d_gemms, p->d_gemvs, p->d_counters,
            p->mode == OMB_MODE_SPLIT_BF16_P3 || f16 ? 3 : 6, f16 ? 1 : 0, p->chunk,
            (int)ceil_div(p->nitems, p->chunk)};
    if (p->big)
        hipLaunchKernelGGL(grouped_kernel<true>, dim3(p->grid), dim3(NT), 0, s, a);
    else
        hipLaunchKernelGGL(grouped_kernel<false>, dim3(p->grid), dim3(NT), 0, s, a);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

// Build a plan; force_cls >= 0 puts every GEMM in that class (autotuner).
omb_status build_plan(const omb_gemm_desc* gemms, int32_t n_gemm, const omb_gemv_desc* gemvs,
                      int32_t n_gemv, int32_t mode, int force_cls, bool capture,
                      omb_group_plan** out) {
    *out = nullptr;
    if (mode != OMB_MODE_PLAIN_F32 && mode != OMB_MODE_SPLIT_BF16 &&
        mode != OMB_MODE_SPLIT_BF16_P3 && mode != OMB_MODE_SPLIT_FP16) {
        set_error("group plan: unknown mode " + std::to_string(mode));
        return OMB_EINVAL;
    }
    OMB_TRY(validate(gemms, n_gemm, gemvs, n_gemv));
    auto* p = new omb_group_plan();
    p->mode = mode;
    std::vector<DGemm> g(n_gemm);
    std::vector<DGemv> v(n_gemv);
    std::vector<Item> items;
    int per_cls[NCLS] = {0, 0, 0, 0, 0};
    std::map<std::string, int> chosen;
    for (int i = 0; i < n_gemm; ++i) {
        const omb_gemm_desc& d = gemms[i];
        g[i] = DGemm{d.M, d.N, d.K, d.A, d.lda, d.B, d.ldb, d.C, d.ldc, d.alpha, d.beta,
                     nullptr, nullptr};
        if (d.M == 0 || d.N == 0) continue;
        const int c = force_cls >= 0 ? force_cls : choose_class(mode, d);
        const int64_t tmn = ceil_div(d.M, kTileM[c]), tnn = ceil_div(d.N, kTileN[c]);
        if (tmn > 0x7fffffff || tnn > 0x7fffffff || (int64_t)items.size() + tmn * tnn > 0x7fffffff) {
            free_plan(p);
            set_error("group plan: too many tiles");
            return OMB_EUNSUPPORTED;
        }
        for (int64_t a = 0; a < tmn; ++a)
            for (int64_t b = 0; b < tnn; ++b) items.push_back(Item{c, i, (int)a, (int)b});
        per_cls[c] += (int)(tmn * tnn);
        if (c == C128x64) p->big = true;
        const Key k = key_of(mode, d);
        chosen["gemm:" + std::to_string(1 << k.bm) + "x" + std::to_string(1 << k.bn) + "x" +
               std::to_string(1 << k.bk) + "=" + kClsName[c]] += 1;
    }
    for (int i = 0; i < n_gemv; ++i) {
        const omb_gemv_desc& d = gemvs[i];
        v[i] = DGemv{d.M, d.N, d.A, d.lda, d.x, d.y, d.alpha, d.beta};
        for (int64_t c = 0; c < ceil_div(d.M, 64); ++c) items.push_back(Item{CGEMV, i, (int)c, 0});
        per_cls[CGEMV] += (int)ceil_div(d.M, 64);
    }
    std::vector<double> cost(items.size());
    for (size_t i = 0; i < items.size(); ++i) cost[i] = item_cost(items[i], g, v);
    std::vector<size_t> order(items.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    std::vector<Item> sorted(items.size());
    for (size_t i = 0; i < order.size(); ++i) sorted[i] = items[order[i]];
    p->nitems = (int)sorted.size();

    omb_status st = OMB_OK;
    if (mode == OMB_MODE_SPLIT_FP16) {
        // one exponent per row of A and column of B, in a plan-owned buffer
        std::vector<ExpItem> ex;
        int64_t lines = 0;
        for (int i = 0; i < n_gemm; ++i)
            if (g[i].M > 0 && g[i].N > 0) lines += g[i].M + g[i].N;
        if (lines > 0 && hipMalloc(&p->d_exps, lines * sizeof(int)) != hipSuccess) {
            (void)hipGetLastError();
            free_plan(p);
            set_error("group plan: exponent allocation failed");
            return OMB_ENOMEM;
        }
        int64_t off = 0;
        for (int i = 0; i < n_gemm; ++i) {
            if (g[i].M == 0 || g[i].N == 0) continue;
            g[i].ea = p->d_exps + off;
            g[i].eb = p->d_exps + off + g[i].M;
            off += g[i].M + g[i].N;
            for (int64_t r = 0; r < g[i].M; r += kExpRows) ex.push_back(ExpItem{i, 0, r});
            for (int64_t c = 0; c < g[i].N; c += kExpCols) ex.push_back(ExpItem{i, 1, c});
        }
        if (ex.size() > 0x7fffffff) {
            free_plan(p);
            set_error("group plan: too many exponent items");
            return OMB_EUNSUPPORTED;
        }
        p->nexp = (int)ex.size();
        if ((st = upload(&p->d_exp_items, ex)) != OMB_OK) {
            free_plan(p);
            return st;
        }
    }
    if ((st = upload(&p->d_gemms, g)) != OMB_OK || (st = upload(&p->d_gemvs, v)) != OMB_OK ||
        (st = upload(&p->d_items, sorted)) != OMB_OK) {
        free_plan(p);
        return st;
    }
    if (hipMalloc(&p->d_counters, 2 * sizeof(unsigned)) != hipSuccess ||
        hipMemset(p->d_counters, 0, 2 * sizeof(unsigned)) != hipSuccess) {
        (void)hipGetLastError();
        free_plan(p);
        set_error("group plan: counter allocation failed");
        return OMB_ENOMEM;
    }
    int dev = 0, cus = 0, per_cu = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        free_plan(p);
        set_error("group plan: device query failed");
        return OMB_EDEVICE;
    }
    cus = prop.multiProcessorCount;
    hipError_t e = p->big ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, grouped_kernel<true>, NT, 0)
                          : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, grouped_kernel<false>, NT, 0);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    p->dev = dev;
    p->grid = std::max(1, std::min(p->nitems, cus * per_cu));
    // up to 4 items per chunk (about 8 chunks per workgroup at most): 4x
    // fewer tickets; longer chunks of the cost-sorted list measured slower
    // on mixed GEMM + GEMV plans
    p->chunk = std::max(1, std::min(4, p->nitems / (8 * p->grid)));

    if (capture && p->nitems > 0) {
        hipStream_t cs = nullptr;
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
            free_plan(p);
            set_error("group plan: stream creation failed");
            return OMB_EDEVICE;
        }
        bool ok = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess;
        ok = ok && launch_plan(p, cs) == OMB_OK;
        ok = hipStreamEndCapture(cs, &p->graph) == hipSuccess && ok;
        ok = ok && hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) == hipSuccess;
        (void)hipStreamDestroy(cs);
        if (!ok) {
            (void)hipGetLastError();
            free_plan(p);
            set_error("group plan: graph capture failed");
            return OMB_EDEVICE;
        }
    }
    std::ostringstream os;
    os << "items=" << p->nitems << " grid=" << p->grid << " chunk=" << p->chunk
       << " kernel=" << (p->big ? "big" : "lite")
       << " mode=" << mode;
    for (int c = 0; c < NCLS; ++c)
        if (per_cls[c]) os << " " << kClsName[c] << "=" << per_cls[c];
    for (auto& kv : chosen) os << " " << kv.first << "(" << kv.second << ")";
    p->text = os.str();
    *out = p;
    return OMB_OK;
}

}  // namespace
}  // namespace omb

using namespace omb;

extern "C" omb_status omb_group_plan_create(const omb_gemm_desc* gemms, int32_t n_gemm,
                                            const omb_gemv_desc* gemvs, int32_t n_gemv,
                                            int32_t mode, omb_stream_t stream,
                                            omb_group_plan** plan) {
    clear_error();
    (void)stream;  // the graph is captured on a private stream and replayed on any
    if (!plan) {
        set_error("omb_group_plan_create: null plan pointer");
        return OMB_EINVAL;
    }
    return build_plan(gemms, n_gemm, gemvs, n_gemv, mode, -1, true, plan);
}

extern "C" omb_status omb_group_plan_launch(omb_group_plan* plan, omb_stream_t stream) {
    clear_error();
    if (!plan) {
        set_error("omb_group_plan_launch: null plan");
        return OMB_EINVAL;
    }
    if (plan->nitems == 0) return OMB_OK;
    OMB_HIP_TRY(hipGraphLaunch(plan->exec, (hipStream_t)stream));
    return OMB_OK;
}

extern "C" omb_status omb_group_plan_launch_direct(omb_group_plan* plan, omb_stream_t stream) {
    clear_error();
    if (!plan) {
        set_error("omb_group_plan_launch_direct: null plan");
        return OMB_EINVAL;
    }
    return launch_plan(plan, (hipStream_t)stream);
}

extern "C" omb_status omb_group_plan_describe(const omb_group_plan* plan, char* buf, size_t len) {
    clear_error();
    if (!plan || !buf || len == 0) {
        set_error("omb_group_plan_describe: invalid argument");
        return OMB_EINVAL;
    }
    std::snprintf(buf, len, "%s", plan->text.c_str());
    return OMB_OK;
}

extern "C" omb_status omb_group_plan_destroy(omb_group_plan* plan) {
    clear_error();
    free_plan(plan);
    return OMB_OK;
}

// ---------------------------------------------------------------- tuning --
namespace omb {
namespace {

__global__ void tune_fill(float* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        p[i] = rng_uniform(0x5eedULL, (uint64_t)i);
}

// Time `reps` direct launches of a plan that puts `batch` copies of the
// shape in class `cls`; returns milliseconds per launch.
omb_status time_class(int mode, int64_t M, int64_t N, int64_t K, int cls, int batch, float* A,
                      float* B, float* C, hipStream_t s, float* ms) {
    std::vector<omb_gemm_desc> d(batch);
    for (int i = 0; i < batch; ++i)
        d[i] = omb_gemm_desc{M, N, K, A, K, B, N, C + (int64_t)i * M * N, N, 1.f, 0.f};
    omb_group_plan* p = nullptr;
    OMB_TRY(build_plan(d.data(), batch, nullptr, 0, mode, cls, false, &p));
    hipEvent_t e0, e1;
    omb_status st = OMB_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        free_plan(p);
        set_error("autotune: event creation failed");
        return OMB_EDEVICE;
    }
    for (int w = 0; w < 2 && st == OMB_OK; ++w) st = launch_plan(p, s);
    const int reps = 5;
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps && st == OMB_OK; ++r) st = launch_plan(p, s);
    (void)hipEventRecord(e1, s);
    if (st == OMB_OK && hipEventSynchronize(e1) == hipSuccess) {
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e0, e1);
        *ms = t / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    free_plan(p);
    return st;
}

}  // namespace
}  // namespace omb

extern "C" omb_status omb_autotune_group(const omb_gemm_desc* gemms, int32_t n_gemm, int32_t mode,
                                         const char* table_path, omb_stream_t stream) {
    clear_error();
    OMB_TRY(validate(gemms, n_gemm, nullptr, 0));
    if (mode != OMB_MODE_PLAIN_F32 && mode != OMB_MODE_SPLIT_BF16 &&
        mode != OMB_MODE_SPLIT_BF16_P3 && mode != OMB_MODE_SPLIT_FP16) {
        set_error("omb_autotune_group: unknown mode");
        return OMB_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    // bucket -> shapes present (the median M*N*K shape represents the bucket)
    std::map<Key, std::vector<omb_gemm_desc>> buckets;
    for (int i = 0; i < n_gemm; ++i)
        if (gemms[i].M > 0 && gemms[i].N > 0) buckets[key_of(mode, gemms[i])].push_back(gemms[i]);
    const int batch = 64;
    // 128x64 is not a candidate: timed alone it can win a bucket, but it
    // needs the larger-register kernel (2 workgroups per CU instead of 3),
    // which slows every other item of the plan (see default_class); a
    // table loaded with omb_autotune_load may still name it.
    const std::vector<int> cands = mode == OMB_MODE_PLAIN_F32 ? std::vector<int>{C32, C64}
                                                              : std::vector<int>{CSPLIT};
    int64_t maxM = 1, maxN = 1, maxK = 1;
    for (auto& kv : buckets)
        for (auto& d : kv.second) {
            maxM = std::max(maxM, d.M);
            maxN = std::max(maxN, d.N);
            maxK = std::max(maxK, d.K);
        }
    float *A = nullptr, *B = nullptr, *C = nullptr;
    if (cands.size() > 1 && !buckets.empty()) {
        if (hipMalloc(&A, maxM * maxK * sizeof(float)) != hipSuccess ||
            hipMalloc(&B, maxK * maxN * sizeof(float)) != hipSuccess ||
            hipMalloc(&C, (size_t)batch * maxM * maxN * sizeof(float)) != hipSuccess) {
            (void)hipGetLastError();
            if (A) (void)hipFree(A);
            if (B) (void)hipFree(B);
            set_error("omb_autotune_group: scratch allocation failed");
            return OMB_ENOMEM;
        }
        hipLaunchKernelGGL(tune_fill, dim3(1024), dim3(256), 0, s, A, maxM * maxK);
        hipLaunchKernelGGL(tune_fill, dim3(1024), dim3(256), 0, s, B, maxK * maxN);
    }
    omb_status st = OMB_OK;
    std::map<Key, int> result;
    for (auto& kv : buckets) {
        if (cands.size() == 1) {
            result[kv.first] = cands[0];
            continue;
        }
        auto shapes = kv.second;
        std::sort(shapes.begin(), shapes.end(), [](const omb_gemm_desc& a, const omb_gemm_desc& b) {
            return a.M * a.N * a.K < b.M * b.N * b.K;
        });
        const omb_gemm_desc rep = shapes[shapes.size() / 2];
        int best = cands[0];
        float best_ms = 1e30f;
        for (int c : cands) {
            float ms = 1e30f;
            st = time_class(mode, rep.M, rep.N, rep.K, c, batch, A, B, C, s, &ms);
            if (st != OMB_OK) break;
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
        if (st != OMB_OK) break;
        result[kv.first] = best;
    }
    if (A) (void)hipFree(A);
    if (B) (void)hipFree(B);
    if (C) (void)hipFree(C);
    if (st != OMB_OK) return st;
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        for (auto& kv : result) tune_table()[kv.first] = kv.second;
    }
    if (table_path) {
        std::ofstream f(table_path);
        if (!f) {
            set_error(std::string("omb_autotune_group: cannot write ") + table_path);
            return OMB_EINVAL;
        }
        f << "# mode log2M log2N log2K class\n";
        std::lock_guard<std::mutex> lock(g_tune_mu);
        for (auto& kv : tune_table())
            f << kv.first.mode << " " << kv.first.bm << " " << kv.first.bn << " " << kv.first.bk
              << " " << kClsName[kv.second] << "\n";
    }
    return OMB_OK;
}

extern "C" omb_status omb_autotune_load(const char* table_path) {
    clear_error();
    if (!table_path) {
        set_error("omb_autotune_load: null path");
        return OMB_EINVAL;
    }
    std::ifstream f(table_path);
    if (!f) {
        set_error(std::string("omb_autotune_load: cannot read ") + table_path);
        return OMB_EINVAL;
    }
    std::map<Key, int> loaded;
    std::string err;
    if (!parse_tune_table(f, &loaded, &err)) {
        set_error("omb_autotune_load: " + err);
        return OMB_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_tune_mu);
    for (auto& kv : loaded) tune_table()[kv.first] = kv.second;
    return OMB_OK;
}

// ---------------------------------------------------------- request batches --
// A grouped launch without a plan object, for the serving loop: the
// descriptors and the cost-sorted item list are built on the host into a
// pinned staging block, copied to the batch's device block on the launch
// stream, and the lite persistent kernel runs them.  Results are bitwise
// those of the grouped plans (and so of the direct calls, see above).  The
// staging block is reused only after the previous launch from it has
// completed, which the caller guarantees (one batch per slot at a time).
namespace omb {

struct GroupBatch {
    char* dev = nullptr;
    char* host = nullptr;
    unsigned* counters = nullptr;
    size_t bytes = 0;
};

namespace {
constexpr int kBatchMaxGemms = 1024, kBatchMaxGemvs = 1024, kBatchMaxItems = 65536;

size_t batch_block_bytes() {
    return round_up(sizeof(DGemm) * kBatchMaxGemms, 256) +
           round_up(sizeof(DGemv) * kBatchMaxGemvs, 256) + sizeof(Item) * kBatchMaxItems;
}

int lite_per_cu() {
    static int per_cu = -1;
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    if (per_cu < 0) {
        int v = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, grouped_kernel<false>, NT, 0) !=
                hipSuccess ||
            v < 1) {
            (void)hipGetLastError();
            v = 1;
        }
        per_cu = v;
    }
    return per_cu;
}
}  // namespace

omb_status batch_create(GroupBatch** out) {
    *out = nullptr;
    auto* b = new GroupBatch();
    b->bytes = batch_block_bytes();
    if (hipMalloc(reinterpret_cast<void**>(&b->dev), b->bytes) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&b->host), b->bytes) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&b->counters), 2 * sizeof(unsigned)) != hipSuccess ||
        hipMemset(b->counters, 0, 2 * sizeof(unsigned)) != hipSuccess) {
        (void)hipGetLastError();
        batch_destroy(b);
        set_error("serve: batch buffer allocation failed");
        return OMB_ENOMEM;
    }
    *out = b;
    return OMB_OK;
}

void batch_destroy(GroupBatch* b) {
    if (!b) return;
    if (b->dev) (void)hipFree(b->dev);
    if (b->host) (void)hipHostFree(b->host);
    if (b->counters) (void)hipFree(b->counters);
    delete b;
}

bool batch_mode_supported(int mode) {
    return mode == OMB_MODE_PLAIN_F32 || mode == OMB_MODE_SPLIT_BF16 ||
           mode == OMB_MODE_SPLIT_BF16_P3;
}

omb_status batch_launch(GroupBatch* b, const omb_gemm_desc* gemms, int n_gemm,
                        const omb_gemv_desc* gemvs, int n_gemv, int mode, hipStream_t s) {
    if (!batch_mode_supported(mode)) {
        set_error("serve batch: mode " + std::to_string(mode) + " is not batched");
        return OMB_EUNSUPPORTED;
    }
    if (n_gemm < 0 || n_gemv < 0 || n_gemm > kBatchMaxGemms || n_gemv > kBatchMaxGemvs) {
        set_error("serve batch: too many instances");
        return OMB_EUNSUPPORTED;
    }
    OMB_TRY(validate(gemms, n_gemm, gemvs, n_gemv));
    DGemm* g = reinterpret_cast<DGemm*>(b->host);
    const size_t voff = round_up(sizeof(DGemm) * kBatchMaxGemms, 256);
    DGemv* v = reinterpret_cast<DGemv*>(b->host + voff);
    const size_t ioff = voff + round_up(sizeof(DGemv) * kBatchMaxGemvs, 256);
    Item* items = reinterpret_cast<Item*>(b->host + ioff);
    std::vector<Item> list;
    std::vector<DGemm> gv(n_gemm);
    std::vector<DGemv> vv(n_gemv);
    for (int i = 0; i < n_gemm; ++i) {
        const omb_gemm_desc& d = gemms[i];
        gv[i] = DGemm{d.M, d.N, d.K, d.A, d.lda, d.B, d.ldb, d.C, d.ldc, d.alpha, d.beta,
                      nullptr, nullptr};
        if (d.M == 0 || d.N == 0) continue;
        int c = choose_class(mode, d);
        if (c == C128x64) c = C64;  // same bits; keeps the lite kernel
        const int64_t tmn = ceil_div(d.M, kTileM[c]), tnn = ceil_div(d.N, kTileN[c]);
        if ((int64_t)list.size() + tmn * tnn > kBatchMaxItems) {
            set_error("serve batch: too many tiles");
            return OMB_EUNSUPPORTED;
        }
        for (int64_t a = 0; a < tmn; ++a)
            for (int64_t bb = 0; bb < tnn; ++bb) list.push_back(Item{c, i, (int)a, (int)bb});
    }
    for (int i = 0; i < n_gemv; ++i) {
        const omb_gemv_desc& d = gemvs[i];
        vv[i] = DGemv{d.M, d.N, d.A, d.lda, d.x, d.y, d.alpha, d.beta};
        const int64_t chunks = ceil_div(d.M, 64);
        if ((int64_t)list.size() + chunks > kBatchMaxItems) {
            set_error("serve batch: too many GEMV chunks");
            return OMB_EUNSUPPORTED;
        }
        for (int64_t c = 0; c < chunks; ++c) list.push_back(Item{CGEMV, i, (int)c, 0});
    }
    const int nitems = (int)list.size();
    if (nitems == 0) return OMB_OK;
    std::vector<double> cost(list.size());
    for (size_t i = 0; i < list.size(); ++i) cost[i] = item_cost(list[i], gv, vv);
    std::vector<size_t> order(list.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return cost[x] > cost[y]; });
    if (n_gemm) std::memcpy(g, gv.data(), sizeof(DGemm) * n_gemm);
    if (n_gemv) std::memcpy(v, vv.data(), sizeof(DGemv) * n_gemv);
    for (size_t i = 0; i < order.size(); ++i) items[i] = list[order[i]];
    // One copy of the used block (the unused tails of the descriptor arrays
    // travel too; the block is at most a few hundred KB).
    const size_t used = ioff + sizeof(Item) * (size_t)nitems;
    OMB_HIP_TRY(hipMemcpyAsync(b->dev, b->host, used, hipMemcpyHostToDevice, s));
    const int grid = std::max(1, std::min(nitems, device_cu_count() * lite_per_cu()));
    const int chunk = std::max(1, std::min(4, nitems / (8 * grid)));
    PArgs a{reinterpret_cast<const Item*>(b->dev + ioff), nitems,
            reinterpret_cast<const DGemm*>(b->dev), reinterpret_cast<const DGemv*>(b->dev + voff),
            b->counters, mode == OMB_MODE_SPLIT_BF16_P3 ? 3 : 6, 0, chunk,
            (int)ceil_div(nitems, chunk)};
    hipLaunchKernelGGL(grouped_kernel<false>, dim3(grid), dim3(NT), 0, s, a);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

// The steps of batch_launch over no items: the occupancy query (cached for
// the process), a copy of the whole pinned block (copies this size run as a
// runtime kernel, whose first launch is the cost), and the launch, in which one
// workgroup finds no chunk (nchunks = 0), reads no item and, as the last
// workgroup to finish, leaves both counters at zero.
omb_status batch_warm(GroupBatch* b, hipStream_t s) {
    (void)lite_per_cu();
    OMB_HIP_TRY(hipMemcpyAsync(b->dev, b->host, b->bytes, hipMemcpyHostToDevice, s));
    PArgs a{reinterpret_cast<const Item*>(b->dev), 0, reinterpret_cast<const DGemm*>(b->dev),
            reinterpret_cast<const DGemv*>(b->dev), b->counters, 6, 0, 1, 0};
    hipLaunchKernelGGL(grouped_kernel<false>, dim3(1), dim3(NT), 0, s, a);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

}  // namespace omb
