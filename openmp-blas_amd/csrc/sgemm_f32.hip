// Plain FP32 SGEMM for gfx950: C = alpha*A*B + beta*C on the vector ALUs.
//
// Counterpart of amt::mtm_helper / impl::mtm_kernel / simd_loop<MTM>
// (include/mtm.hpp:83-206, include/simd_loop.hpp:80-192): the reference packs
// MR x KC and KC x NR panels into cache-sized buffers and runs a 16x6 register
// micro-kernel.  Here a 256-thread workgroup owns a BM x BN tile of C (three
// classes: 128x128, 128x64, 64x64 -- a3's cache-level block choice becomes a
// choice of tile class and K split per shape), stages BM x 16 and 16 x BN
// slices of A and B through LDS (double buffered; the staging plays the role
// of `pack`, including the layout change for either operand order), and
// every thread keeps a (BM/16) x (BN/16) register micro-tile.  Global loads
// run two slices ahead in registers and LDS fragments one k ahead.
//
// Accumulation order: each thread sums 128 consecutive k (8 LDS slices) into
// an inner register tile, then adds it to an outer tile -- the same for every
// tile class, so the class never changes an output's bits.  This two-level sum
// keeps the rounding error of long K at or below the reference's (which sums
// KC consecutive k and then does C += block, simd_loop.hpp:120-128,160-190).
//
// Row-major C only: a column-major C is handled by the caller through
// C^T = B^T A^T (swap operands and their strides).
#include "common.hpp"

namespace omb {
namespace {

constexpr int BK = 16;
constexpr int NT = 256;         // threads per workgroup (a 16 x 16 thread grid)
constexpr int INNER_K = 128;    // k steps per inner accumulation block
constexpr int PAD = 4;          // LDS row padding (floats), keeps float4 alignment
constexpr int kGroupM = 8;      // tile rows per L2 group
constexpr int64_t kSplitMinK = 256;    // least k per coarse split (K > kSplitExactK)
constexpr int64_t kSplitExactK = 512;  // up to this K, splits are only single accumulation blocks
constexpr int64_t kSplitMost = 16;     // most coarse splits considered
constexpr double kSplitMaxBytes = 128.0 * (1 << 20);  // most partial-sum bytes
// FP32 flop/s of one 128x128 workgroup, measured: with two per CU (86
// TFLOP/s over 512 at 8192^3) and alone on its CU (50 TFLOP/s over 256 at
// 2048^3: one wave per SIMD loses the dual issue of two).
constexpr double kTileRateF32 = 0.17e12;
constexpr double kTileRateF32Alone = 0.2e12;
// The same for the 128x64 class (three per CU) and the 64x64 class (four
// per CU), scaled from one measurement of all three classes (8192 x 8192 x
// 2048 for the full rate, 16 x 16 tiles with K = 4096 alone): per workgroup
// 0.154 / 0.103 / 0.072 TFLOP/s full and 0.196 / 0.155 / 0.126 alone.
constexpr double kTileRateF32_128x64 = 0.114e12;
constexpr double kTileRateF32_128x64Alone = 0.158e12;
constexpr double kTileRateF32_64x64 = 0.079e12;
constexpr double kTileRateF32_64x64Alone = 0.129e12;
constexpr double kSplitHbm = 4.0e12;      // bytes/s for the partials' traffic
constexpr double kSplitLaunch = 4.0e-6;   // the reduction kernel's launch and tail
constexpr double kSplitReduceStep = 1.0e-6;  // one round trip per 8 partials of an element

struct Args {
    int64_t M, N, K;
    float alpha, beta;
    const float* A;
    int64_t lda;
    const float* B;
    int64_t ldb;
    float* C;
    int64_t ldc;
    int64_t tiles_m, tiles_n;
    int64_t kchunk;  // k per split (blockIdx.y), a multiple of INNER_K
    float* part;     // split partial sums [split][M][N], or null (one split)
};

// Load up to 4 consecutive floats starting at p (n valid, 0..4); zero-fill.
template <bool VEC>
__device__ inline float4 load4(const float* p, int n) {
    if (VEC && n == 4) return *reinterpret_cast<const float4*>(p);
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (n > 0) v.x = p[0];
    if (n > 1) v.y = p[1];
    if (n > 2) v.z = p[2];
    if (n > 3) v.w = p[3];
    return v;
}

__device__ inline int clamp4(int64_t r) { return r <= 0 ? 0 : (r >= 4 ? 4 : (int)r); }

// Global -> registers for one BK slice of a BM x BN tile: LA = BM/64 float4
// of A and LB = BN/64 float4 of B per thread, each along the operand's
// unit-stride dimension.
// FULL: the slice lies wholly inside A and B and the operands are 16-byte
// aligned, so every load is an unguarded float4.
template <int BM, int BN, bool A_KC, bool B_NC, bool VEC, bool FULL = false>
__device__ inline void load_slice(const Args& p, int64_t m0, int64_t n0, int64_t k0, int tid,
                                  float4 (&ra)[BM / 64], float4 (&rb)[BN / 64]) {
#pragma unroll
    for (int i = 0; i < BM / 64; ++i) {
        const int e = tid + i * NT;
        if (A_KC) {  // A[m*lda + k]: float4 along k
            int row = e >> 2, kq = (e & 3) * 4;
            int64_t m = m0 + row, k = k0 + kq;
            if constexpr (FULL) {
                ra[i] = *reinterpret_cast<const float4*>(p.A + m * p.lda + k);
            } else {
                int n = m < p.M ? clamp4(p.K - k) : 0;
                ra[i] = load4<VEC>(p.A + (n ? m * p.lda + k : 0), n);
            }
        } else {     // A[k*lda + m]: float4 along m
            int mq = (e % (BM / 4)) * 4, kk = e / (BM / 4);
            int64_t m = m0 + mq, k = k0 + kk;
            if constexpr (FULL) {
                ra[i] = *reinterpret_cast<const float4*>(p.A + k * p.lda + m);
            } else {
                int n = k < p.K ? clamp4(p.M - m) : 0;
                ra[i] = load4<VEC>(p.A + (n ? k * p.lda + m : 0), n);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < BN / 64; ++i) {
        const int e = tid + i * NT;
        if (B_NC) {  // B[k*ldb + n]: float4 along n
            int nq = (e % (BN / 4)) * 4, kk = e / (BN / 4);
            int64_t nn = n0 + nq, k = k0 + kk;
            if constexpr (FULL) {
                rb[i] = *reinterpret_cast<const float4*>(p.B + k * p.ldb + nn);
            } else {
                int n = k < p.K ? clamp4(p.N - nn) : 0;
                rb[i] = load4<VEC>(p.B + (n ? k * p.ldb + nn : 0), n);
            }
        } else {     // B[n*ldb + k]: float4 along k
            int col = e >> 2, kq = (e & 3) * 4;
            int64_t nn = n0 + col, k = k0 + kq;
            if constexpr (FULL) {
                rb[i] = *reinterpret_cast<const float4*>(p.B + nn * p.ldb + k);
            } else {
                int n = nn < p.N ? clamp4(p.K - k) : 0;
                rb[i] = load4<VEC>(p.B + (n ? nn * p.ldb + k : 0), n);
            }
        }
    }
}

// Registers -> LDS, laid out [k][m] and [k][n].
template <int BM, int BN, bool A_KC, bool B_NC>
__device__ inline void store_slice(float (*As)[BM + PAD], float (*Bs)[BN + PAD], int tid,
                                   const float4 (&ra)[BM / 64], const float4 (&rb)[BN / 64]) {
#pragma unroll
    for (int i = 0; i < BM / 64; ++i) {
        const int e = tid + i * NT;
        if (A_KC) {
            int row = e >> 2, kq = (e & 3) * 4;
            As[kq + 0][row] = ra[i].x;
            As[kq + 1][row] = ra[i].y;
            As[kq + 2][row] = ra[i].z;
            As[kq + 3][row] = ra[i].w;
        } else {
            int mq = (e % (BM / 4)) * 4, kk = e / (BM / 4);
            *reinterpret_cast<float4*>(&As[kk][mq]) = ra[i];
        }
    }
#pragma unroll
    for (int i = 0; i < BN / 64; ++i) {
        const int e = tid + i * NT;
        if (B_NC) {
            int nq = (e % (BN / 4)) * 4, kk = e / (BN / 4);
            *reinterpret_cast<float4*>(&Bs[kk][nq]) = rb[i];
        } else {
            int col = e >> 2, kq = (e & 3) * 4;
            Bs[kq + 0][col] = rb[i].x;
            Bs[kq + 1][col] = rb[i].y;
            Bs[kq + 2][col] = rb[i].z;
            Bs[kq + 3][col] = rb[i].w;
        }
    }
}

// Workgroups per CU each class is built for: the 128x128 class holds 2 x 64
// accumulators per thread (2 workgroups per CU), 128x64 2 x 32 (3), 64x64
// 2 x 16 (4).
constexpr int occ_of(int tile_elems) {
    return tile_elems >= 128 * 128 ? 2 : (tile_elems >= 128 * 64 ? 3 : 4);
}

// Thread (ty, tx) of the 16 x 16 grid owns rows {g*64 + ty*4 + i} (g < GR =
// BM/64, i < 4) and columns {h*64 + tx*4 + j} (h < GC = BN/64, j < 4).
template <int BM, int BN, bool A_KC, bool B_NC, bool VEC>
__global__ __launch_bounds__(NT, occ_of(BM * BN)) void sgemm_f32_kernel(Args q) {
    constexpr int GR = BM / 64, GC = BN / 64, TR = 4 * GR, TC = 4 * GC;
    __shared__ float As[2][BK][BM + PAD];
    __shared__ float Bs[2][BK][BN + PAD];

    // Split blockIdx.y takes k in [y * kchunk, min(K, (y + 1) * kchunk)): the
    // operands are offset to that range and K becomes its length.  kchunk is
    // a multiple of INNER_K, so the 128-k accumulation blocks are the same
    // as without the split.
    Args p = q;
    const int64_t kbeg = (int64_t)blockIdx.y * q.kchunk;
    if (q.part) {
        p.K = q.K - kbeg < q.kchunk ? q.K - kbeg : q.kchunk;
        p.A = q.A + (A_KC ? kbeg : kbeg * q.lda);
        p.B = q.B + (B_NC ? kbeg * q.ldb : kbeg);
    }

    const int tid = threadIdx.x;
    const int64_t ntiles = p.tiles_m * p.tiles_n;
    int64_t t = xcd_remap(blockIdx.x, ntiles);
    int64_t tm, tn;
    group_tile(t, p.tiles_m, p.tiles_n, kGroupM, &tm, &tn);
    const int64_t m0 = tm * BM, n0 = tn * BN;

    const int tx = tid & 15, ty = tid >> 4;
    float inner[TR][TC], outer[TR][TC];
#pragma unroll
    for (int i = 0; i < TR; ++i)
#pragma unroll
        for (int j = 0; j < TC; ++j) inner[i][j] = outer[i][j] = 0.f;

    const int64_t nk = ceil_div(p.K, BK);
    constexpr int kFold = INNER_K / BK;
    // Slices kt and kt+1 live in LDS buffers 0 and 1; global loads run two
    // slices ahead in two register sets.
    auto compute = [&](int cur, int64_t kt) {
        float4 fa[2][GR], fb[2][GC];
        auto fragment = [&](int k, int set) {
#pragma unroll
            for (int g = 0; g < GR; ++g)
                fa[set][g] = *reinterpret_cast<const float4*>(&As[cur][k][g * 64 + ty * 4]);
#pragma unroll
            for (int h = 0; h < GC; ++h)
                fb[set][h] = *reinterpret_cast<const float4*>(&Bs[cur][k][h * 64 + tx * 4]);
        };
        fragment(0, 0);
#pragma unroll
        for (int k = 0; k < BK; ++k) {
            if (k + 1 < BK) fragment(k + 1, (k + 1) & 1);
            float a[TR], b[TC];
#pragma unroll
            for (int g = 0; g < GR; ++g) {
                const float4 v = fa[k & 1][g];
                a[4 * g] = v.x; a[4 * g + 1] = v.y; a[4 * g + 2] = v.z; a[4 * g + 3] = v.w;
            }
#pragma unroll
            for (int h = 0; h < GC; ++h) {
                const float4 v = fb[k & 1][h];
                b[4 * h] = v.x; b[4 * h + 1] = v.y; b[4 * h + 2] = v.z; b[4 * h + 3] = v.w;
            }
#pragma unroll
            for (int i = 0; i < TR; ++i)
#pragma unroll
                for (int j = 0; j < TC; ++j) inner[i][j] = fmaf(a[i], b[j], inner[i][j]);
        }
        if ((kt + 1) % kFold == 0 || kt + 1 == nk) {
#pragma unroll
            for (int i = 0; i < TR; ++i)
#pragma unroll
                for (int j = 0; j < TC; ++j) {
                    outer[i][j] += inner[i][j];
                    inner[i][j] = 0.f;
                }
        }
    };
    float4 ra0[GR], rb0[GC], ra1[GR], rb1[GC];
    const bool full = VEC && m0 + BM <= p.M && n0 + BN <= p.N && p.K > 0 && p.K % BK == 0;
    if (VEC && full) {
        // Interior tile: unguarded loads with the slice index clamped to the
        // last one, whole slice pairs, an odd last slice after the loop.  One
        // path through the loop keeps the compiler's wait counts exact (a
        // store waits only for its own loads), so the loads stay two slices
        // ahead.  Same slices and fold order as the guarded loop below.
        const int64_t klast = nk - 1;
        auto kc = [&](int64_t k) { return k < klast ? k : klast; };
        load_slice<BM, BN, A_KC, B_NC, VEC, true>(p, m0, n0, 0, tid, ra0, rb0);
        store_slice<BM, BN, A_KC, B_NC>(As[0], Bs[0], tid, ra0, rb0);
        load_slice<BM, BN, A_KC, B_NC, VEC, true>(p, m0, n0, kc(1) * BK, tid, ra1, rb1);
        __syncthreads();
        int64_t kt = 0;
        for (; kt + 1 < nk; kt += 2) {
            load_slice<BM, BN, A_KC, B_NC, VEC, true>(p, m0, n0, kc(kt + 2) * BK, tid, ra0, rb0);
            __builtin_amdgcn_sched_barrier(0);
            compute(0, kt);
            store_slice<BM, BN, A_KC, B_NC>(As[1], Bs[1], tid, ra1, rb1);
            __syncthreads();
            load_slice<BM, BN, A_KC, B_NC, VEC, true>(p, m0, n0, kc(kt + 3) * BK, tid, ra1, rb1);
            __builtin_amdgcn_sched_barrier(0);
            compute(1, kt + 1);
            store_slice<BM, BN, A_KC, B_NC>(As[0], Bs[0], tid, ra0, rb0);
            __syncthreads();
        }
        if (kt < nk) compute(0, kt);
    } else {
        load_slice<BM, BN, A_KC, B_NC, VEC>(p, m0, n0, 0, tid, ra0, rb0);
        store_slice<BM, BN, A_KC, B_NC>(As[0], Bs[0], tid, ra0, rb0);
        if (nk > 1) load_slice<BM, BN, A_KC, B_NC, VEC>(p, m0, n0, BK, tid, ra1, rb1);
        __syncthreads();
        for (int64_t kt = 0; kt < nk; kt += 2) {
            if (kt + 2 < nk)
                load_slice<BM, BN, A_KC, B_NC, VEC>(p, m0, n0, (kt + 2) * BK, tid, ra0, rb0);
            compute(0, kt);
            if (kt + 1 < nk) store_slice<BM, BN, A_KC, B_NC>(As[1], Bs[1], tid, ra1, rb1);
            __syncthreads();
            if (kt + 1 >= nk) break;
            if (kt + 3 < nk)
                load_slice<BM, BN, A_KC, B_NC, VEC>(p, m0, n0, (kt + 3) * BK, tid, ra1, rb1);
            compute(1, kt + 1);
            if (kt + 2 < nk) store_slice<BM, BN, A_KC, B_NC>(As[0], Bs[0], tid, ra0, rb0);
            __syncthreads();
        }
    }

    // Epilogue: rows {g*64 + ty*4 + i}, columns {h*64 + tx*4 + j}.
    if (p.part) {  // raw partial sums of this split, row-major M x N
        float* part = p.part + (int64_t)blockIdx.y * p.M * p.N;
        const bool pvec = p.N % 4 == 0;
#pragma unroll
        for (int i = 0; i < TR; ++i) {
            const int64_t m = m0 + (i / 4) * 64 + ty * 4 + (i % 4);
            if (m >= p.M) continue;
#pragma unroll
            for (int h = 0; h < GC; ++h) {
                const int64_t n = n0 + h * 64 + tx * 4;
                float* c = part + m * p.N + n;
                if (pvec && n + 3 < p.N) {
                    *reinterpret_cast<float4*>(c) = make_float4(outer[i][h * 4], outer[i][h * 4 + 1],
                                                                outer[i][h * 4 + 2], outer[i][h * 4 + 3]);
                } else {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (n + j < p.N) c[j] = outer[i][h * 4 + j];
                }
            }
        }
        return;
    }
    const bool cvec = ((reinterpret_cast<uintptr_t>(p.C) & 15) == 0) && (p.ldc % 4 == 0);
#pragma unroll
    for (int i = 0; i < TR; ++i) {
        const int64_t m = m0 + (i / 4) * 64 + ty * 4 + (i % 4);
        if (m >= p.M) continue;
#pragma unroll
        for (int h = 0; h < GC; ++h) {
            const int64_t n = n0 + h * 64 + tx * 4;
            float* c = p.C + m * p.ldc + n;
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = p.alpha * outer[i][h * 4 + j];
            if (cvec && n + 3 < p.N) {
                float4 o;
                if (p.beta != 0.f) {
                    float4 c4 = *reinterpret_cast<const float4*>(c);
                    o = make_float4(v[0] + p.beta * c4.x, v[1] + p.beta * c4.y,
                                    v[2] + p.beta * c4.z, v[3] + p.beta * c4.w);
                } else {
                    o = make_float4(v[0], v[1], v[2], v[3]);
                }
                *reinterpret_cast<float4*>(c) = o;
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (n + j < p.N) c[j] = p.beta != 0.f ? v[j] + p.beta * c[j] : v[j];
                }
            }
        }
    }
}

// C = alpha * (sum of the splits' partials, in split order) + beta * C.
// Rows over blockIdx.y, columns over threads: VEC takes four adjacent
// columns per thread with 128-bit loads and stores (N, ldc multiples of 4,
// C 16-byte aligned), else one column.  Eight partials are in flight per
// step (a long split list is otherwise one memory round trip per partial);
// every element's adds stay in split order, so VEC changes no bits.
template <bool VEC>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(int64_t M, int64_t N, int64_t splits,
                                                            float alpha, float beta,
                                                            const float* __restrict__ part,
                                                            float* C, int64_t ldc) {
    constexpr int W = VEC ? 4 : 1;
    typedef float vf __attribute__((ext_vector_type(W)));
    const int64_t total = M * N, cols = N / W;
    for (int64_t m = blockIdx.y; m < M; m += gridDim.y) {
        for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < cols;
             q += (int64_t)gridDim.x * blockDim.x) {
            const int64_t i = m * N + q * W;
            vf s = *reinterpret_cast<const vf*>(part + i);
            int64_t t = 1;
            for (; t + 8 <= splits; t += 8) {
                vf v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = *reinterpret_cast<const vf*>(part + (t + u) * total + i);
#pragma unroll
                for (int u = 0; u < 8; ++u) s += v[u];
            }
            for (; t < splits; ++t) s += *reinterpret_cast<const vf*>(part + t * total + i);
            vf* c = reinterpret_cast<vf*>(C + m * ldc + q * W);
            const vf v = alpha * s;
            *c = beta != 0.f ? v + beta * *c : v;
        }
    }
}

template <int BM, int BN, bool A_KC, bool B_NC>
void launch_variant(const Args& a, int64_t splits, bool vec, hipStream_t s) {
    dim3 grid((unsigned)(a.tiles_m * a.tiles_n), (unsigned)splits), block(NT);
    if (vec)
        hipLaunchKernelGGL((sgemm_f32_kernel<BM, BN, A_KC, B_NC, true>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((sgemm_f32_kernel<BM, BN, A_KC, B_NC, false>), grid, block, 0, s, a);
}

template <int BM, int BN>
void launch_class(const Args& a, bool a_kc, bool b_nc, int64_t splits, bool vec, hipStream_t s) {
    if (a_kc && b_nc) launch_variant<BM, BN, true, true>(a, splits, vec, s);
    else if (a_kc && !b_nc) launch_variant<BM, BN, true, false>(a, splits, vec, s);
    else if (!a_kc && b_nc) launch_variant<BM, BN, false, true>(a, splits, vec, s);
    else launch_variant<BM, BN, false, false>(a, splits, vec, s);
}

// The plain tile classes: tile shape, workgroups per CU, and the FP32 flop/s
// of one workgroup with the CU full of its class and alone on its CU.
struct F32Class {
    int bm, bn, occ;
    double rate_full, rate_alone;
};
constexpr F32Class kF32Classes[3] = {
    {128, 128, 2, kTileRateF32, kTileRateF32Alone},
    {128, 64, 3, kTileRateF32_128x64, kTileRateF32_128x64Alone},
    {64, 64, 4, kTileRateF32_64x64, kTileRateF32_64x64Alone},
};

__global__ void scale_kernel(int64_t M, int64_t N, float beta, float* C, int64_t ldc) {
    int64_t total = M * N;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t m = i / N, n = i % N;
        float* c = C + m * ldc + n;
        *c = beta != 0.f ? beta * *c : 0.f;
    }
}

}  // namespace

int64_t gemm_k_splits_tiles(int64_t M, int64_t N, int64_t K, int64_t align, int bm, int bn,
                            int occ, double rate_full, double rate_alone, int64_t* kchunk,
                            double* time) {
    *kchunk = K;
    const int64_t tiles = ceil_div(M, bm) * ceil_div(N, bn);
    const int64_t cus = device_cu_count();
    // Rounds of workgroups (`occ` resident per CU) times the k each covers,
    // at the rate of a workgroup sharing its CU or, when there are no more
    // workgroups than CUs, alone on it; plus the partials' write and read,
    // the reduction launch and its round trips.
    auto cost = [&](int64_t splits, int64_t kc) {
        const int64_t w = tiles * splits;
        const double work = 2.0 * bm * bn * (double)kc;
        double t = w <= cus ? work / rate_alone : (double)ceil_div(w, occ * cus) * work / rate_full;
        if (splits > 1)
            t += (double)(splits + 1) * (double)(M * N) * 4.0 / kSplitHbm + kSplitLaunch +
                 (double)ceil_div(splits - 1, 8) * kSplitReduceStep;
        return t;
    };
    if (tiles <= 0 || K <= align) {
        if (time) *time = tiles > 0 ? cost(1, K) : 0.0;
        return 1;
    }
    const double part_bytes = (double)(M * N) * 4.0;
    // Exact choices first: no split, or one accumulation block per split.
    // The in-order sum of single blocks is the unsplit kernel's own fold
    // (outer = ((b0 + b1) + b2) ...), so either gives the unsplit bits, on
    // any device.  The split is taken when the model finds it 5% faster.
    int64_t best = 1;
    double best_t = cost(1, K);
    const int64_t nblk = ceil_div(K, align);
    if ((double)nblk * part_bytes <= kSplitMaxBytes) {
        const double t = cost(nblk, align);
        if (t < 0.95 * best_t) {
            best = nblk;
            best_t = t;
            *kchunk = align;
        }
    }
    if (K <= kSplitExactK) {
        if (time) *time = best_t;
        return best;
    }
    // Coarser splits (at least 256 k each) round differently: one is taken
    // only when it is more than 10% faster than the exact choice.
    const double exact_t = best_t;
    const int64_t most = std::min<int64_t>(kSplitMost, K / kSplitMinK);
    for (int64_t sp = 2; sp <= most; ++sp) {
        const int64_t kc = round_up(ceil_div(K, sp), align);
        const int64_t n = ceil_div(K, kc);
        if (kc == align) continue;  // the exact choice above
        if ((double)n * part_bytes > kSplitMaxBytes) break;  // partials too large
        const double t = cost(n, kc);
        if (n > 1 && t < 0.9 * exact_t && t < best_t) {
            best = n;
            best_t = t;
            *kchunk = kc;
        }
    }
    if (best == 1) *kchunk = K;
    if (time) *time = best_t;
    return best;
}

int64_t gemm_k_splits(int64_t M, int64_t N, int64_t K, int64_t align, double rate_full,
                      double rate_alone, int64_t* kchunk) {
    return gemm_k_splits_tiles(M, N, K, align, 128, 128, 2, rate_full, rate_alone, kchunk,
                               nullptr);
}

// Tile class and K split of a plain GEMM: the class whose modelled time is
// least, a smaller class only when the model finds it 15% faster than the
// larger one before it (with a 5% margin the model's optimism for partly
// filled rounds of small tiles cost 3-8% at some sizes from 1216 to 2432).
// Any class and any exact split give the same bits.
int sgemm_f32_choose(int64_t M, int64_t N, int64_t K, int64_t* splits, int64_t* kchunk) {
    int best = 0;
    double best_t = 0.0;
    for (int c = 0; c < 3; ++c) {
        const F32Class& k = kF32Classes[c];
        int64_t kc = K;
        double t = 0.0;
        const int64_t sp = gemm_k_splits_tiles(M, N, K, INNER_K, k.bm, k.bn, k.occ, k.rate_full,
                                               k.rate_alone, &kc, &t);
        if (c == 0 || t < 0.85 * best_t) {
            best = c;
            best_t = t;
            *splits = sp;
            *kchunk = kc;
        }
    }
    return best;
}

int64_t sgemm_f32_splits(int64_t M, int64_t N, int64_t K, int64_t* kchunk) {
    int64_t splits = 1;
    sgemm_f32_choose(M, N, K, &splits, kchunk);
    return splits;
}

omb_status launch_splitk_reduce(int64_t M, int64_t N, int64_t splits, float alpha, float beta,
                                const float* part, float* C, int64_t ldc, hipStream_t s) {
    if (M <= 0 || N <= 0) return OMB_OK;
    const bool vec = N % 4 == 0 && ldc % 4 == 0 && (reinterpret_cast<uintptr_t>(C) & 15) == 0 &&
                     (reinterpret_cast<uintptr_t>(part) & 15) == 0;
    const int64_t cols = vec ? N / 4 : N;
    const int64_t bx = std::min<int64_t>(ceil_div(cols, 256), 4 * device_cu_count());
    const int64_t by = std::min<int64_t>(
        std::min<int64_t>(M, 65535), std::max<int64_t>(1, 4 * device_cu_count() / bx));
    const dim3 grid((unsigned)bx, (unsigned)by);
    if (vec)
        hipLaunchKernelGGL(splitk_reduce_kernel<true>, grid, dim3(256), 0, s, M, N, splits, alpha,
                           beta, part, C, ldc);
    else
        hipLaunchKernelGGL(splitk_reduce_kernel<false>, grid, dim3(256), 0, s, M, N, splits,
                           alpha, beta, part, C, ldc);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

omb_status launch_sgemm_f32(int64_t M, int64_t N, int64_t K, float alpha,
                            const float* A, int64_t lda, bool a_kc,
                            const float* B, int64_t ldb, bool b_nc,
                            float beta, float* C, int64_t ldc, hipStream_t s, float* part) {
    // Tile class per shape; split K over blockIdx.y when the tiles alone
    // leave CUs idle (the splits' partials are then added in split order).
    int64_t splits = 1, kchunk = K;
    const int cls = sgemm_f32_choose(M, N, K, &splits, &kchunk);
    const F32Class& fc = kF32Classes[cls];
    Args a{M, N, K, alpha, beta, A, lda, B, ldb, C, ldc, ceil_div(M, fc.bm), ceil_div(N, fc.bn),
           kchunk, nullptr};
    const int64_t tiles = a.tiles_m * a.tiles_n;
    if (tiles > 0x7fffffffLL) {
        set_error("sgemm: too many tiles for one launch");
        return OMB_EUNSUPPORTED;
    }
    if (splits > 1) {
        if (!part) {
            Workspace* ws = nullptr;
            OMB_TRY(get_workspace(s, (size_t)(splits * M * N) * sizeof(float), &ws));
            part = reinterpret_cast<float*>(ws->buf);
        }
        a.part = part;
    }
    const bool vec = ((reinterpret_cast<uintptr_t>(A) & 15) == 0) &&
                     ((reinterpret_cast<uintptr_t>(B) & 15) == 0) && (lda % 4 == 0) &&
                     (ldb % 4 == 0);
    if (cls == 0) launch_class<128, 128>(a, a_kc, b_nc, splits, vec, s);
    else if (cls == 1) launch_class<128, 64>(a, a_kc, b_nc, splits, vec, s);
    else launch_class<64, 64>(a, a_kc, b_nc, splits, vec, s);
    OMB_HIP_TRY(hipGetLastError());
    if (splits > 1) return launch_splitk_reduce(M, N, splits, alpha, beta, a.part, C, ldc, s);
    return OMB_OK;
}

omb_status launch_scale(int64_t M, int64_t N, float beta, float* C, int64_t ldc, hipStream_t s) {
    if (M <= 0 || N <= 0 || beta == 1.f) return OMB_OK;
    int64_t total = M * N;
    unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(total, 256), 4096);
    hipLaunchKernelGGL(scale_kernel, dim3(blocks), dim3(256), 0, s, M, N, beta, C, ldc);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

}  // namespace omb
