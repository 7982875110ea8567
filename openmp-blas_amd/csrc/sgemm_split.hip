// Split-BF16 SGEMM for gfx950: an FP32 GEMM computed on the BF16 MFMA cores.
//
// Every FP32 operand element x is split exactly into three BF16 terms
//   x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1)
// (each subtraction is exact in FP32; 3 x 8 significand bits cover FP32's 24),
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
// (double buffered, bank-swizzled).  All products of 8 slices (128 k) are
// accumulated by the MFMAs into an inner register tile, smallest first, which
// is then added into an outer tile with FP32 adds: the running sum over K
// never absorbs MFMA rounding at its own magnitude.
#include <algorithm>

#include "common.hpp"

namespace omb {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int TM = 128, TN = 128, TK = 16;
constexpr int NT = 256;
constexpr int kGroupM = 8;
constexpr int SLICE_A = 3 * TM * TK;  // bf16 elements per A slice (3 planes)
constexpr int SLICE_B = 3 * TN * TK;

// ----------------------------------------------------------------- pass 1 --
// One thread: 8 consecutive k of one row -> one 16-byte store per plane.
__global__ __launch_bounds__(NT) void split_pack_kernel(const float* __restrict__ X, int64_t R,
                                                        int64_t K, int64_t rs, int64_t cs,
                                                        int64_t Rp, int64_t Kp, int br,
                                                        int k_major, __bf16* __restrict__ out) {
    const int64_t groups = Kp >> 3;
    const int64_t total = Rp * groups;
    const int64_t nkt = Kp / TK;
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * NT) {
        int64_t r, g;
        if (k_major) {
            r = i / groups;
            g = i - r * groups;
        } else {
            g = i / Rp;
            r = i - g * Rp;
        }
        const int64_t k0 = g << 3;
        float v[8];
        if (r < R && cs == 1 && k0 + 8 <= K &&
            ((reinterpret_cast<uintptr_t>(X + r * rs + k0) & 15) == 0)) {
            const float4* p = reinterpret_cast<const float4*>(X + r * rs + k0);
            float4 a = p[0], b = p[1];
            v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
            v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                v[j] = (r < R && k0 + j < K) ? X[r * rs + (k0 + j) * cs] : 0.f;
        }
        bf16x8 h0, h1, h2;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            __bf16 t0 = (__bf16)v[j];
            float r1 = v[j] - (float)t0;
            __bf16 t1 = (__bf16)r1;
            float r2 = r1 - (float)t1;
            h0[j] = t0;
            h1[j] = t1;
            h2[j] = (__bf16)r2;
        }
        const int64_t rt = r / br, rr = r - rt * br;
        const int64_t kt = k0 / TK, kk = k0 - kt * TK;
        __bf16* base = out + ((rt * nkt + kt) * 3) * (int64_t)(br * TK) + rr * TK + kk;
        *reinterpret_cast<bf16x8*>(base) = h0;
        *reinterpret_cast<bf16x8*>(base + br * TK) = h1;
        *reinterpret_cast<bf16x8*>(base + 2 * br * TK) = h2;
    }
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
};

__device__ inline f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// LDS position of 16-byte chunk h (0/1) of slice row r: the chunk index is
// flipped by bits 2 and 3 of the row so that the 8 or 16 lanes a ds_read_b128
// serves per cycle (rows r..r+15, same h) hit distinct banks.
__device__ inline int lds_chunk(int r, int h) { return (r << 1) | (h ^ ((r >> 2) & 1) ^ ((r >> 3) & 1)); }

constexpr int kFoldSlices = 8;  // inner accumulation spans 8 slices = 128 k
constexpr int VA = SLICE_A / 8 / NT;  // 16-byte loads per thread per slice (3)
constexpr int VB = SLICE_B / 8 / NT;
constexpr int V = VA + VB;

// One slice in flight is 6 uint4 per thread (3 of A, 3 of B), kept in named
// registers (arrays of them were demoted to scratch by the compiler).
#define OMB_GLOAD(kt, r0, r1, r2, r3, r4, r5)                            \
    do {                                                                  \
        const uint4* a_ = ga + (kt) * (SLICE_A / 8) + tid;                \
        const uint4* b_ = gb + (kt) * (SLICE_B / 8) + tid;                \
        r0 = a_[0];                                                       \
        r1 = a_[NT];                                                      \
        r2 = a_[2 * NT];                                                  \
        r3 = b_[0];                                                       \
        r4 = b_[NT];                                                      \
        r5 = b_[2 * NT];                                                  \
    } while (0)
#define OMB_LSTORE(l, r0, r1, r2, r3, r4, r5)                             \
    do {                                                                  \
        (l)[q0] = r0;                                                     \
        (l)[q1] = r1;                                                     \
        (l)[q2] = r2;                                                     \
        (l)[q3] = r3;                                                     \
        (l)[q4] = r4;                                                     \
        (l)[q5] = r5;                                                     \
    } while (0)

// Slice chunk q of A (0 .. 3*TM*2-1): plane q / (2*TM), row (q / 2) % TM,
// half q % 2 (same for B after the A part) -> its swizzled LDS position.
__device__ __forceinline__ int lds_index(int q) {
    const int pl = q / (2 * TM), rem = q - pl * 2 * TM;
    return pl * 2 * TM + lds_chunk(rem >> 1, rem & 1);
}

template <int NPROD>
__global__ __launch_bounds__(NT, 2) void sgemm_bf16x3_kernel(MArgs p) {
    __shared__ __attribute__((aligned(16))) uint4 lds[2][(SLICE_A + SLICE_B) / 8];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t ntiles = p.tiles_m * p.tiles_n;
    int64_t t = xcd_remap(blockIdx.x, ntiles);
    int64_t tm, tn;
    group_tile(t, p.tiles_m, p.tiles_n, kGroupM, &tm, &tn);

    const uint4* ga = reinterpret_cast<const uint4*>(p.Ap + tm * p.nkt * SLICE_A);
    const uint4* gb = reinterpret_cast<const uint4*>(p.Bp + tn * p.nkt * SLICE_B);
    static_assert(VA == 3 && VB == 3, "staging macros assume 3 + 3 loads per slice");
    // LDS slots of this thread's chunks (loop invariant).
    const int q0 = lds_index(tid), q1 = lds_index(tid + NT), q2 = lds_index(tid + 2 * NT);
    const int q3 = SLICE_A / 8 + q0, q4 = SLICE_A / 8 + q1, q5 = SLICE_A / 8 + q2;

    f32x16 inner[2][2], outer[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) inner[i][j][r] = outer[i][j][r] = 0.f;

    constexpr int NPLANES = NPROD == 6 ? 3 : 2;
    const int frow = lane & 31, fh = lane >> 5;
    auto compute = [&](int buf, int64_t kt) {
        const uint4* la = lds[buf];
        const uint4* lb = lds[buf] + SLICE_A / 8;
        bf16x8 a[NPLANES][2], b[NPLANES][2];
#pragma unroll
        for (int pl = 0; pl < NPLANES; ++pl)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                uint4 va = la[pl * 2 * TM + lds_chunk(wm * 64 + i * 32 + frow, fh)];
                uint4 vb = lb[pl * 2 * TN + lds_chunk(wn * 64 + i * 32 + frow, fh)];
                a[pl][i] = *reinterpret_cast<const bf16x8*>(&va);
                b[pl][i] = *reinterpret_cast<const bf16x8*>(&vb);
            }
        // smallest products first; everything goes into the inner sum
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (NPROD == 6) {
                    inner[i][j] = mfma(a[0][i], b[2][j], inner[i][j]);
                    inner[i][j] = mfma(a[1][i], b[1][j], inner[i][j]);
                    inner[i][j] = mfma(a[2][i], b[0][j], inner[i][j]);
                }
                inner[i][j] = mfma(a[0][i], b[1][j], inner[i][j]);
                inner[i][j] = mfma(a[1][i], b[0][j], inner[i][j]);
                inner[i][j] = mfma(a[0][i], b[0][j], inner[i][j]);
            }
        if ((kt + 1) % kFoldSlices == 0 || kt + 1 == p.nkt) {
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    outer[i][j] += inner[i][j];
                    inner[i][j] = f32x16{};
                }
        }
    };

    const int64_t nkt = p.nkt;
    uint4 x0, x1, x2, x3, x4, x5;  // even slices
    uint4 y0, y1, y2, y3, y4, y5;  // odd slices
    OMB_GLOAD(0, x0, x1, x2, x3, x4, x5);
    if (nkt > 1) OMB_GLOAD(1, y0, y1, y2, y3, y4, y5);
    OMB_LSTORE(lds[0], x0, x1, x2, x3, x4, x5);
    __syncthreads();
    for (int64_t kt = 0; kt < nkt; kt += 2) {
        // even slice kt from LDS 0; y* hold kt+1; load kt+2 into x*
        if (kt + 2 < nkt) OMB_GLOAD(kt + 2, x0, x1, x2, x3, x4, x5);
        compute(0, kt);
        if (kt + 1 < nkt) OMB_LSTORE(lds[1], y0, y1, y2, y3, y4, y5);
        __syncthreads();
        if (kt + 1 >= nkt) break;
        // odd slice kt+1 from LDS 1; x* hold kt+2; load kt+3 into y*
        if (kt + 3 < nkt) OMB_GLOAD(kt + 3, y0, y1, y2, y3, y4, y5);
        compute(1, kt + 1);
        if (kt + 2 < nkt) OMB_LSTORE(lds[0], x0, x1, x2, x3, x4, x5);
        __syncthreads();
    }

    // Epilogue.  32x32 accumulator layout: lane l holds column l % 32, value r
    // holds row 8*(r/4) + 4*(l/32) + r%4.
    const int64_t m_base = tm * TM + wm * 64, n_base = tn * TN + wn * 64;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int64_t n = n_base + j * 32 + (lane & 31);
            if (n >= p.N) continue;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t m = m_base + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (m < p.M) {
                    float* c = p.C + m * p.ldc + n;
                    float v = p.alpha * outer[i][j][r];
                    *c = p.beta != 0.f ? v + p.beta * *c : v;
                }
            }
        }
}

}  // namespace

int64_t split_packed_elems(int64_t R, int64_t K) {
    if (R <= 0 || K <= 0) return 0;
    return 3 * round_up(R, TM) * round_up(K, TK);
}

omb_status launch_split_pack(int64_t R, int64_t K, const float* X, int64_t rs, int64_t cs,
                             __bf16* out, hipStream_t s) {
    if (R <= 0 || K <= 0) return OMB_OK;
    const int64_t Rp = round_up(R, TM), Kp = round_up(K, TK);
    const int64_t total = Rp * (Kp >> 3);
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(ceil_div(total, NT), 8192));
    const int k_major = (cs == 1) ? 1 : 0;
    hipLaunchKernelGGL(split_pack_kernel, dim3((unsigned)blocks), dim3(NT), 0, s, X, R, K, rs, cs,
                       Rp, Kp, (int)TM, k_major, out);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

omb_status launch_sgemm_packed(int64_t M, int64_t N, int64_t K, float alpha, const __bf16* ap,
                               const __bf16* bp, float beta, float* C, int64_t ldc, int nprod,
                               hipStream_t s) {
    const int64_t tiles_m = ceil_div(M, TM), tiles_n = ceil_div(N, TN);
    if (tiles_m * tiles_n > 0x7fffffffLL) {
        set_error("sgemm split: too many tiles for one launch");
        return OMB_EUNSUPPORTED;
    }
    MArgs a{ap, bp, ceil_div(K, TK), C, ldc, M, N, alpha, beta, tiles_m, tiles_n};
    dim3 grid((unsigned)(tiles_m * tiles_n)), block(NT);
    if (nprod == 3)
        hipLaunchKernelGGL(sgemm_bf16x3_kernel<3>, grid, block, 0, s, a);
    else
        hipLaunchKernelGGL(sgemm_bf16x3_kernel<6>, grid, block, 0, s, a);
    OMB_HIP_TRY(hipGetLastError());
    return OMB_OK;
}

omb_status launch_sgemm_split(int64_t M, int64_t N, int64_t K, float alpha,
                              const float* A, int64_t rsA, int64_t csA,
                              const float* B, int64_t rsB, int64_t csB,
                              float beta, float* C, int64_t ldc, int nprod, hipStream_t s) {
    const size_t a_elems = (size_t)split_packed_elems(M, K);
    const size_t b_elems = (size_t)split_packed_elems(N, K);
    Workspace* ws = nullptr;
    OMB_TRY(get_workspace(s, (a_elems + b_elems) * sizeof(__bf16), &ws));
    __bf16* ap = reinterpret_cast<__bf16*>(ws->buf);
    __bf16* bp = ap + a_elems;
    // A as rows m, columns k; B^T as rows n, columns k (element (n,k) = B(k,n)).
    OMB_TRY(launch_split_pack(M, K, A, rsA, csA, ap, s));
    OMB_TRY(launch_split_pack(N, K, B, csB, rsB, bp, s));
    return launch_sgemm_packed(M, N, K, alpha, ap, bp, beta, C, ldc, nprod, s);
}

}  // namespace omb
