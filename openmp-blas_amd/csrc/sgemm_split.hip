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
// Slices are staged through LDS (double buffered).  a0b0 accumulates into
// one register set and the smaller products into a second one, so adding
// the small terms never rounds at the magnitude of the running sum.
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

template <int NPROD>
__global__ __launch_bounds__(NT, 2) void sgemm_bf16x3_kernel(MArgs p) {
    __shared__ __attribute__((aligned(16))) __bf16 lds[2][SLICE_A + SLICE_B];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t ntiles = p.tiles_m * p.tiles_n;
    int64_t t = xcd_remap(blockIdx.x, ntiles);
    int64_t tm, tn;
    group_tile(t, p.tiles_m, p.tiles_n, kGroupM, &tm, &tn);

    const uint4* ga = reinterpret_cast<const uint4*>(p.Ap + tm * p.nkt * SLICE_A);
    const uint4* gb = reinterpret_cast<const uint4*>(p.Bp + tn * p.nkt * SLICE_B);
    constexpr int VA = SLICE_A / 8 / NT;  // 16-byte loads per thread per slice (3)
    constexpr int VB = SLICE_B / 8 / NT;
    uint4 ra[VA], rb[VB];

    auto gload = [&](int64_t kt) {
        const uint4* a = ga + kt * (SLICE_A / 8);
        const uint4* b = gb + kt * (SLICE_B / 8);
#pragma unroll
        for (int i = 0; i < VA; ++i) ra[i] = a[tid + i * NT];
#pragma unroll
        for (int i = 0; i < VB; ++i) rb[i] = b[tid + i * NT];
    };
    auto lstore = [&](int buf) {
        uint4* l = reinterpret_cast<uint4*>(lds[buf]);
#pragma unroll
        for (int i = 0; i < VA; ++i) l[tid + i * NT] = ra[i];
#pragma unroll
        for (int i = 0; i < VB; ++i) l[SLICE_A / 8 + tid + i * NT] = rb[i];
    };

    f32x16 hi[2][2], lo[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) hi[i][j][r] = lo[i][j][r] = 0.f;

    gload(0);
    lstore(0);
    __syncthreads();

    constexpr int NPLANES = NPROD == 6 ? 3 : 2;
    // Fragment offsets inside a slice: row (lane % 32), k half (lane / 32).
    const int frag = (lane & 31) * TK + (lane >> 5) * 8;
    for (int64_t kt = 0; kt < p.nkt; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < p.nkt;
        if (more) gload(kt + 1);
        const __bf16* la = lds[cur];
        const __bf16* lb = lds[cur] + SLICE_A;
        bf16x8 a[NPLANES][2], b[NPLANES][2];
#pragma unroll
        for (int pl = 0; pl < NPLANES; ++pl)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                a[pl][i] = *reinterpret_cast<const bf16x8*>(
                    la + pl * TM * TK + (wm * 64 + i * 32) * TK + frag);
                b[pl][i] = *reinterpret_cast<const bf16x8*>(
                    lb + pl * TN * TK + (wn * 64 + i * 32) * TK + frag);
            }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (NPROD == 6) {
                    lo[i][j] = mfma(a[0][i], b[2][j], lo[i][j]);
                    lo[i][j] = mfma(a[1][i], b[1][j], lo[i][j]);
                    lo[i][j] = mfma(a[2][i], b[0][j], lo[i][j]);
                }
                lo[i][j] = mfma(a[0][i], b[1][j], lo[i][j]);
                lo[i][j] = mfma(a[1][i], b[0][j], lo[i][j]);
                hi[i][j] = mfma(a[0][i], b[0][j], hi[i][j]);
            }
        if (more) lstore(cur ^ 1);
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
                    float v = p.alpha * (hi[i][j][r] + lo[i][j][r]);
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
