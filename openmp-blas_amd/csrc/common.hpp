// Internal declarations shared by the kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <string>

#include "ompblas.h"

namespace omb {

// ------------------------------------------------------------------ errors --
void set_error(const std::string& msg);
void clear_error();

#define OMB_HIP_TRY(expr)                                                       \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess) {                                                 \
            ::omb::set_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
            return OMB_EDEVICE;                                                 \
        }                                                                       \
    } while (0)

#define OMB_TRY(expr)                      \
    do {                                   \
        omb_status s_ = (expr);            \
        if (s_ != OMB_OK) return s_;       \
    } while (0)

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
__host__ __device__ inline int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

// --------------------------------------------------------------- workspace --
// Per-stream scratch memory: one grow-only device buffer.
struct Workspace {
    void* buf = nullptr;
    size_t bytes = 0;
};

// Compute units of the current device (cached per device).
int device_cu_count();

// Workspace of `stream` on the current device with at least `bytes` of buffer.
omb_status get_workspace(hipStream_t stream, size_t bytes, Workspace** ws);

// ------------------------------------------------------- matrix operands --
// A validated operand view: element (i, j) at p[i*rs + j*cs], one stride 1.
struct MatView {
    const float* p;
    int64_t rows, cols, rs, cs;
};

// Resolve a view into (row_contig?, ld): row_contig means the column index
// is the unit-stride one (cs == 1, "last_order"); ld is the other stride.
omb_status resolve_layout(const char* name, int64_t rows, int64_t cols, int64_t rs, int64_t cs,
                          bool* unit_col, int64_t* ld);

// ------------------------------------------------------------ kernels -----
// Plain FP32 SGEMM, row-major C (ldc), A K-contiguous iff a_kc, B N-contiguous
// iff b_nc.  C = alpha*A*B + beta*C.
// K splits for an M x N x K GEMM of 128 x 128 tiles (1: none) and the k per
// split, a multiple of `align` (the kernel's accumulation block), from a
// model of rounds x k per workgroup (at `rate_full` flop/s per workgroup
// with two per CU, `rate_alone` with one) plus the partials' traffic.  The
// first choice is exact: one block per split, whose in-order sum is the
// unsplit kernel's own fold, so the bits are the unsplit ones on any device
// (taken when the partials fit in 128 MiB and the model finds it 5%
// faster than no split).  For K > 512 a coarser split (at least 256 k each,
// deterministic, with a different rounding) replaces it only when it is
// more than 10% faster.  A launch with more than one split needs
// splits * M * N floats of partial sums, from `part` when given, else from
// the stream's workspace; launch_splitk_reduce adds them in split order.
int64_t gemm_k_splits(int64_t M, int64_t N, int64_t K, int64_t align, double rate_full,
                      double rate_alone, int64_t* kchunk);
int64_t sgemm_f32_splits(int64_t M, int64_t N, int64_t K, int64_t* kchunk);
int64_t sgemm_split_splits(int64_t M, int64_t N, int64_t K, int64_t* kchunk);
omb_status launch_splitk_reduce(int64_t M, int64_t N, int64_t splits, float alpha, float beta,
                                const float* part, float* C, int64_t ldc, hipStream_t s);
omb_status launch_sgemm_f32(int64_t M, int64_t N, int64_t K, float alpha,
                            const float* A, int64_t lda, bool a_kc,
                            const float* B, int64_t ldb, bool b_nc,
                            float beta, float* C, int64_t ldc, hipStream_t s,
                            float* part = nullptr);

// Split-BF16 SGEMM (row-major C after layout folding); operands as strided
// views: both are split into the packed form in the stream's workspace, then
// multiplied by the packed kernel.
omb_status launch_sgemm_split(int64_t M, int64_t N, int64_t K, float alpha,
                              const float* A, int64_t rsA, int64_t csA,
                              const float* B, int64_t rsB, int64_t csB,
                              float beta, float* C, int64_t ldc, int nprod, hipStream_t s);

// Packed split-BF16 operands (3 planes, 128-row x 16-k tiles, zero padded;
// planes = 2 keeps the first two, all the 3-product form reads, in
// split16_packed_elems elements).
int64_t split_packed_elems(int64_t R, int64_t K);
omb_status launch_split_pack(int64_t R, int64_t K, const float* X, int64_t rs, int64_t cs,
                             __bf16* out, hipStream_t s, int planes = 3);
// Both operands of a product, A (M x K) and B^T (N x K, as strided views), in
// one launch; the same packed form as two launch_split_pack calls.
omb_status launch_split_pack_pair(int64_t M, int64_t N, int64_t K, const float* A, int64_t rsA,
                                  int64_t csA, __bf16* ap, const float* Bt, int64_t rsB,
                                  int64_t csB, __bf16* bp, hipStream_t s, int planes = 3);
omb_status launch_sgemm_packed(int64_t M, int64_t N, int64_t K, float alpha, const __bf16* ap,
                               const __bf16* bp, float beta, float* C, int64_t ldc, int nprod,
                               hipStream_t s, float* part = nullptr, const int* ea = nullptr,
                               const int* eb = nullptr, int planes = 3);

// Scaled split-FP16 SGEMM: every row of A and column of B is scaled by a
// power of two that brings its largest magnitude into [2^14, 2^15), split
// into two FP16 pieces, and the
// three leading products (a0b1 + a1b0 + a0b0) run on the FP16 MFMA; the
// sums are scaled back by the row and column exponents.  Workspace: the two
// packed operands (split16_packed_elems each: 2-plane layout), then
// split16_extra_bytes(M, N), then any split partials.
int64_t split16_packed_elems(int64_t R, int64_t K);
size_t split16_extra_bytes(int64_t M, int64_t N);
omb_status launch_sgemm_split16(int64_t M, int64_t N, int64_t K, float alpha,
                                const float* A, int64_t rsA, int64_t csA,
                                const float* B, int64_t rsB, int64_t csB,
                                float beta, float* C, int64_t ldc, hipStream_t s);
// The same on a caller-supplied workspace of split16_ws_bytes(M, N, K) bytes
// (256-byte aligned).
size_t split16_ws_bytes(int64_t M, int64_t N, int64_t K);
omb_status launch_sgemm_split16_ws(int64_t M, int64_t N, int64_t K, float alpha,
                                   const float* A, int64_t rsA, int64_t csA,
                                   const float* B, int64_t rsB, int64_t csB,
                                   float beta, float* C, int64_t ldc, char* ws, hipStream_t s);

// Scale C (row-major) by beta (K == 0 case).
omb_status launch_scale(int64_t M, int64_t N, float beta, float* C, int64_t ldc, hipStream_t s);

omb_status launch_sgemv_row(int64_t M, int64_t N, float alpha, const float* A, int64_t lda,
                            const float* x, float beta, float* y, hipStream_t s);
omb_status launch_sgemv_col(int64_t M, int64_t N, float alpha, const float* A, int64_t lda,
                            const float* x, float beta, float* y, hipStream_t s);
omb_status launch_sdot(int64_t n, const float* x, const float* y, float* result, hipStream_t s);
// SDOT with a caller-supplied partial buffer of sdot_blocks(n) doubles.
int64_t sdot_blocks(int64_t n);
omb_status launch_sdot_ws(int64_t n, const float* x, const float* y, float* result,
                          double* part, hipStream_t s);
// SDOT with its partials in the library's static device scratch (no
// allocation); sdot_scratch_setup finds that scratch's address on the
// current device (once per device) and optionally returns it.
omb_status sdot_scratch_setup(double** out);
omb_status launch_sdot_noalloc(int64_t n, const float* x, const float* y, float* result,
                               hipStream_t s);

omb_status launch_fill_uniform(float* dst, int64_t n, uint64_t seed, uint64_t stream_id,
                               int64_t offset, hipStream_t s);

// Grouped launches without a plan object (the serving loop's batches): a
// device block for up to 1024 GEMMs, 1024 GEMVs and 65536 items, with its
// pinned staging block and counters.  batch_launch builds the items on the
// host, copies them on stream s and runs the persistent kernel; the caller
// launches again from the same batch only after the previous launch has
// completed.  Plain and split-BF16 (6 or 3 products) modes.
struct GroupBatch;
omb_status batch_create(GroupBatch** out);
void batch_destroy(GroupBatch* b);
bool batch_mode_supported(int mode);
omb_status batch_launch(GroupBatch* b, const omb_gemm_desc* gemms, int n_gemm,
                        const omb_gemv_desc* gemvs, int n_gemv, int mode, hipStream_t s);
// One launch of the persistent kernel over no items on stream s: the first
// launch of a kernel in a process costs milliseconds, which a caller timing
// its batches pays here, before its clock starts.
omb_status batch_warm(GroupBatch* b, hipStream_t s);

// ----------------------------------------------------- device helpers -----
#if defined(__HIPCC__)
// Counter-based generator (splitmix64, Steele/Lea/Flood 2014).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t rng_key(uint64_t seed, uint64_t stream) {
    return mix64(seed * 0x9E3779B97F4A7C15ULL + stream * 0xD1B54A32D192ED03ULL +
                 0x632BE59BD9B4E019ULL);
}
__host__ __device__ inline float rng_uniform(uint64_t key, uint64_t i) {
    uint64_t h = mix64(key + (i + 1) * 0x9E3779B97F4A7C15ULL);
    return (float)(h >> 40) * 0x1.0p-23f - 1.0f;
}

// 128-bit loads of 4 consecutive floats at any 4-byte aligned address (the
// type tells the compiler that only 4-byte alignment is known; gfx950 serves
// it with one dwordx4 load).  The _nt form does not keep the line in cache.
typedef float f32x4_u __attribute__((ext_vector_type(4), aligned(4)));
__device__ inline float4 ld4u_nt(const float* p) {
    f32x4_u v = __builtin_nontemporal_load(reinterpret_cast<const f32x4_u*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline float4 ld4u(const float* p) {
    f32x4_u v = *reinterpret_cast<const f32x4_u*>(p);
    return make_float4(v.x, v.y, v.z, v.w);
}
// Global-address-space views of pointers that arrive through memory (work
// descriptors): without them the compiler must emit flat loads, which also
// count against the LDS wait counter, so every LDS wait drains the global
// loads in flight.  Only for pointers into device (global) memory.
#define OMB_GLOBAL __attribute__((address_space(1)))
typedef const OMB_GLOBAL float gcfloat;
typedef OMB_GLOBAL float gfloat;
typedef const OMB_GLOBAL int gcint;
typedef OMB_GLOBAL int gint;
__device__ inline gcfloat* as_global(const float* p) { return (gcfloat*)p; }
__device__ inline gfloat* as_global(float* p) { return (gfloat*)p; }
__device__ inline gcint* as_global(const int* p) { return (gcint*)p; }
__device__ inline gint* as_global(int* p) { return (gint*)p; }
typedef const OMB_GLOBAL f32x4_u gcf32x4_u;
__device__ inline float4 ld4u_nt(gcfloat* p) {
    f32x4_u v = __builtin_nontemporal_load(reinterpret_cast<gcf32x4_u*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline float4 ld4u(gcfloat* p) {
    f32x4_u v = *reinterpret_cast<gcf32x4_u*>(p);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline float dot4(float4 a, float4 b, float acc) {
    acc = fmaf(a.x, b.x, acc);
    acc = fmaf(a.y, b.y, acc);
    acc = fmaf(a.z, b.z, acc);
    acc = fmaf(a.w, b.w, acc);
    return acc;
}

// Butterfly sum over the 64 lanes of a wavefront; every lane gets the total.
// The pairing is fixed, so the result is deterministic.
__device__ inline float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Map a 1-D block id to a tile id so that the blocks the hardware sends to
// one XCD (ids congruent mod 8) get a contiguous range of tiles.
__device__ inline int64_t xcd_remap(int64_t bid, int64_t nblocks) {
    constexpr int64_t kXcd = 8;
    int64_t full = (nblocks / kXcd) * kXcd;
    if (bid >= full) return bid;
    int64_t per = full / kXcd;
    return (bid % kXcd) * per + bid / kXcd;
}

// Tile id -> (tm, tn) in groups of `group` tile rows (column-major inside a
// group) so that tiles in flight together share A and B panels.
__device__ inline void group_tile(int64_t t, int64_t tiles_m, int64_t tiles_n, int64_t group,
                                  int64_t* tm, int64_t* tn) {
    int64_t per_group = group * tiles_n;
    int64_t g = t / per_group;
    int64_t first = g * group;
    int64_t gs = tiles_m - first < group ? tiles_m - first : group;
    int64_t r = t - g * per_group;
    *tm = first + r % gs;
    *tn = r / gs;
}
#endif

}  // namespace omb
