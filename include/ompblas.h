/*
 * ompblas.h -- C ABI of the MI355X (gfx950) BLAS hot path: mtm (SGEMM),
 * mtv / vtm (SGEMV) and dot (SDOT), plus the split-BF16 GEMM mode, the
 * grouped persistent launch and the request-serving loop.
 *
 * Conventions
 *   - Every matrix is described as in the reference's uBLAS helpers: extents
 *     {rows, cols} and strides {rs, cs}, element (i, j) at p[i*rs + j*cs].
 *     One of rs / cs must be 1 (first_order: rs == 1, last_order: cs == 1)
 *     and the other at least the matching extent.
 *   - All data pointers are DEVICE pointers unless a name says otherwise.
 *   - Calls are asynchronous on `stream` (a hipStream_t; NULL = default
 *     stream).  Calls on different streams may run concurrently: each
 *     stream owns its library workspace.
 *   - No exception crosses this boundary.  Status codes are <= 0;
 *     omb_last_error() gives the calling thread's last message.
 */
#ifndef OMPBLAS_H
#define OMPBLAS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t omb_status;
#define OMB_OK 0
#define OMB_EINVAL (-1)       /* bad argument (null pointer, negative size, ld < extent) */
#define OMB_EUNSUPPORTED (-2) /* valid request this library does not implement */
#define OMB_ENOMEM (-3)       /* device memory could not be obtained */
#define OMB_EDEVICE (-4)      /* a HIP runtime call failed */
#define OMB_EBUSY (-5)        /* admission refused (serving loop) */

typedef void* omb_stream_t; /* hipStream_t */

/* GEMM arithmetic modes. */
#define OMB_MODE_PLAIN_F32 0     /* FP32 FMA on the vector ALUs (no MFMA) */
#define OMB_MODE_SPLIT_BF16 1    /* FP32 operands split into 3 BF16 terms, 6 products on MFMA */
#define OMB_MODE_SPLIT_BF16_P3 2 /* same split, 3 leading products only (faster, less accurate) */
/* Every row of A and column of B scaled by a power of two (largest magnitude
 * into [2^14, 2^15)) and split into 2 FP16 terms; the 3 leading products on
 * the FP16 MFMA, FP32 accumulation, sums scaled back.  Each element x is
 * carried to within max(2^-22 |x|, 2^-39 amax), amax the largest magnitude
 * of its row (A) or column (B): normwise as accurate as plain FP32, while an
 * element far below amax keeps only that absolute accuracy.
 * Every GEMM entry point (omb_sgemm, grouped plans, the serving loop). */
#define OMB_MODE_SPLIT_FP16 3

/* Last error message of the calling thread ("" when none). */
const char* omb_last_error(void);

/* Initialise the library on the current device (optional; idempotent). */
omb_status omb_init(void);

/* Library version string. */
const char* omb_version(void);

/* ---------------------------------------------------------------- mtm ----
 * C = alpha*A*B + beta*C;  A: M x K, B: K x N, C: M x N.
 * Replaces amt::mtm_helper (include/mtm.hpp:116-122), which computes
 * C += A*B (alpha = beta = 1) for any first/last-order mix of A, B, C.
 * beta == 0 does not read C.  C must share no element with A or B (A and B
 * may alias each other); as in BLAS this is not checked, since disjoint
 * submatrix views of one buffer are legitimate and their spans overlap. */
omb_status omb_sgemm(int64_t M, int64_t N, int64_t K, float alpha,
                     const float* A, int64_t rsA, int64_t csA,
                     const float* B, int64_t rsB, int64_t csB,
                     float beta, float* C, int64_t rsC, int64_t csC,
                     int32_t mode, omb_stream_t stream);

/* Split-BF16 in two steps, for callers that reuse an operand across calls.
 * omb_split_pack splits an R x K FP32 operand (element (r,k) at
 * X[r*rs + k*cs]) into the packed 3-plane BF16 form; pack A as (M x K) and
 * B as its transpose (N x K: rs = csB, cs = rsB).  omb_sgemm_packed then
 * computes C = alpha*A*B + beta*C from the two packed operands with
 * nprod = 6 or 3 products.  `packed` needs omb_split_pack_bytes(R, K) bytes.
 * omb_sgemm(mode = OMB_MODE_SPLIT_BF16*) is exactly these three calls. */
size_t omb_split_pack_bytes(int64_t R, int64_t K);
omb_status omb_split_pack(int64_t R, int64_t K, const float* X, int64_t rs, int64_t cs,
                          void* packed, omb_stream_t stream);
omb_status omb_sgemm_packed(int64_t M, int64_t N, int64_t K, float alpha,
                            const void* A_packed, const void* Bt_packed, float beta,
                            float* C, int64_t rsC, int64_t csC, int32_t nprod,
                            omb_stream_t stream);

/* ---------------------------------------------------------------- mtv ----
 * y = alpha*A*x + beta*y;  A: M x N, x: N, y: M (contiguous vectors).
 * Replaces amt::mtv_helper first_order (include/mtv.hpp:15-22; y += A*x,
 * i.e. beta = 1) and last_order (include/mtv.hpp:72-79; y = A*x, beta = 0).
 * vtm (include/mtv.hpp:164-236) is this call on the transposed view
 * (swap M/N and rs/cs).  y must share no element with A or x (not
 * checked). */
omb_status omb_sgemv(int64_t M, int64_t N, float alpha,
                     const float* A, int64_t rsA, int64_t csA,
                     const float* x, float beta, float* y, omb_stream_t stream);

/* ---------------------------------------------------------------- dot ----
 * *result = sum_i x[i]*y[i] (overwrite; result is a device pointer).
 * Replaces amt::dot_prod_helper (include/dot.hpp:143-149).  Per-workgroup
 * sums are FP32; the cross-workgroup combine is FP64 when n >= 2^26. */
omb_status omb_sdot(int64_t n, const float* x, const float* y, float* result,
                     omb_stream_t stream);

/* omb_sdot without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and bits as omb_sdot; the per-workgroup partial sums (at most
 * 1024 doubles, 8 KiB, for any n) live in a static device array of the
 * library instead of the stream's workspace.  Every call on a device shares
 * that array: calls that overlap on one device (other streams, other host
 * threads) race and must be serialised by the caller.  Its address is
 * looked up once per device, in omb_init or on the first call there. */
omb_status omb_sdot_noalloc(int64_t n, const float* x, const float* y, float* result,
                            omb_stream_t stream);

/* --------------------------------------------------------- outer product ----
 * C += alpha * x * y^T, C M x N in either storage order (rsC, csC); every
 * element is fma(alpha * x[i], y[j], C[i][j]).  Replaces
 * amt::outer_prod_helper (include/outer.hpp:15-48; alpha = 1 there). */
omb_status omb_sger(int64_t M, int64_t N, float alpha, const float* x, const float* y,
                    float* C, int64_t rsC, int64_t csC, omb_stream_t stream);

/* ------------------------------------------------------------- transpose ----
 * Out of place: C (N x M, strides rsC, csC) = A^T (A M x N, strides rsA,
 * csA); A and C must not overlap.  Replaces amt::transpose_helper(...,
 * tag::outplace) (include/trans.hpp:23-45).
 * In place: the n x n matrix at A (rows ld apart, either storage order) is
 * replaced by its transpose.  Replaces transpose_helper(..., tag::inplace)
 * (include/trans.hpp:47-71). */
omb_status omb_stranspose(int64_t M, int64_t N, const float* A, int64_t rsA, int64_t csA,
                          float* C, int64_t rsC, int64_t csC, omb_stream_t stream);
omb_status omb_stranspose_inplace(int64_t n, float* A, int64_t ld, omb_stream_t stream);

/* ------------------------------------------------------- input filling ----
 * Counter-based uniform[-1,1) generator: element i of (seed, stream_id) is
 * top24(splitmix64(key(seed, stream_id), offset + i)) * 2^-23 - 1 (exact). */
omb_status omb_fill_uniform(float* dst, int64_t n, uint64_t seed, uint64_t stream_id,
                            int64_t offset, omb_stream_t stream);

/* ---------------------------------------------------- FP64 error checks ----
 * Measurement kernels (not part of the FP32 path): an FP64 GEMM / GEMV / dot
 * of FP32 inputs and an error norm.  C64 is row-major M x N (ld = N).
 * c0 (nullable, strides of C) is added to the FP64 product: C64 = c0 + A*B.
 * Both validated against host FP64 (Higham's gamma_K bound). */
omb_status omb_check_dgemm(int64_t M, int64_t N, int64_t K,
                           const float* A, int64_t rsA, int64_t csA,
                           const float* B, int64_t rsB, int64_t csB,
                           const float* c0, int64_t rsC, int64_t csC,
                           double* C64, omb_stream_t stream);
omb_status omb_check_dgemv(int64_t M, int64_t N, const float* A, int64_t rsA, int64_t csA,
                           const float* x, const float* y0, double* y64, omb_stream_t stream);
omb_status omb_check_ddot(int64_t n, const float* x, const float* y, double* r64,
                          omb_stream_t stream);
/* out[0] = max|C - C64|, out[1] = max|C64|, out[2] = sum (C - C64)^2,
 * out[3] = sum C64^2 (device array of 4 doubles).  Deterministic. */
omb_status omb_check_errnorm(int64_t M, int64_t N, const float* C, int64_t rsC, int64_t csC,
                             const double* C64, double* out, omb_stream_t stream);
/* VALU issue ceiling (measurement only, no counterpart in the reference):
 * `blocks` workgroups of 256 threads each run `iters` rounds of 64
 * independent FP32 multiply-adds per thread (blocks * 256 * iters * 128 FLOP)
 * in the instruction form of the plain SGEMM kernel.  Timing it gives the
 * FP32 FMA rate the CUs sustain at that occupancy.  sink: 4 device bytes. */
omb_status omb_calib_fmac(int64_t blocks, int64_t iters, float* sink, omb_stream_t stream);
/* MFMA throughput ceiling (measurement only, no counterpart in the
 * reference): `blocks` workgroups of 256 threads each run `iters` rounds of
 * 64 v_mfma_f32_32x32x16_bf16 per wavefront (blocks * 4 * iters * 64 * 32768
 * FLOP) on register operands -- random BF16 values (random != 0), each MFMA
 * pairing other fragments, or all zero -- with no memory traffic in the
 * loop.  Timing both gives the BF16 MFMA rate this GPU sustains with and
 * without the data-dependent power of random operands. */
omb_status omb_calib_mfma(int64_t blocks, int64_t iters, int32_t random, float* sink,
                          omb_stream_t stream);

/* ----------------------------------------------------- grouped launch ----
 * Thousands of independent GEMMs (C_i = alpha*A_i*B_i + beta*C_i, every
 * operand row-major with the given leading dimensions) and GEMVs
 * (y_j = alpha*A_j*x_j + beta*y_j, A_j row-major) executed by ONE
 * persistent kernel launch, captured in a hipGraph.  No counterpart in the
 * reference (it has no batched API); the per-instance semantics are those
 * of omb_sgemm / omb_sgemv, and each output is bitwise the direct call's,
 * except where omb_sgemm splits K into coarse blocks (K > 512 with few
 * output tiles, e.g. 2048^3 plain): grouped tiles keep the unsplit fold,
 * which rounds differently within the same error bound. */
typedef struct {
    int64_t M, N, K;
    const float* A; int64_t lda;
    const float* B; int64_t ldb;
    float* C; int64_t ldc;
    float alpha, beta;
} omb_gemm_desc;

typedef struct {
    int64_t M, N;
    const float* A; int64_t lda;
    const float* x;
    float* y;
    float alpha, beta;
} omb_gemv_desc;

typedef struct omb_group_plan omb_group_plan;

/* Build a plan: tile shapes from the autotuner table (see
 * omb_autotune_group), work list on the device, graph captured on `stream`. */
omb_status omb_group_plan_create(const omb_gemm_desc* gemms, int32_t n_gemm,
                                 const omb_gemv_desc* gemvs, int32_t n_gemv,
                                 int32_t mode, omb_stream_t stream, omb_group_plan** plan);
/* Replay the captured graph: one persistent kernel, preceded in the scaled
 * split-FP16 mode by one kernel that computes every row and column exponent
 * (plan-owned memory: launches of one plan must not overlap). */
omb_status omb_group_plan_launch(omb_group_plan* plan, omb_stream_t stream);
/* Run the same work without the graph (direct kernel launches). */
omb_status omb_group_plan_launch_direct(omb_group_plan* plan, omb_stream_t stream);
/* Describe the plan: number of work items, tiles per class, the chosen
 * tile class per shape bucket ("gemm:<cls> ..."), into buf (NUL-terminated). */
omb_status omb_group_plan_describe(const omb_group_plan* plan, char* buf, size_t len);
omb_status omb_group_plan_destroy(omb_group_plan* plan);

/* Autotuner: time every tile class on representative shapes of each
 * (M, N, K) log2 bucket present in `gemms`, keep the fastest per bucket in
 * the process-wide table, and optionally persist it to `table_path`
 * (text, one line per bucket; NULL = do not write).  Loading an existing
 * table: omb_autotune_load. */
omb_status omb_autotune_group(const omb_gemm_desc* gemms, int32_t n_gemm, int32_t mode,
                              const char* table_path, omb_stream_t stream);
omb_status omb_autotune_load(const char* table_path);

/* ------------------------------------------------------ serving loop ----
 * Continuous batching of mixed GEMM / GEMV / dot requests over device
 * resident operands.  Requests are replayed from a trace of arrival times;
 * each request is admitted when its HBM need fits the budget and a slot
 * (stream) is free, runs concurrently with requests already on the GPU, and
 * its completion time is recorded.  The call first waits for all queued
 * device work (so operands written on any stream are ready) and returns when
 * every admitted request has finished.  A request whose workspace exceeds
 * the budget is refused with OMB_EBUSY in its result; the others still run.
 * Each result is bitwise the same as the request's solo run through
 * omb_sgemm / omb_sgemv / omb_sdot.  No counterpart in the reference. */
#define OMB_REQ_GEMM 0
#define OMB_REQ_GEMV 1
#define OMB_REQ_DOT 2

typedef struct {
    int32_t kind;       /* OMB_REQ_* */
    int32_t mode;       /* GEMM arithmetic mode (OMB_MODE_*) */
    int64_t m, n, k;    /* GEMM: M,N,K; GEMV: rows=m, cols=n; dot: n */
    double arrival_s;   /* arrival time, seconds from the start of the run */
    const float* a;     /* device operands (row-major, contiguous) */
    const float* b;     /* GEMM: B (K x N); GEMV: x; dot: y */
    float* c;           /* GEMM: C (M x N); GEMV: y; dot: result (1 float) */
    int64_t ws_bytes;   /* HBM the request needs beyond its operands (filled in) */
} omb_request;

typedef struct {
    int32_t slots;          /* max requests in flight (one HIP stream each) */
    int64_t hbm_budget;     /* bytes of workspace the loop may hold at once (0 = free HBM - reserve) */
    int64_t hbm_reserve;    /* bytes kept free when hbm_budget == 0 */
    int32_t spin_poll;      /* 1 = busy-poll completions, 0 = yield between polls */
    int32_t no_batching;    /* 0 = while at least `slots` requests wait, small GEMM / GEMV
                             * requests eligible together (two or more) share one grouped
                             * launch on one slot; 1 = a stream per request */
} omb_serve_config;

typedef struct {
    double start_s;     /* when the request was launched (from run start) */
    double done_s;      /* when its completion was observed */
    double latency_s;   /* done_s - arrival_s */
    int32_t status;     /* OMB_OK or an error code */
    int32_t slot;       /* stream slot it ran on */
} omb_request_result;

typedef struct {
    double wall_s;
    double p50_s, p90_s, p99_s, max_s, mean_s;
    double total_flop;      /* sum of 2MNK / 2MN / 2n */
    int32_t max_in_flight;  /* most requests observed on the GPU at once */
    int32_t admitted_late;  /* requests that waited for a slot or for memory */
    int32_t batches;        /* grouped launches of small requests */
    int32_t batched_requests; /* requests that ran in them */
} omb_serve_stats;

omb_status omb_serve_run(omb_request* reqs, int32_t n, const omb_serve_config* cfg,
                         omb_request_result* results, omb_serve_stats* stats);
/* Live submission: the same admission, batching and slots as omb_serve_run,
 * on a worker thread that takes requests as they are submitted instead of
 * from a trace (no counterpart in the reference, whose calls are
 * synchronous; each request runs one of its mtm / mtv / dot calls,
 * include/mtm.hpp:116-206, mtv.hpp:72-98, dot.hpp:143-174).
 * omb_server_open makes the slots' streams and a workspace arena of
 * cfg->slots x max_request_bytes (within the HBM budget) on the current
 * device; a request needing more is refused with OMB_EBUSY in its result.
 * omb_server_submit queues a copy of *r (arrival_s is set to the submission
 * time) and returns its ticket (0, 1, 2, ... in submission order); its
 * operands must be complete on the device when it is submitted (the slots'
 * streams do not wait for other streams) and stay valid until its result is
 * published.  omb_server_poll sets *done (and *result once done);
 * omb_server_wait blocks until the ticket's result is published.
 * omb_server_close waits for every submitted request, frees the server and
 * fills *stats (latencies from submission to observed completion).
 * With cfg->spin_poll the worker polls without sleeping, also while idle
 * (lowest latency, one host core busy); without it, it sleeps until a
 * submission arrives.  Submit, poll and wait may be called from any
 * threads; close once, after the last submit and after every wait and poll
 * has returned (it frees the server).  A server keeps each submission's
 * request and result (about 200 bytes) until it is closed, and takes at
 * most INT32_MAX submissions: a long-lived caller reopens it. */
typedef struct omb_server omb_server;
omb_status omb_server_open(const omb_serve_config* cfg, int64_t max_request_bytes,
                           omb_server** server);
omb_status omb_server_submit(omb_server* server, const omb_request* r, int64_t* ticket);
omb_status omb_server_poll(omb_server* server, int64_t ticket, int32_t* done,
                           omb_request_result* result);
omb_status omb_server_wait(omb_server* server, int64_t ticket, omb_request_result* result);
omb_status omb_server_close(omb_server* server, omb_serve_stats* stats);
/* The workspace bytes omb_serve_run will hold for request r while it runs
 * (the value it fills into ws_bytes), for sizing hbm_budget; -1 for an
 * invalid request.  Split-BF16 GEMMs hold their packed operands, GEMMs
 * whose K is split hold the partial sums, dots their partials. */
int64_t omb_serve_request_bytes(const omb_request* r);

/* ----------------------------------------------- host-pointer helpers ----
 * Used by the C++ drop-in layer to accept the reference's host buffers. */
int32_t omb_is_device_pointer(const void* p);
omb_status omb_device_alloc(void** p, size_t bytes);
omb_status omb_device_free(void* p);
omb_status omb_memcpy_h2d(void* dst, const void* src, size_t bytes, omb_stream_t stream);
omb_status omb_memcpy_d2h(void* dst, const void* src, size_t bytes, omb_stream_t stream);
omb_status omb_stream_synchronize(omb_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* OMPBLAS_H */
