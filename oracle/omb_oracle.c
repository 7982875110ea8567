/*
 * omb_oracle.c -- CPU restatement of the OpenMP-BLAS reference hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Used by tests/ (parity checker), bench.py's
 * cpu_baseline leg and __graft_entry__.smoke(); never by the product library.
 *
 * Written from the reference's documented algorithm; the loop structure,
 * blocking formulas and OpenMP constructs follow the cited lines because the
 * floating-point result of the reference is defined by them (accumulation
 * order inside a KC block, C += block, lane-parallel simd reductions).
 *
 * Build: oracle/Makefile (gcc -O3 -march=x86-64-v3 -ffast-math -fopenmp).
 */
#include "omb_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* Compile-time micro-tile of the float path: the reference's Skylake AVX2
 * model always yields MR=16, NR=6 for float (include/cpuinfo.hpp:210-223,
 * doc/matrix_times_matrix.tex:290-302); oracle_blocking recomputes it. */
#define OMB_MR 16
#define OMB_NR 6

static inline int64_t imin(int64_t a, int64_t b) { return a < b ? a : b; }
static inline int64_t imax(int64_t a, int64_t b) { return a > b ? a : b; }

/* ---------------------------------------------------------------------------
 * Cache probe: include/cache_manager.hpp:107-150.  Sizes and line sizes come
 * from sysconf; associativity defaults to 1.  The reference stores the L3
 * associativity into the L2 slot (:133-135) and leaves L3's at 1; that is
 * reproduced because it defines the reference's MC/NC on Linux.
 * ------------------------------------------------------------------------- */
void oracle_cache_probe(oracle_cache_geom* g) {
    memset(g, 0, sizeof(*g));
    for (int i = 0; i < 3; ++i) g->l[i].assoc = 1;
    long v;
    if ((v = sysconf(_SC_LEVEL1_DCACHE_SIZE)) > 0) g->l[0].size = v;
    if ((v = sysconf(_SC_LEVEL2_CACHE_SIZE)) > 0) g->l[1].size = v;
    if ((v = sysconf(_SC_LEVEL3_CACHE_SIZE)) > 0) g->l[2].size = v;
    if ((v = sysconf(_SC_LEVEL1_DCACHE_ASSOC)) > 0) g->l[0].assoc = (int8_t)v;
    if ((v = sysconf(_SC_LEVEL2_CACHE_ASSOC)) > 0) g->l[1].assoc = (int8_t)v;
    if ((v = sysconf(_SC_LEVEL3_CACHE_ASSOC)) > 0) g->l[1].assoc = (int8_t)v; /* sic */
    if ((v = sysconf(_SC_LEVEL1_DCACHE_LINESIZE)) > 0) g->l[0].line = v;
    if ((v = sysconf(_SC_LEVEL2_CACHE_LINESIZE)) > 0) g->l[1].line = v;
    if ((v = sysconf(_SC_LEVEL3_CACHE_LINESIZE)) > 0) g->l[2].line = v;
}

/* include/utils.hpp:185-201: bisection where the lower bound jumps to mid+1. */
double oracle_ct_sqrt(double res) {
    double l = 1.0, r = res;
    for (int it = 0; it < 100000 && l != r; ++it) {
        double mid = (r + l) / 2.0;
        if (mid * mid >= res) {
            if (mid == r) break;
            r = mid;
        } else {
            l = mid + 1.0;
        }
    }
    return r;
}

/* include/cpuinfo.hpp:186-191 */
static int64_t ref_ceil(double num) {
    int64_t ip = (int64_t)num;
    return (int64_t)(num > (double)ip ? num + 1.0 : num);
}

/* include/cpuinfo.hpp:36-45 (Skylake table) and :168-223 (mr/nr). */
static void skylake_mr_nr(int elem_size, int64_t* mr, int64_t* nr) {
    const double vec_len = 256.0;
    const double fma_lat = 4.0, mul_lat = 4.0, add_lat = 4.0;
    const double load_lat = 7.0, load_thr = 0.5, fma_thr = 0.5;
    double lat = (mul_lat != 0.0 && add_lat != 0.0) ? fma_lat + 2.0 * load_lat
                                                    : mul_lat + add_lat + 2.0 * load_lat;
    double tload = 2.0 / load_thr;
    double thr = 3.0 / (1.0 / fma_thr + tload);
    int64_t elements = (int64_t)(vec_len / (elem_size * 8));
    int64_t mrxnr = (int64_t)((double)elements * lat * thr);
    int64_t m = ref_ceil(oracle_ct_sqrt((double)mrxnr) / (double)elements) * elements;
    int64_t n = ref_ceil(((double)elements * lat * thr) / (double)m);
    *mr = m;
    *nr = n + (n & 1);
}

/* include/mtm.hpp:19-81 (kc, mc, nc), include/mtv.hpp:30-37, include/dot.hpp:
 * 155-160.  Zero line sizes or sizes would divide by zero in the reference;
 * here they make the call fail (return -1) instead. */
int oracle_blocking(const oracle_cache_geom* g, int elem_size, int last_order_c,
                    oracle_blocking_t* o) {
    memset(o, 0, sizeof(*o));
    int64_t mr, nr;
    skylake_mr_nr(elem_size, &mr, &nr);
    if (elem_size == 8 && last_order_c) { int64_t t = mr; mr = nr; nr = t; }
    o->mr = mr;
    o->nr = nr;
    for (int i = 0; i < 3; ++i)
        if (g->l[i].line <= 0 || g->l[i].size <= 0 || g->l[i].assoc <= 0) return -1;

    /* kc: L1 */
    int64_t a0 = g->l[0].assoc == 1 ? 8 : g->l[0].assoc;
    int64_t sets0 = g->l[0].size / (g->l[0].line * a0);
    double ratio = (double)nr / (double)mr;
    int64_t ca = (int64_t)ceil(((double)a0 - 1.0) / (1.0 + ratio));
    int64_t kc = (ca * sets0 * g->l[0].line) / (mr * elem_size);
    /* mc: L2 */
    int64_t a1 = g->l[1].assoc;
    int64_t sets1 = g->l[1].size / (g->l[1].line * a1);
    int64_t ca1 = (a1 - 1) >> 1;
    int64_t szm = kc > 0 ? (ca1 * sets1 * g->l[1].line) / (kc * elem_size) : 0;
    int64_t mc = szm - szm % mr;
    /* nc: L3 */
    int64_t a2 = g->l[2].assoc == 1 ? 16 : g->l[2].assoc;
    int64_t sets2 = g->l[2].size / (g->l[2].line * a2);
    int64_t cb = (a2 - 1) >> 1;
    int64_t szn = kc > 0 ? (cb * sets2 * g->l[2].line) / (kc * elem_size) : 0;
    int64_t nc = szn - szn % nr;
    o->kc = kc;
    o->mc = mc;
    o->nc = nc;

    /* first-order mtv (include/mtv.hpp:30-37); MR there is the float/double
     * calculate_mr value regardless of layout. */
    int64_t mr_v, nr_v;
    skylake_mr_nr(elem_size, &mr_v, &nr_v);
    int64_t kb = ((a0 - 1) * (sets0 * g->l[0].line)) / (elem_size * mr_v);
    int64_t mbv = kb > 0 ? ((g->l[1].assoc - 1) * sets1 * g->l[1].line) / (elem_size * kb) : 0;
    o->mtv_kb = kb;
    o->mtv_mb = mbv - mbv % mr_v;

    /* dot (include/dot.hpp:155-160) */
    int64_t l1el = g->l[0].size / elem_size;
    int64_t l2el = g->l[1].size / elem_size;
    o->dot_sec1 = l1el << 1;
    o->dot_sec2 = l1el >> 1;
    o->dot_sec3 = l2el >> 1;
    o->dot_l3el = g->l[2].size / elem_size;
    return 0;
}

static const oracle_blocking_t* host_blocking(void) {
    static oracle_blocking_t blk;
    static int done = 0;
    if (!done) {
        oracle_cache_geom g;
        oracle_cache_probe(&g);
        if (oracle_blocking(&g, 4, 0, &blk) != 0) {
            /* Fall back to the doc laptop geometry (doc/intro.tex:182-210). */
            oracle_cache_geom d = {{{32768, 8, 64}, {262144, 4, 64}, {16777216, 16, 64}}};
            oracle_blocking(&d, 4, 0, &blk);
        }
        done = 1;
    }
    return &blk;
}

/* ---------------------------------------------------------------------------
 * Packing: include/utils.hpp:99-118 (A panel: per column k, m consecutive rows)
 * and :120-141 (B panel, transposed: per row k, m consecutive columns).
 * ------------------------------------------------------------------------- */
void oracle_pack_f32(float* out, int64_t wo, const float* in, const int64_t* wi,
                     int64_t m, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        const float* in1 = in + wi[1] * i;
        float* out1 = out + wo * i;
#pragma omp simd
        for (int64_t j = 0; j < m; ++j) out1[j] = in1[j * wi[0]];
    }
}

void oracle_pack_trans_f32(float* out, int64_t wo, const float* in, const int64_t* wi,
                           int64_t m, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        const float* in1 = in + i * wi[0];
        float* out1 = out + wo * i;
#pragma omp simd
        for (int64_t j = 0; j < m; ++j) out1[j] = in1[j * wi[1]];
    }
}

/* ---------------------------------------------------------------------------
 * Micro-kernel: include/simd_loop.hpp:104-130 (float).  A full MR x NR tile is
 * always accumulated, as in the reference; lanes beyond mr/nr read panel slack
 * and are discarded.  Each C element gets sum_k a*b in k order (FMA under
 * -ffast-math), then C += buff (copy_from_buff, :160-190).
 * ------------------------------------------------------------------------- */
static inline void micro_f32(float* c, int64_t ldc, const float* a, const float* b,
                             int64_t K, int64_t mr, int64_t nr, int last_order_c) {
    float buff[OMB_MR * OMB_NR];
    memset(buff, 0, sizeof(buff));
    const float* ak = a;
    const float* bk = b;
    const int half = OMB_NR >> 1;
    for (int64_t k = 0; k < K; ++k, ak += mr, bk += nr) {
        for (int j = 0; j < half; ++j) {
#pragma omp simd
            for (int i = 0; i < OMB_MR; ++i) {
                buff[(j + 0) * OMB_MR + i] += bk[j + 0] * ak[i];
                buff[(j + half) * OMB_MR + i] += bk[j + half] * ak[i];
            }
        }
    }
    if (!last_order_c) {
        for (int64_t j = 0; j < nr; ++j) {
            float* cj = c + j * ldc;
            const float* bj = buff + j * OMB_MR;
#pragma omp simd
            for (int64_t i = 0; i < mr; ++i) cj[i] += bj[i];
        }
    } else {
        for (int64_t i = 0; i < mr; ++i) {
            float* ci = c + i * ldc;
#pragma omp simd
            for (int64_t j = 0; j < nr; ++j) ci[j] += buff[j * OMB_MR + i];
        }
    }
}

/* Macro-kernel: include/mtm.hpp:83-110. */
static void macro_f32(float* c, const int64_t* wc, const float* a, const float* b,
                      int64_t M, int64_t N, int64_t K, int last_order_c) {
    const int64_t ldc = imax(wc[0], wc[1]);
    for (int64_t j = 0; j < N; j += OMB_NR) {
        const float* bi = b + j * K;
        float* ci = c + wc[1] * j;
        int64_t jb = imin(OMB_NR, N - j);
        for (int64_t i = 0; i < M; i += OMB_MR) {
            const float* ak = a + i * K;
            float* ck = ci + wc[0] * i;
            int64_t ib = imin(OMB_MR, M - i);
            micro_f32(ck, ldc, ak, bi, K, ib, jb, last_order_c);
        }
    }
}

/* Process-lifetime packing storage, as the reference's static aligned_buff
 * (include/mtm.hpp:147-154); grow-only, zeroed once so slack reads are finite. */
static float* g_pa = NULL;
static size_t g_pa_n = 0;
static float* g_pb = NULL;
static size_t g_pb_n = 0;

static float* grow(float** p, size_t* have, size_t need) {
    if (*have < need) {
        free(*p);
        *p = (float*)aligned_alloc(64, ((need * sizeof(float) + 63) / 64) * 64);
        if (!*p) { *have = 0; return NULL; }
        memset(*p, 0, need * sizeof(float));
        *have = need;
    }
    return *p;
}

/* include/mtm.hpp:116-206 */
void oracle_smtm(float* c, const int64_t* nc, const int64_t* wc,
                 const float* a, const int64_t* na, const int64_t* wa,
                 const float* b, const int64_t* nb, const int64_t* wb,
                 int last_order_c, const oracle_blocking_t* blk, int nthreads) {
    (void)nc;
    if (!blk) blk = host_blocking();
    const int64_t M = na[0], K = na[1], N = nb[1];
    if (M <= 0 || N <= 0 || K <= 0) return;
    const int64_t MB = imax(blk->mc, OMB_MR), NB = imax(blk->nc, OMB_NR), KB = imax(blk->kc, 1);
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    const int64_t WA0 = wa[0], WA1 = wa[1], WB0 = wb[0], WB1 = wb[1], WC0 = wc[0], WC1 = wc[1];
    const size_t slack = (size_t)(OMB_MR * OMB_NR + 64);
    float* pA = grow(&g_pa, &g_pa_n, (size_t)KB * (size_t)(MB + 1) * (size_t)nt + slack);
    float* pB = grow(&g_pb, &g_pb_n, (size_t)KB * (size_t)(NB + 1) + slack);
    if (!pA || !pB) abort();

#pragma omp parallel num_threads(nt)
    {
        for (int64_t j = 0; j < N; j += NB) {
            const float* bk = b + WB1 * j;
            float* ck = c + WC1 * j;
            const int64_t jb = imin(NB, N - j);
            for (int64_t k = 0; k < K; k += KB) {
                const float* ai = a + WA1 * k;
                const float* bi = bk + WB0 * k;
                const int64_t kb = imin(KB, K - k);
#pragma omp for schedule(dynamic)
                for (int64_t jj = 0; jj < jb; jj += OMB_NR) {
                    int64_t jjb = imin(jb - jj, OMB_NR);
                    oracle_pack_trans_f32(pB + jj * kb, jjb, bi + jj * WB1, wb, jjb, kb);
                }
#pragma omp for schedule(dynamic)
                for (int64_t i = 0; i < M; i += MB) {
                    const int64_t tid = omp_get_thread_num();
                    const int64_t ib = imin(MB, M - i);
                    const float* aii = ai + WA0 * i;
                    float* aptr = pA + tid * kb * MB;
                    for (int64_t ii = 0; ii < ib; ii += OMB_MR) {
                        int64_t iib = imin(ib - ii, OMB_MR);
                        oracle_pack_f32(aptr + ii * kb, iib, aii + ii * WA0, wa, iib, kb);
                    }
                    macro_f32(ck + WC0 * i, wc, aptr, pB, ib, jb, kb, last_order_c);
                }
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * mtv, first order: include/mtv.hpp:15-70 with simd_loop<MTV,MR>
 * (include/simd_loop.hpp:52-78).  y += A*x.
 * ------------------------------------------------------------------------- */
static inline void mtv_tile(float* c, const float* a, int64_t W, const float* b, int64_t n) {
#pragma omp simd reduction(+ : c[0:OMB_MR])
    for (int64_t k = 0; k < n; ++k) {
        for (int64_t i = 0; i < OMB_MR; ++i) c[i] += a[i + k * W] * b[k];
    }
}

/* The reference reduces into c[0:MR] even when fewer than MR rows remain,
 * which touches y past its end on the last block; a private tile keeps the
 * same per-row arithmetic (private sum from 0, then c[i] += sum) in bounds. */
static inline void mtv_rem(float* c, const float* a, int64_t m, int64_t W, const float* b,
                           int64_t n) {
    float t[OMB_MR];
    memset(t, 0, sizeof(t));
#pragma omp simd reduction(+ : t[0:OMB_MR])
    for (int64_t i = 0; i < m; ++i) {
        const float* at = a + i;
        for (int64_t k = 0; k < n; ++k) t[i] += at[k * W] * b[k];
    }
    for (int64_t i = 0; i < m; ++i) c[i] += t[i];
}

void oracle_smtv_first(float* y, const float* a, int64_t M, int64_t N, const int64_t* wa,
                       const float* x, const oracle_blocking_t* blk, int nthreads) {
    if (!blk) blk = host_blocking();
    if (M <= 0 || N <= 0) return;
    const int64_t KB = imax(blk->mtv_kb, 1), MB = imax(blk->mtv_mb, OMB_MR);
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
    {
#pragma omp for
        for (int64_t i = 0; i < M; i += MB) {
            const int64_t ib = imin(M - i, MB);
            const float* ai = a + i * wa[0];
            float* ci = y + i;
            for (int64_t k = 0; k < N; k += KB) {
                const int64_t kb = imin(N - k, KB);
                const float* ak = ai + k * wa[1];
                const float* bk = x + k;
                float* ck = ci;
                const int64_t mitr = ib / OMB_MR, mrem = ib % OMB_MR;
                mtv_rem(ck, ak, mrem, wa[1], bk, kb);
                ak += mrem * wa[0];
                ck += mrem;
                for (int64_t ii = 0; ii < mitr; ++ii)
                    mtv_tile(ck + ii * OMB_MR, ak + ii * wa[0] * OMB_MR, wa[1], bk, kb);
            }
        }
    }
}

/* simd_loop<INNER> (include/simd_loop.hpp:21-35) */
static inline float inner_f32(const float* a, const float* b, int64_t n) {
    float sum = 0.0f;
#pragma omp simd reduction(+ : sum)
    for (int64_t i = 0; i < n; ++i) sum += a[i] * b[i];
    return sum;
}

/* mtv, last order: include/mtv.hpp:72-98.  y = A*x (overwrite). */
void oracle_smtv_last(float* y, const float* a, int64_t M, int64_t N, const int64_t* wa,
                      const float* x, int nthreads) {
    if (M <= 0) return;
    const int64_t WA = imax(wa[0], wa[1]);
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(static) if (M > 256) num_threads(nt)
    for (int64_t i = 0; i < M; ++i) y[i] = inner_f32(a + i * WA, x, N);
}

/* ---------------------------------------------------------------------------
 * outer product: include/outer.hpp:15-48 (outer_prod_helper) with
 * simd_loop<OUTER> (include/simd_loop.hpp:38-50): line i of C (WC apart)
 * gets c[k] += a[k] * b[i] for k < NA; lines in parallel when NB > 256.
 * The closure (include/outer.hpp:104-120) passes first-order C as columns
 * (wc = {1, M}) and last-order C as rows with a and b swapped.
 * ------------------------------------------------------------------------- */
void oracle_souter_lines(float* c, int64_t WC, const float* a, int64_t NA, const float* b,
                         int64_t NB, int nthreads) {
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(static) if (NB > 256) num_threads(nt)
    for (int64_t i = 0; i < NB; ++i) {
        float* cj = c + i * WC;
        const float cst = b[i];
#pragma omp simd
        for (int64_t k = 0; k < NA; ++k) cj[k] += a[k] * cst;
    }
}

void oracle_souter(float* c, int64_t M, int64_t N, int last_order, const float* a,
                   const float* b, int nthreads) {
    if (M <= 0 || N <= 0) return;
    if (last_order)
        oracle_souter_lines(c, N, b, N, a, M, nthreads);
    else
        oracle_souter_lines(c, M, a, M, b, N, nthreads);
}

/* ---------------------------------------------------------------------------
 * transpose: include/trans.hpp:11-71 with simd_loop<TRANS>
 * (include/simd_loop.hpp:193-240).  Block edge = sqrt_pow_of_two(L1 bytes /
 * (2 * sizeof(float))) (include/utils.hpp:49-58).
 * ------------------------------------------------------------------------- */
int64_t oracle_transpose_block(int64_t l1_bytes) {
    int64_t n = l1_bytes > 0 ? l1_bytes : sysconf(_SC_LEVEL1_DCACHE_SIZE);
    if (n <= 0) n = 32768;
    n /= 2 * (int64_t)sizeof(float);
    int64_t p = 0;
    n >>= 1;
    while (n) {
        n >>= 1;
        ++p;
    }
    p >>= 1;
    return (int64_t)1 << p;
}

void oracle_stranspose_out(float* c, const int64_t* wc, const float* a, const int64_t* na,
                           const int64_t* wa, int64_t block) {
    if (block <= 0) block = oracle_transpose_block(0);
    for (int64_t i = 0; i < na[0]; i += block) {
        const int64_t ib = imin(na[0] - i, block);
        for (int64_t j = 0; j < na[1]; j += block) {
            const int64_t jb = imin(na[1] - j, block);
            float* ck = c + i * wc[1] + j * wc[0];
            const float* ak = a + i * wa[0] + j * wa[1];
            for (int64_t ii = 0; ii < ib; ++ii)
                for (int64_t jj = 0; jj < jb; ++jj)
                    ck[ii * wc[1] + jj * wc[0]] = ak[ii * wa[0] + jj * wa[1]];
        }
    }
}

void oracle_stranspose_in(float* a, int64_t n, int64_t block) {
    if (block <= 0) block = oracle_transpose_block(0);
    /* the closure's strides (include/trans.hpp:158-163): wc = {1, n}, wa = {1, n} */
    for (int64_t i = 0; i < n; i += block) {
        const int64_t ib = imin(n - i, block);
        for (int64_t j = i; j < n; j += block) {
            const int64_t jb = imin(n - j, block);
            const int same = i == j;
            float* ck = a + i * n + j;  /* c + i*wc[1] + j*wc[0] */
            float* ak = a + i + j * n;  /* a + i*wa[0] + j*wa[1] */
            for (int64_t ii = 0; ii < ib && jb > ii * same; ++ii)
                for (int64_t jj = ii * same; jj < jb; ++jj) {
                    float* x = ck + ii * n + jj;
                    float* y = ak + ii + jj * n;
                    const float t = *x;
                    *x = *y;
                    *y = t;
                }
        }
    }
}

/* ---------------------------------------------------------------------------
 * dot: include/dot.hpp:16-53 (section two / three loops) and :143-174.
 * The reference changes the global OpenMP team size in section two
 * (:166-168); here the team size is a num_threads clause instead.
 * ------------------------------------------------------------------------- */
float oracle_sdot(const float* a, const float* b, int64_t n, const oracle_blocking_t* blk,
                  int nthreads) {
    if (!blk) blk = host_blocking();
    if (n <= 0) return 0.0f;
    int maxt = nthreads > 0 ? nthreads : omp_get_max_threads();
    const int64_t sec2 = imax(blk->dot_sec2, 1), sec3 = imax(blk->dot_sec3, 1);
    if (n < blk->dot_sec1) return inner_f32(a, b, n);
    if (n < blk->dot_l3el) {
        int64_t q = n / sec2;
        int nt = (int)imax(1, imin(maxt, q));
        float sum = 0.0f;
#pragma omp parallel for schedule(static) reduction(+ : sum) num_threads(nt)
        for (int64_t i = 0; i < n; i += sec2) {
            int64_t ib = imin(sec2, n - i);
            sum += inner_f32(a + i, b + i, ib);
        }
        return sum;
    }
    float sum = 0.0f;
#pragma omp parallel reduction(+ : sum) num_threads(maxt)
    {
        for (int64_t i = 0; i < n; i += sec3) {
            int64_t ib = imin(sec3, n - i);
            const float* ai = a + i;
            const float* bi = b + i;
#pragma omp for schedule(dynamic)
            for (int64_t j = 0; j < ib; j += sec2) {
                int64_t jb = imin(sec2, ib - j);
                sum += inner_f32(ai + j, bi + j, jb);
            }
        }
    }
    return sum;
}

/* ---------------------------------------------------------------------------
 * Input generator (not part of the reference; shared bit-for-bit with the
 * device fill kernel).  splitmix64 (Steele, Lea, Flood: "Fast splittable
 * pseudorandom number generators", OOPSLA 2014): output i of a generator with
 * state s is mix(s + (i+1)*gamma).
 * ------------------------------------------------------------------------- */
#define OMB_GAMMA 0x9E3779B97F4A7C15ULL

static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t omb_splitmix64_at(uint64_t state, uint64_t i) { return mix64(state + (i + 1) * OMB_GAMMA); }

static inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
    return mix64(seed * OMB_GAMMA + stream * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL);
}

uint64_t omb_rng_u64(uint64_t seed, uint64_t stream, uint64_t i) {
    return omb_splitmix64_at(stream_key(seed, stream), i);
}

/* The reference tests' inputs: rand() % 100 from glibc's process-wide
 * generator, never reseeded (test/test_utils.hpp:4-9). */
void oracle_rand100_fill(float* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = (float)(rand() % 100);
}

void oracle_srand(unsigned seed) { srand(seed); }

/* uniform[-1,1): top 24 bits u -> u * 2^-23 - 1, exact in FP32. */
void omb_rng_fill_uniform(float* out, int64_t n, uint64_t seed, uint64_t stream, int64_t offset) {
    const uint64_t key = stream_key(seed, stream);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        uint64_t h = omb_splitmix64_at(key, (uint64_t)(i + offset));
        out[i] = (float)(h >> 40) * 0x1.0p-23f - 1.0f;
    }
}
