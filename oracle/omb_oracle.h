/*
 * omb_oracle.h -- CPU restatement of the OpenMP-BLAS reference algorithms
 * (mtm / mtv / vtm / dot) used ONLY as test infrastructure: the parity
 * checker in tests/, the CPU leg of bench.py and __graft_entry__.smoke().
 * The product library never links, loads or calls anything declared here.
 *
 * Every function cites the reference file:line it restates.  The reference
 * itself (header-only C++20 over Boost.uBLAS tensor) cannot be compiled in
 * this image (no Boost headers), so this restatement is the oracle.
 *
 * Parity pins (see tests/test_oracle.py):
 *   - oracle_cache_probe    : the host's cache parameters as getconf(1) reports
 *                             them, placed in the slots of
 *                             include/cache_manager.hpp:107-150 (L3
 *                             associativity in the L2 slot, L3's left at 1).
 *   - oracle_blocking       : doc tables (KB 384/256, MB 32/32, NB 4776/3582,
 *                             MR/NR 16x6 / 8x6) from doc/matrix_times_matrix.tex
 *                             :290-302,378-388,440-450,487-497 with the doc
 *                             laptop's cache geometry (doc/intro.tex:182-210).
 *   - oracle_pack*          : known answers of test/test.pack.cpp:35-43,74-82.
 *   - oracle_smtm           : integer-exact brute force on the 8 layout combos of
 *                             test/test.mtm.cpp:30-460 (sizes 2..31, rand()%100),
 *                             plus FP64 error bounds and the row-slab identity.
 *   - oracle_smtv/_svtm     : integer-exact brute force (test/test.mtv.cpp,
 *                             test/test.vtm.cpp sizes 2..511).
 *   - oracle_sdot           : closed form s(s+1)(2s+1)/6 (test/test.dot_prod.cpp).
 *   - oracle_souter*        : integer-exact brute force on the sizes of
 *                             test/test.outer_prod.cpp (2..31, both orders).
 *   - oracle_stranspose_*   : element-for-element brute force on the sizes of
 *                             test/test.trans.cpp (2..31, out of and in place)
 *                             and the block edge against the doc laptop L1.
 *   - omb_rng_*             : splitmix64 known-answer vectors (published
 *                             reference outputs of splitmix64 for seed 0).
 */
#ifndef OMB_ORACLE_H
#define OMB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One cache level as the reference's cache_info holds it
 * (include/cache_manager.hpp:30-40): size in bytes, associativity (int8 in
 * the reference), line size in bytes. */
typedef struct {
    int64_t size;
    int64_t assoc;
    int64_t line;
} oracle_cache_level;

typedef struct {
    oracle_cache_level l[3];
} oracle_cache_geom;

/* Blocking parameters the reference derives (include/mtm.hpp:19-81,
 * include/cpuinfo.hpp:168-223, include/mtv.hpp:30-37, include/dot.hpp:155-160). */
typedef struct {
    int64_t mr, nr, kc, mc, nc;        /* mtm */
    int64_t mtv_kb, mtv_mb;            /* first-order mtv */
    int64_t dot_sec1, dot_sec2, dot_sec3, dot_l3el; /* dot regimes */
} oracle_blocking_t;

/* Linux sysconf probe, including the reference's L3-associativity write into
 * the L2 slot (include/cache_manager.hpp:107-150). */
void oracle_cache_probe(oracle_cache_geom* g);

/* MR/NR from the Skylake AVX2 latency model, KC/MC/NC from the geometry.
 * elem_size = 4 (float) or 8 (double). */
int oracle_blocking(const oracle_cache_geom* g, int elem_size, int last_order_c,
                    oracle_blocking_t* out);

/* The reference's ct_sqrt (include/utils.hpp:185-201). */
double oracle_ct_sqrt(double x);

/* pack / pack(trans) (include/utils.hpp:99-141); wi = {row stride, col stride}. */
void oracle_pack_f32(float* out, int64_t wo, const float* in, const int64_t* wi,
                     int64_t m, int64_t n);
void oracle_pack_trans_f32(float* out, int64_t wo, const float* in, const int64_t* wi,
                           int64_t m, int64_t n);

/* mtm_helper (include/mtm.hpp:116-206): C += A*B, float.
 * n* = {rows, cols}; w* = {row stride, col stride} (uBLAS strides).
 * last_order_c selects the micro-kernel variant for C's layout.
 * blk may be NULL (then probed from this host's sysconf once).
 * nthreads <= 0 means omp_get_max_threads(). */
void oracle_smtm(float* c, const int64_t* nc, const int64_t* wc,
                 const float* a, const int64_t* na, const int64_t* wa,
                 const float* b, const int64_t* nb, const int64_t* wb,
                 int last_order_c, const oracle_blocking_t* blk, int nthreads);

/* mtv_helper first_order (include/mtv.hpp:15-70): y += A*x with col-major A.
 * mtv_helper last_order (include/mtv.hpp:72-98): y = A*x with row-major A.
 * M = rows (len y), N = cols (len x), wa = {row stride, col stride}. */
void oracle_smtv_first(float* y, const float* a, int64_t M, int64_t N,
                       const int64_t* wa, const float* x,
                       const oracle_blocking_t* blk, int nthreads);
void oracle_smtv_last(float* y, const float* a, int64_t M, int64_t N,
                      const int64_t* wa, const float* x, int nthreads);

/* dot_prod_helper (include/dot.hpp:143-174): returns sum a[i]*b[i]. */
float oracle_sdot(const float* a, const float* b, int64_t n,
                  const oracle_blocking_t* blk, int nthreads);

/* outer_prod_helper (include/outer.hpp:15-48): NB lines of C, WC apart,
 * line i += a * b[i].  oracle_souter: the closure's C (M x N) += a b^T for
 * either storage order (include/outer.hpp:104-120). */
void oracle_souter_lines(float* c, int64_t WC, const float* a, int64_t NA, const float* b,
                         int64_t NB, int nthreads);
void oracle_souter(float* c, int64_t M, int64_t N, int last_order, const float* a,
                   const float* b, int nthreads);

/* transpose_helper (include/trans.hpp:23-71): out of place C = A^T with
 * C strides wc, A extents na and strides wa; in place for a square
 * column-major n x n matrix.  block <= 0: the reference's edge for this
 * host's L1; oracle_transpose_block(l1_bytes) gives it for any L1 size
 * (l1_bytes <= 0: this host's). */
int64_t oracle_transpose_block(int64_t l1_bytes);
void oracle_stranspose_out(float* c, const int64_t* wc, const float* a, const int64_t* na,
                           const int64_t* wa, int64_t block);
void oracle_stranspose_in(float* a, int64_t n, int64_t block);

/* Counter-based input generator shared bit-for-bit with the device fill
 * kernel: value(seed, stream, i) = splitmix64 finaliser chain. */
uint64_t omb_splitmix64_at(uint64_t state, uint64_t i);
uint64_t omb_rng_u64(uint64_t seed, uint64_t stream, uint64_t i);
void omb_rng_fill_uniform(float* out, int64_t n, uint64_t seed, uint64_t stream,
                          int64_t offset);

/* Inputs of the reference's Catch2 tests: glibc rand() % 100, process-wide
 * stream (test/test_utils.hpp:4-9); oracle_srand resets it. */
void oracle_rand100_fill(float* out, int64_t n);
void oracle_srand(unsigned seed);

#ifdef __cplusplus
}
#endif
#endif
