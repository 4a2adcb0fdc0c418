/*
 * lu_oracle.h -- CPU restatement of charmlu's block LU with partial pivoting.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (charmlu_amd/) links,
 * loads or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the test oracle / the timed
 * CPU baseline.
 *
 * Every function restates the reference algorithm and cites the reference
 * file:line it follows.  All matrices are row-major, as the reference's
 * blocks are (lu.h:151-153, getIndex(i,j) = i*b + j).
 *
 * Parity pins (tests/test_oracle.py):
 *   matgen      -- C++ std::minstd_rand known answer (10000th value from
 *                  seed 1 is 399268537) and the first outputs of seed 9934835
 *   gen_uniform -- library-defined counter generator (splitmix64 finalizer);
 *                  pinned by a published splitmix64 test vector and by the
 *                  GPU generator producing the same bits
 *   blas ops    -- MKL cblas_* (when present), brute-force loops, closed forms
 *   getrf       -- scipy.linalg.lu_factor (LAPACK dgetrf) ipiv, PA=LU
 *                  backward error, brute-force unblocked LU at small n,
 *                  pivot cycles == sequential swaps
 *   solve       -- scipy.linalg.lu_solve, residual
 *   residual    -- closed form on hand-built inputs
 */
#ifndef CHARMLU_ORACLE_H
#define CHARMLU_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- generators ---------------------------------------------------------- */
/* b generator: MatGen (benchmark.C:27-74), Park-Miller minimal standard via
 * Schrage.  Writes count doubles starting after skip draws. */
void orc_matgen_doubles(int seed, long skip, long count, double *out);
int orc_matgen_nth_int(int seed, long k); /* k-th draw, k >= 1 */

/* A generator: library-defined counter-based uniform [-0.5, 0.5).
 * value(i,j) = (splitmix64(seed*phi ^ (i<<32 | j)) >> 11) * 2^-53 - 0.5 */
double orc_gen_value(uint64_t seed, long i, long j);
void orc_gen_uniform(uint64_t seed, long row0, long col0, long m, long n,
                     double *A, long lda);
/* diagonally dominant variant: a_ii = sum_j |a_ij| over the full row. */
void orc_gen_diagdom(uint64_t seed, long n, double *A, long lda);

/* ---- BLAS surface of platformBlas.h (row-major CBLAS semantics) ---------- */
/* C = alpha*A*B + beta*C  (lu.C:36-41 calls it with alpha=-1, beta=1) */
void orc_dgemm(int M, int N, int K, double alpha, const double *A, int lda,
               const double *B, int ldb, double beta, double *C, int ldc);
/* B = L^{-1} B, L unit lower (lu.C:187) */
void orc_dtrsm_llnu(int M, int N, const double *L, int ldl, double *B, int ldb);
/* A += alpha * x * y^T (lu.C:247-252) */
void orc_dger(int M, int N, double alpha, const double *x, int incx,
              const double *y, int incy, double *A, int lda);
/* y = alpha*A*x + beta*y (benchmark.C:494-496) */
void orc_dgemv(int M, int N, double alpha, const double *A, int lda,
               const double *x, double beta, double *y);

/* Switch the four block ops to a vendor CBLAS loaded with dlopen (MKL's
 * libmkl_rt.so).  Returns 0 on success; on failure the plain loops stay. */
int orc_use_cblas(const char *libpath);
int orc_using_cblas(void);

/* ---- pivot records and cycles (lu.C:300-352, lu.C:46-163) --------------- */
/* Turn a sequence of swaps (pos[k] <-> row[k]) into the reference's cycle
 * form.  seq_index has room for count+1 ints, seq for 2*count ints.
 * Returns the number of cycles. */
int orc_swaps_to_cycles(int count, const int *pos, const int *row,
                        int *seq_index, int *seq);
/* Apply cycles to the rows of a row-major matrix: new[s_i] = old[s_{i+1}]
 * circularly, columns [c0, c1). */
void orc_apply_cycles(int ncycles, const int *seq_index, const int *seq,
                      double *A, long lda, int c0, int c1);

/* ---- factorization, solve, residual ------------------------------------- */
/* Blocked right-looking LU of the n x n row-major A (lda >= n), block b,
 * pivot batch size batch (<=0 means b/4 as benchmark.C:95-98), rhs permuted
 * along with the rows (may be NULL).  ipiv[j] = global pivot row of column j.
 * Returns 0, or j+1 for the first column whose pivot has |u_jj| <= 100*eps
 * (the reference aborts there, lu.ci:339-341).  Returns -1 if n % b != 0
 * (benchmark.C:143-147).  nthreads > 1 runs the independent block updates
 * of a step on that many OpenMP threads (one block op per thread, as one
 * Charm++ PE per core). */
int orc_getrf(int n, int b, int batch, double *A, long lda, double *rhs,
              int *ipiv, int nthreads);
/* Near-tie log of the next orc_getrf calls (NULL pointers: off): for each
 * global column j, mag1[j] = |pivot|, row2[j] / mag2[j] = the best other
 * candidate row at that column (-1 / 0 if none).  Arrays of n entries. */
void orc_set_tie_log(int *row2, double *mag2, double *mag1);
/* Unblocked textbook LU with partial pivoting (LAPACK dgetf2 with division),
 * lowest index on ties.  Used as a brute-force pin for orc_getrf. */
int orc_getf2(int n, double *A, long lda, int *ipiv);
/* Block forward + backward solve on an already permuted rhs
 * (lu.ci:387-458). */
void orc_solve(int n, int b, const double *LU, long lda, double *rhs);
/* Apply ipiv (sequential swaps) to a vector. */
void orc_apply_ipiv(int n, const int *ipiv, double *v);
/* Reference scaled residual (benchmark.C:385-400, 476-573):
 * max|b-Ax| / ((max|a_ij| * max|x| + max|b|) * n * eps).
 * Also returns, via out4 if non-NULL, {max|a|, max|b|, max|x|, max|b-Ax|}
 * and via ns if non-NULL the normwise form ||Ax-b||_inf/(||A||_inf ||x||_inf n eps). */
double orc_residual(int n, const double *A, long lda, const double *x,
                    const double *b, double *out4, double *ns);

double orc_wtime(void);
/* Block steps of the next orc_getrf calls (0: all); a bounded sample. */
void orc_set_max_steps(int k);
/* The vendor LAPACK's multithreaded dgetrf + dgetrs (dlopen'ed from
 * libpath) on row-major A and b: the library baseline beside the
 * restatement.  Returns LAPACK's info; *seconds = factor + solve time. */
int orc_lapack_gesv(const char *libpath, int n, double *A, long lda, double *b, int *ipiv,
                    int nthreads, double *seconds);

#ifdef __cplusplus
}
#endif
#endif
