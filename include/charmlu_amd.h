/*
 * charmlu_amd.h -- C ABI of the MI355X-native dense LU (libcharmlu_amd.so).
 *
 * Drop-in boundary for charmlu's block-update path.  The reference reaches
 * its arithmetic through the four CBLAS routines that platformBlas.h:4-38
 * selects; section 1 below exports those routines with the same argument
 * lists (row-major CBLAS semantics, host pointers, synchronous), so a
 * maintainer can point the reference's calls at them.  Section 2 gives the
 * same operations on device pointers, ordered on a HIP stream.  Section 3
 * replaces the whole chare-array factorization, solve and validation with
 * device-resident entry points.
 *
 * Conventions
 *   - Every matrix is row-major with a leading dimension in elements, as the
 *     reference's blocks are (lu.h:151-153, getIndex(i,j) = i*b + j).
 *   - Pivot indices are 0-based global rows: ipiv[j] = p means row j was
 *     exchanged with row p when column j was factored (the reference's
 *     recordPivot(j, p), lu.C:342-352).
 *   - stream is a hipStream_t passed as void* (NULL = the null stream).
 *   - Return value: 0 on success, -i if argument i is invalid (LAPACK
 *     style), CHARMLU_ERR_HIP if a HIP call failed, CHARMLU_ERR_NOMEM if the
 *     library could not allocate its workspace, CHARMLU_ERR_CONFIG if an
 *     environment setting of the library is out of range (CHARMLU_PANEL_CUS:
 *     0, or a multiple of 8 from 8 to 64), CHARMLU_ERR_UNSUPPORTED if the
 *     problem is too large for the
 *     panel CUs (the first panel strip needs more workgroups than can be
 *     resident at once: with 32 panel CUs, n <= 131072 in FP64 and
 *     n <= 262144 in FP32; charmlu_panel_plan tells).  charmlu_last_error()
 *     gives a message.  A factorization that fails after it has enqueued
 *     work orders the caller's stream after all of it before returning.
 *     Factorization status (singular pivot) is reported through an info
 *     word, never by aborting (the reference aborts, lu.ci:339-341).
 *   - Device-resident calls on one device share that device's workspace:
 *     calls that overlap in time on one device race.
 */
#ifndef CHARMLU_AMD_H
#define CHARMLU_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHARMLU_OK 0
#define CHARMLU_ERR_HIP (-100)
#define CHARMLU_ERR_NOMEM (-101)
#define CHARMLU_ERR_CONFIG (-102)
#define CHARMLU_ERR_UNSUPPORTED (-103)

/* CBLAS enumerations (public CBLAS values) */
#define CHARMLU_ROW_MAJOR 101
#define CHARMLU_COL_MAJOR 102
#define CHARMLU_NO_TRANS 111
#define CHARMLU_TRANS 112
#define CHARMLU_UPPER 121
#define CHARMLU_LOWER 122
#define CHARMLU_NON_UNIT 131
#define CHARMLU_UNIT 132
#define CHARMLU_LEFT 141
#define CHARMLU_RIGHT 142

/* precision of the trailing update in the mixed-precision factorization */
#define CHARMLU_PREC_FP64 0
#define CHARMLU_PREC_FP32 1
#define CHARMLU_PREC_BF16 2
#define CHARMLU_PREC_FP16 3
#define CHARMLU_PREC_FP8 4
/* or-ed into prec of charmlu_sgetrf_mixed: factor without interchanges */
#define CHARMLU_MIXED_NOPIV 0x100

int charmlu_version(void);
const char *charmlu_last_error(void);

/* Per-launch timing of charmlu_dgetrf's kernel classes with HIP events on
 * the streams they run on (no counterpart in the reference, which times the
 * whole run only, benchmark.C:242-245, 343-349).  charmlu_profile(1) starts
 * recording, charmlu_profile(0) stops; charmlu_profile_read fills out[15] =
 * {count, work, milliseconds} for the classes 0 trailing-update GEMM of the
 * direct-to-LDS 128x128 kernel (work = flops), 1 other trailing-update GEMMs
 * (flops), 2 panels (rows x columns), 3 row interchanges (columns moved per
 * step), 4 U-row solves (flops), then clears the records.  With nout >= 20,
 * out[15 + k] is the time during which at least one record of class k was
 * running (the two parts of a split trailing update overlap).
 * Synchronizes. */
int charmlu_profile(int on);
int charmlu_profile_read(double *out, int nout);
/* Make device `device` current for this thread and set up its workspace. */
int charmlu_init(int device);
/* Wait for every device the library has used, then release its streams,
 * events and workspaces (no counterpart in the reference, whose Charm++
 * runtime owns its resources).  The next call sets them up again.  The
 * Python binding calls it at interpreter exit. */
int charmlu_finalize(void);
/* Launch shape of one panel strip of m rows and w columns (no counterpart
 * in the reference; host arithmetic only, callable without a device):
 * prec CHARMLU_PREC_FP64 or CHARMLU_PREC_FP32, panel_cus as
 * CHARMLU_PANEL_CUS (0: no CU masks), ncu the device's CU count, pivot 0 or
 * 1.  out[6] = {kernel strip width, rows per thread, threads per
 * workgroup, workgroups, workgroups resident at once, widest strip for m
 * rows (0: none)}.  Returns 0 or CHARMLU_ERR_UNSUPPORTED. */
int charmlu_panel_plan(int prec, int m, int w, int panel_cus, int ncu, int pivot, int *out);
/* Co-residency check of the panel stream's CU mask (no counterpart in the
 * reference): launches `groups` workgroups of a kernel that holds one
 * workgroup per CU on the stream the panel strips run on; each sleeps
 * `sleeps` x 8128 cycles and writes {XCC id, hardware id, start, end} (the
 * end stamps on the 100 MHz constant clock) to dout[4 * groups] (device).
 * A workgroup that starts after another one ended was not resident with
 * it.  Ordered on `stream`; asynchronous. */
int charmlu_panel_probe(int groups, int sleeps, unsigned long long *dout, void *stream);

/* ---- 1. platformBlas.h surface: host pointers, synchronous ------------- */

/* Replaces cblas_dgemm(CblasRowMajor, CblasNoTrans, CblasNoTrans, b, b, b,
 * -1.0, L, b, U, b, 1.0, C, b) in LUBlk::updateMatrix (lu.C:36-41) and the
 * testdgemm call (benchmark.C:324-329).  order: ROW_MAJOR or COL_MAJOR;
 * transa/transb: NO_TRANS or TRANS. */
int charmlu_cblas_dgemm(int order, int transa, int transb, int M, int N, int K,
                        double alpha, const double *A, int lda, const double *B,
                        int ldb, double beta, double *C, int ldc);

/* Replaces cblas_dtrsm(CblasRowMajor, CblasLeft, CblasLower, CblasNoTrans,
 * CblasUnit, b, b, 1.0, L, b, A, b) in LUBlk::computeU (lu.C:187).
 * Supported: side LEFT; any uplo/trans/diag; M <= 256 per call on device
 * (larger M is solved in 256-row blocks). */
int charmlu_cblas_dtrsm(int order, int side, int uplo, int transa, int diag,
                        int M, int N, double alpha, const double *A, int lda,
                        double *B, int ldb);

/* Replaces cblas_dger(CblasRowMajor, M, N, -1.0, x, b, y, 1, A, b) in
 * LUBlk::updateLsubBlock (lu.C:247-252). */
int charmlu_cblas_dger(int order, int M, int N, double alpha, const double *x,
                       int incx, const double *y, int incy, double *A, int lda);

/* Replaces cblas_dgemv(CblasRowMajor, CblasNoTrans, b, b, 1.0, A, b, x, 1,
 * 0.0, y, 1) in BenchmarkLUBlk::recvXvec (benchmark.C:494-496). */
int charmlu_cblas_dgemv(int order, int trans, int M, int N, double alpha,
                        const double *A, int lda, const double *x, int incx,
                        double beta, double *y, int incy);

/* ---- 2. the same block ops on device pointers, stream-ordered ---------- */

/* C = alpha*op(A)*op(B) + beta*C, row-major. (lu.C:26-43) */
int charmlu_dgemm_dev(int transa, int transb, int M, int N, int K, double alpha,
                      const double *dA, long lda, const double *dB, long ldb,
                      double beta, double *dC, long ldc, void *stream);
/* B = op(T)^{-1} B for a triangular T (M <= 256), left side. (lu.C:180-189) */
int charmlu_dtrsm_dev(int uplo, int transa, int diag, int M, int N,
                      const double *dT, long ldt, double *dB, long ldb,
                      void *stream);
/* A += alpha * x * y^T. (lu.C:230-254) */
int charmlu_dger_dev(int M, int N, double alpha, const double *dx, long incx,
                     const double *dy, long incy, double *dA, long lda,
                     void *stream);

/* ---- 3. whole factorization, solve, generation, validation ------------- */

/* Blocked right-looking LU with partial pivoting of the n x n matrix dA in
 * place (the whole chare-array factorization: factor/updateUntilActive/
 * factorDiagBlock/factorBelowDiagBlock/factorAboveDiagBlock, lu.ci:31-370).
 * nb: block size (the reference's b; n % nb must be 0, benchmark.C:143-147).
 * dipiv: n device ints.  dinfo: one device int, 0 or the first column j+1
 * whose pivot has |u_jj| <= 100*eps (lu.ci:339-341), or -1 if the
 * cross-workgroup exchange of a panel strip timed out (a workgroup of the
 * strip was not resident; the factors are then invalid).  Asynchronous. */
int charmlu_dgetrf(int n, int nb, double *dA, long lda, int *dipiv, int *dinfo,
                   void *stream);

/* Solve op(A) X = B with the factors of charmlu_dgetrf: the row
 * interchanges, then forward and backward block substitution
 * (forwardSolve/backwardSolve/offDiagSolve, lu.ci:387-458).
 * trans: NO_TRANS or TRANS.  B is n x nrhs row-major, overwritten by X.
 * One right-hand side (nrhs = 1, ldb = 1) with NO_TRANS and n a multiple of
 * the solve's diagonal block (256) runs each triangle as one launch whose
 * workgroups wait for each other's published values; such a wait is bounded,
 * and one that gives up leaves NaN in the rows that depend on it (and makes
 * charmlu_dgesv_ir / _ex return CHARMLU_ERR_HIP).  CHARMLU_TRSV_TILES=0
 * selects two launches per diagonal block instead. */
int charmlu_dgetrs(int trans, int n, int nrhs, const double *dLU, long lda,
                   const int *dipiv, double *dB, long ldb, void *stream);

/* ---- 4. mixed precision ------------------------------------------------ */

/* The factorization of charmlu_dgetrf on an FP32 working matrix (the
 * reference is FP64 only: same algorithm, lu.ci:31-370).  The panel is
 * factored in FP32; the trailing update (updateMatrix, lu.C:26-43) runs with
 * its operands rounded to prec (PREC_FP32, PREC_BF16, PREC_FP16, or
 * PREC_FP8: e4m3 with a power-of-two scale per 128-row tile of L21 and per
 * 128-column tile of U12, nb a multiple of 128; FP32 accumulation).  dA: n
 * x n FP32 row-major, overwritten by L and U.  The pivots and info follow
 * charmlu_dgetrf.  prec | CHARMLU_MIXED_NOPIV factors without row
 * interchanges (ipiv[j] = j), for matrices that need none, such as
 * diagonally dominant ones; nb <= 2048 then.  The panel is then the LU of
 * the diagonal block plus L21 = A21 inv(U11).  If every computed |l_ij| <
 * 1, partial pivoting would have chosen the same (diagonal) pivots.
 * With partial pivoting and a BF16, FP16 or FP8 update (nb <= 2048), each
 * panel is first eliminated that way and kept if every multiplier is below
 * 1/2 in magnitude and no pivot is tiny (partial pivoting then keeps every
 * diagonal row); at the first panel that fails, the diagonal block is
 * restored and the factorization continues with exchanging panels (the
 * pivots are partial pivoting's either way; the check makes the host wait
 * once per panel, so the call returns when the last panel has been
 * checked; CHARMLU_SPECULATE=0 exchanges at every panel).  The unpivoted
 * low-precision schedule updates the rest of the matrix once per group of
 * steps (CHARMLU_GROUP, 1 to 16; 16 for FP8, 12 otherwise).  Asynchronous
 * otherwise. */
int charmlu_sgetrf_mixed(int n, int nb, int prec, float *dA, long lda, int *dipiv,
                         int *dinfo, void *stream);

/* dS (m x n FP32, row-major) = dA (FP64) rounded to nearest: the working
 * copy factored by charmlu_sgetrf_mixed (LAPACK's dlag2s, without its
 * overflow test).  Asynchronous. */
int charmlu_dlag2s(long m, long n, const double *dA, long lda, float *dS, long lds,
                   void *stream);

/* charmlu_dgemm_dev in FP32 (the FP32 form of the differentiable solve
 * forms its matrix gradient with it; no reference counterpart). */
int charmlu_sgemm_dev(int transa, int transb, int M, int N, int K, float alpha,
                      const float *dA, long lda, const float *dB, long ldb,
                      float beta, float *dC, long ldc, void *stream);

/* The low-precision trailing update of charmlu_sgetrf_mixed as a block op
 * (updateMatrix, lu.C:26-43, at reduced input precision):
 * C = alpha * A * B^T + beta * C with A (M x K) and B (N x K) row-major BF16
 * (prec = PREC_BF16) or FP16 (PREC_FP16), C FP32, products accumulated in
 * FP32.  K a multiple of 32, lda and ldb multiples of 8, A and B 16-byte
 * aligned. */
int charmlu_gemm_lp_dev(int prec, int M, int N, int K, float alpha, const void *dA,
                        long lda, const void *dB, long ldb, float beta, float *dC,
                        long ldc, void *stream);

/* The FP8 form of the update as a block op: C = alpha * q(A) * q(B) + beta *
 * C for FP32 A (M x K), B (K x N), C, where q() rounds to e4m3 after
 * scaling each 128-row tile of A and each 128-column tile of B by the
 * largest power of two that keeps it within +-448; products accumulate in
 * FP32, the scales undone exactly by the matrix cores' block scales.  Uses
 * the device workspace. */
int charmlu_gemm_fp8_dev(int M, int N, int K, float alpha, const float *dA, long lda,
                         const float *dB, long ldb, float beta, float *dC, long ldc,
                         void *stream);

/* charmlu_gemm_fp8_dev without any allocation (no counterpart in the
 * reference: for callers that cannot allocate device memory, such as code
 * under HIP graph capture, tight serving loops and harnesses that forbid
 * hipMalloc).  Same arguments, kernels and results as charmlu_gemm_fp8_dev,
 * but the quantized operands and their scales live in a static device
 * array of CHARMLU_NOALLOC_SCRATCH_BYTES that every device using the
 * library holds until the process ends: (M + N) * ceil(K / 64) * 64 +
 * 36 * (ceil(M / 128) + ceil(N / 128)) bytes must fit, else
 * CHARMLU_ERR_UNSUPPORTED.  It calls no allocating routine and never
 * touches the library's workspace.  Calls on one device share that array:
 * calls that overlap in time on one device race.  Asynchronous. */
#define CHARMLU_NOALLOC_SCRATCH_BYTES (32u << 20)
int charmlu_gemm_fp8_dev_noalloc(int M, int N, int K, float alpha, const float *dA, long lda,
                                 const float *dB, long ldb, float beta, float *dC, long ldc,
                                 void *stream);

/* charmlu_dgetrs on FP32 factors and an FP32 right-hand side. */
int charmlu_sgetrs(int trans, int n, int nrhs, const float *dLU, long lda,
                   const int *dipiv, float *dB, long ldb, void *stream);

/* Solve A x = b in FP64 by GMRES-based iterative refinement preconditioned
 * with the FP32 factors (dLU, dipiv) of charmlu_sgetrf_mixed.  dA is the
 * original FP64 matrix; x, residuals, the Krylov basis and inner products
 * are FP64.  Refinement stops once the reference's scaled residual
 * (benchmark.C:385-400) is <= tol, or after max_iter GMRES cycles of at most
 * restart (1..200) iterations.  stats (optional, host, 4 doubles) =
 * {cycles, inner iterations, final reference residual, final normwise
 * residual ||b-Ax||_inf / (||A||_inf ||x||_inf n eps)}.  Synchronous. */
int charmlu_dgesv_ir(int n, const double *dA, long lda, const float *dLU, long ldlu,
                     const int *dipiv, const double *db, double *dx, double tol,
                     int max_iter, int restart, double *stats, void *stream);

/* charmlu_dgesv_ir for op(A) X = B with several right-hand sides: trans
 * NO_TRANS or TRANS (A^T X = B, the adjoint solve of the differentiable
 * solve, from the same stored low-precision factors; the transposed
 * block solve of forwardSolve/backwardSolve, lu.ci:387-458).  B and X are
 * n x nrhs row-major (ldb, ldx >= nrhs); each column is refined on its own
 * to the same stopping test, with r = b - op(A) x in FP64.  stats
 * (optional, host, 4 doubles) = {most cycles of any column, inner
 * iterations of all columns, largest final reference residual, largest
 * final normwise residual (with ||A^T||_inf for TRANS)}.  Synchronous. */
int charmlu_dgesv_ir_ex(int trans, int n, int nrhs, const double *dA, long lda, const float *dLU,
                        long ldlu, const int *dipiv, const double *dB, long ldb, double *dX,
                        long ldx, double tol, int max_iter, int restart, double *stats,
                        void *stream);

/* Seeded inputs of the driver (benchmark.C:430-460).
 * gen_matrix: entry (i,j) of the library's counter generator, kind 0 =
 * uniform [-0.5,0.5), kind 1 = same with a_ii = sum_j |a_ij|.
 * gen_rhs: MatGen draws skip+1 .. skip+n of seed (benchmark.C:27-74). */
int charmlu_dgen_matrix(int kind, uint64_t seed, long n, double *dA, long lda,
                        void *stream);
int charmlu_dgen_rhs(int seed, long skip, long n, double *db, void *stream);

/* Scaled-residual inputs (benchmark.C:385-400, 462-573): out[5] (device) =
 * {max|a_ij|, max|b_i|, max|x_i|, max|b-Ax|_i, max_i sum_j |a_ij|}; dr
 * (optional, n doubles) receives b - A x.  Reference residual =
 * out[3] / ((out[0]*out[2] + out[1]) * n * eps). */
int charmlu_dresidual(int n, const double *dA, long lda, const double *dx,
                      const double *db, double *dr, double *dout, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CHARMLU_AMD_H */
