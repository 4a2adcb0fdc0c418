// kernels.h -- host-side launchers of the charmlu_amd device kernels.
// All matrices are row-major with a leading dimension in elements, as the
// reference's blocks are (lu.h:151-153).  Every launcher is asynchronous on
// the given stream and returns CLU_OK or a negative status.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clu {

// ---- GEMM: C = alpha * op(A) * op(B) + beta * C (FP64 MFMA) -------------
// ta: A is stored K x M (op(A) = A^T); tb: B is stored N x K.
int dgemm(bool ta, bool tb, int M, int N, int K, double alpha, const double *A,
          long lda, const double *B, long ldb, double beta, double *C, long ldc,
          hipStream_t s);
// dgemm NN for the factorization's look-ahead block column (alpha = +-1):
// the same kernels, the direct-to-LDS one under its own name for profiles
int dgemm_lookahead(int M, int N, int K, double alpha, const double *A, long lda,
                    const double *B, long ldb, double beta, double *C, long ldc, hipStream_t s);
// The factorization's trailing update (alpha = +-1): the direct-to-LDS
// kernel whenever dgemm_trailing_lds holds (16-byte aligned operands and C,
// M and N multiples of 128, K of 8), whatever the number of tiles.
int dgemm_trailing(int M, int N, int K, double alpha, const double *A, long lda,
                   const double *B, long ldb, double beta, double *C, long ldc, hipStream_t s);
bool dgemm_trailing_lds(int M, int N, int K, const double *A, long lda, const double *B,
                        long ldb, const double *C, long ldc);

// The same on FP32 operands (FP32 matrix cores).
int sgemm(bool ta, bool tb, int M, int N, int K, float alpha, const float *A, long lda,
          const float *B, long ldb, float beta, float *C, long ldc, hipStream_t s);

// C (FP32) = alpha * A * B^T + beta * C with A (M x K) and B (N x K) in a
// 16-bit format (T16 = __bf16 or _Float16), rows K-contiguous, K % 32 == 0,
// lda and ldb multiples of 8, 16-byte aligned bases.
template <typename T16>
int hgemm_nt(int M, int N, int K, float alpha, const T16 *A, long lda, const T16 *B, long ldb,
             float beta, float *C, long ldc, hipStream_t s);
// C = alpha * A * B^T + beta * C with FP8 e4m3 operands (A: M x K, B:
// N x K, K a multiple of 64) carrying power-of-two scales per 128-row tile
// (escA[M/128], escB[N/128]: the stored values are v * 2^esc).  With nseg
// > 1 (at most 16), K is nseg segments of kseg indices (a multiple of 64),
// segment q scaled by escA[q * estride + ...] and escB[q * estride + ...]:
// up to 16 products of their own scaling summed in one pass over C.
int fgemm_nt(int M, int N, int K, float alpha, const uint8_t *A, long lda, const int *escA,
             const uint8_t *B, long ldb, const int *escB, float beta, float *C, long ldc,
             hipStream_t s, int kseg = 0, int nseg = 1, long estride = 0);
// FP8 quantization with per-tile power-of-two scales: rows of src (M x K)
// in 128-row tiles (times sign), or columns of src (K x N) in 128-column
// tiles stored transposed; dst rows hold Kp bytes (Kp % 64 == 0, zero pad);
// esc gets one exponent per tile, part (scratch) FP8_PART_PER_TILE floats
// per tile.  cpy (optional, ld ldc): also receives src itself, in src's
// layout (the pass that reads src for the operand stores the factor there).
// ldd: distance between dst rows (0: Kp; larger: dst is one K segment of a
// wider operand).
constexpr int FP8_PART_PER_TILE = 8;
int quant_rows_fp8(int M, int K, int Kp, float sign, const float *src, long lds, uint8_t *dst,
                   int *esc, float *part, hipStream_t s, float *cpy = nullptr, long ldc = 0,
                   long ldd = 0);
int quant_cols_fp8(int K, int N, int Kp, const float *src, long lds, uint8_t *dst, int *esc,
                   float *part, hipStream_t s, float *cpy = nullptr, long ldc = 0,
                   long ldd = 0);
// dst (M x Kp) = 16-bit src (M x K) padded with zeros; dst (N x Kp) = 16-bit
// transpose of src (K x N) padded with zeros; cpy and ldd as for the FP8
// quantizers.
template <typename T16>
int pack_rows(int M, int K, int Kp, float scale, const float *src, long lds, T16 *dst,
              hipStream_t s, float *cpy = nullptr, long ldc = 0, long ldd = 0);
template <typename T16>
int pack_cols(int K, int N, int Kp, const float *src, long lds, T16 *dst, hipStream_t s,
              float *cpy = nullptr, long ldc = 0, long ldd = 0);

// ---- pivoting -------------------------------------------------------------
// Row move list of a run of sequential swaps ipiv[k1..k2) (global rows):
// moves[i] = {dst, src} means "row dst receives old row src"; count[0] is
// the number of moves.  Up to 2*(k2-k1) moves.
int pivots_to_moves(const int *ipiv, int k1, int k2, int2 *moves, int *count,
                    hipStream_t s);
// Inverse direction: row src receives old row dst (undoes the permutation).
// Apply a move list to columns [c0, c1) of a row-major matrix, skipping the
// columns [skip0, skip1); nruns > 1: the lists at moves + r * run_stride
// (lengths count[r]) in order, in one launch (reverse order when inverse).
template <typename T>
int apply_moves(const int2 *moves, const int *count, int max_moves, T *A,
                long lda, int c0, int c1, int skip0, int skip1, bool inverse,
                hipStream_t s, int nruns = 1, long run_stride = 0);
// Move lists of every run of 256 pivots of ipiv[k1..k2) at once (one
// workgroup per run): run r at moves + r*512, its length at counts[r].
int pivots_to_moves_all(const int *ipiv, int k1, int k2, int2 *moves, int *counts,
                        hipStream_t s);
// the interchanges ipiv[0..n) (or their inverse) applied to one vector;
// moves / counts: scratch for the move lists of all (n + 255) / 256 runs
template <typename T>
int permute_vec(const int *ipiv, int n, T *v, bool inverse, int2 *moves, int *counts,
                hipStream_t s);

// ---- panel factorization (FP64 / FP32) ------------------------------------
// 8-byte words of the panel exchange records (2 parities x 64 workgroups x
// 136 words): PanelWork::slots holds at least this many.
constexpr long PANEL_SLOT_WORDS = 2L * 64 * 136;
struct PanelWork {
  void *slots;        // per-workgroup candidate records, 2 parities x 256
  unsigned epoch;     // sequence number of the next column (host-side)
  int max_groups;     // workgroups of one strip that are resident at once
  int *info;          // device int, first near-singular column + 1
  unsigned *abortw;   // device word: a strip's exchange gave up (cleared per factorization)
};
// Launch shape of one strip: G workgroups of NT threads, R rows per thread,
// sw columns (the kernel's strip width, >= the strip's own width).
struct PanelShape {
  int sw, R, NT, G;
};
// Widest strip (32, 16 or 8 columns) panel_base takes for a strip of m rows
// when at most max_groups workgroups can be resident at once; 0 if none
// fits (every narrower or shorter strip then fits too: the factorizations
// check the tallest, first, strip before they enqueue anything).
template <typename T>
int panel_strip_width(int m, int max_groups, bool pivot = true);
// The launch shape panel_base uses for an m x w strip (0), or -1 if none.
template <typename T>
int panel_plan(int m, int w, int max_groups, bool pivot, PanelShape *ps);
// Factor rows [j0, m_end) x columns [j0, j0+w) (w <= panel_strip_width) of A
// in place with partial pivoting, one column at a time; ipiv[j0..j0+w) get
// global rows.
// The strip's row interchanges are also applied to the other columns
// [pc0, j0) and [j0+w, pc1) of the panel.  pivot = false: no interchanges
// (the diagonal is the pivot; one workgroup, m_end - j0 <= 1024).
template <typename T>
int panel_base(T *A, long lda, int j0, int m_end, int w, int pc0, int pc1, int *ipiv,
               PanelWork &pw, hipStream_t s, bool pivot = true);

// ---- triangular solves ----------------------------------------------------
// B = L^{-1} B with L unit lower (M x M), B is M x N (M <= 256).
template <typename T>
int trsm_llnu(int M, int N, const T *L, long ldl, T *B, long ldb, hipStream_t s);
// B = U^{-1} B with U non-unit upper (M x M), B is M x N (M <= 256).
template <typename T>
int trsm_lunn(int M, int N, const T *U, long ldu, T *B, long ldb, hipStream_t s);
// Transposed solves: B = L^{-T} B (L unit lower) and B = U^{-T} B.
template <typename T>
int trsm_lltu(int M, int N, const T *L, long ldl, T *B, long ldb, hipStream_t s);
template <typename T>
int trsm_lutn(int M, int N, const T *U, long ldu, T *B, long ldb, hipStream_t s);

// General small triangular solve, left side: B = alpha * op(F)^{-1} B
// (M <= 256); upper/trans/unit select the triangle, op and diagonal.
template <typename T>
int trsm_small(bool upper, bool trans, bool unit, int M, int N, T alpha, const T *F, long ldf,
               T *B, long ldb, hipStream_t s);

// A += alpha * x * y^T; y = alpha * op(A) * x + beta * y (CBLAS increments)
int dger(int M, int N, double alpha, const double *x, long incx, const double *y, long incy,
         double *A, long lda, hipStream_t s);
int dgemv(bool trans, int M, int N, double alpha, const double *A, long lda, const double *x,
          long incx, double beta, double *y, long incy, hipStream_t s);

// y[0..M) -= A[0..M, 0..K) * x  (trans: y[0..M) -= A[0..K, 0..M)^T * x)
// X = inverse of the n x n triangle of A (unit lower, or upper; unit
// selects an implied unit diagonal), every element of X written (zeros on
// the other side).  ntri triangles at once: triangle i at A + i*sa, its
// inverse at X + i*sx.  tbuf: scratch of ntri * n * n elements.
template <typename T>
int tri_inv(bool upper, bool unit, int n, int ntri, const T *A, long lda, long sa, T *X, long ldx,
            long sx, T *tbuf, hipStream_t s);
// strided-batched NN GEMM (grid rows y and layers z, strides per operand)
template <typename T>
int gemm_batched(int M, int N, int K, T alpha, const T *A, long lda, const T *B, long ldb,
                 T beta, T *C, long ldc, int ny, int nz, long a_y, long a_z, long b_y, long b_z,
                 long c_y, long c_z, hipStream_t s);
// y = op(G) x, G one m x m block inverse (m <= 1024)
template <typename T>
int blk_gemv(bool trans, int m, const T *G, long ldg, const T *x, T *y, hipStream_t s);
template <typename T>
int gemv_sub(bool trans, int M, int K, const T *A, long lda, const T *x, T *y,
             hipStream_t s);
// one triangle of the block solve of one right-hand side in one launch:
// out = op(T)^{-1} rhs, op(T) a triangle of the n x n factors F (trans: its
// transpose), solved from the last block up when upper is set (op(T) upper
// triangular), with the inverses of op's S x S diagonal blocks in Inv (S =
// 256 or 1024, n a multiple of S); tagx / tagt hold VW n published words
// each, seq is new for every call, abortw receives seq if a wait gave up
template <typename T>
int trsv_tiles(bool upper, bool trans, int n, int S, const T *F, long lda, const T *Inv,
               const T *rhs, T *out, unsigned long long *tagx, unsigned long long *tagt,
               unsigned seq, unsigned *abortw, hipStream_t s);
// in-place transpose of nblk contiguous S x S blocks (S a multiple of 32)
template <typename T>
int blk_transpose(int S, int nblk, T *G, hipStream_t s);

// ---- generation, residual, conversion -------------------------------------
int gen_uniform(uint64_t seed, long row0, long col0, long m, long n, double *A,
                long lda, hipStream_t s);
int gen_diagdom(uint64_t seed, long n, double *A, long lda, hipStream_t s);
int gen_rhs(int seed, long skip, long n, double *b, hipStream_t s);
// out[0..5] = {max|a_ij|, max|b|, max|x|, max|b-Ax|, max_i sum_j |a_ij|}
// r (optional) = b - A x.
int residual(int n, const double *A, long lda, const double *x, const double *b,
             double *r, double *out, hipStream_t s);
// out[0] = max|a_ij|, out[1] = max_j sum_i |a_ij| (out: 5 doubles, zeroed first)
int abs_colsum(int n, const double *A, long lda, double *out, hipStream_t s);
template <typename TD, typename TS>
int convert(long m, long n, const TS *src, long lds, TD *dst, long ldd, hipStream_t s);
int fill_int(int *p, int v, long n, hipStream_t s);
// *flag = 1 unless a panel eliminated without interchanges has every
// multiplier (strictly lower nb x nb A11, M x nb L21) below thr in magnitude
// and every |u_jj| above tiny (flag cleared first)
int spec_check(int M, int nb, const float *L21, long ld21, const float *A11, long lda,
               float thr, float tiny, int *flag, hipStream_t s);
// out[4 g + {0..3}] = {XCC id, HW_ID, start, end (100 MHz clock)} of each of
// `groups` one-per-CU workgroups that sleep `sleeps` x 8128 cycles
int residency_probe(int groups, int sleeps, unsigned long long *out, hipStream_t s);
// info[0] = -1 if abortw[0] != 0 (a panel strip's exchange gave up)
int finish_info(int *info, const unsigned *abortw, hipStream_t s);

// ---- FP64 vector kernels of the iterative refinement ----------------------
// h[0..k) = V w (V: k rows of n); w += alpha * V^T h; out = {||x||_2^2,
// ||x||_inf}; y = a x; y32 = (float)(a x); y = a (double)x32.
int basis_dot(int k, long n, const double *V, long ldv, const double *w, double *h,
              hipStream_t s);
int basis_axpy(int k, long n, double alpha, const double *V, long ldv, const double *h,
               double *w, hipStream_t s);
int vec_norms(long n, const double *x, double *out, hipStream_t s);
int vec_scale(long n, double a, const double *x, double *y, hipStream_t s);
int vec_d2s(long n, double a, const double *x, float *y, hipStream_t s);
int vec_s2d(long n, double a, const float *x, double *y, hipStream_t s);

}  // namespace clu
