/*
 * lu_oracle.c -- CPU restatement of charmlu's block LU with partial pivoting.
 *
 * TEST INFRASTRUCTURE ONLY (the test oracle and the CPU baseline); the product
 * never links or calls it.  See lu_oracle.h for what pins each function.
 *
 * The reference (Charm++/SDAG, lu.ci + lu.C) distributes b x b blocks over
 * processes; this file runs the same per-block operations in the order the
 * reference's dependencies force (restated in global indices), with a fixed
 * order wherever the reference's order is decided at run time (argmax ties:
 * lowest row; reduction sums: ascending block index).
 */
#define _GNU_SOURCE
#include "lu_oracle.h"

#include <dlfcn.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* generators                                                                */
/* ------------------------------------------------------------------------ */

/* MatGen::nextRndInt (benchmark.C:37-41): Schrage's method for
 * 48271 * s mod (2^31 - 1) with q = 44488, r = 3399 (benchmark.C:31-35). */
static int matgen_next(int *cur) {
  const int a = 48271, m = 2147483647; /* MAXINT = ~(1<<31), benchmark.C:28 */
  const int q = m / a, r = m % a;
  int s = a * (*cur % q) - r * (*cur / q);
  if (s < 0) s += m;
  *cur = s;
  return s;
}

int orc_matgen_nth_int(int seed, long k) {
  int cur = seed, v = seed;
  for (long i = 0; i < k; ++i) v = matgen_next(&cur);
  return v;
}

/* genVec (benchmark.C:454-460): block row x skips x*b draws and takes b, so
 * global row i holds draw i+1; toRndDouble = r / MAXINT - 0.5
 * (benchmark.C:44-46). */
void orc_matgen_doubles(int seed, long skip, long count, double *out) {
  int cur = seed;
  for (long i = 0; i < skip; ++i) matgen_next(&cur);
  for (long i = 0; i < count; ++i)
    out[i] = (double)matgen_next(&cur) / 2147483647 - 0.5;
}

/* Library-defined A generator (the reference draws A from Charm++'s CrnDouble,
 * benchmark.C:436-437,446-452, whose source is not available).  splitmix64's
 * output function (Steele, Lea, Flood 2014; constants as published by
 * Vigna) applied to a per-entry counter; 53 high bits -> [0,1) -> -0.5. */
static inline uint64_t splitmix64_mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double orc_gen_value(uint64_t seed, long i, long j) {
  uint64_t key = (seed * 0x9E3779B97F4A7C15ULL) ^
                 (((uint64_t)i << 32) | (uint64_t)(uint32_t)j);
  uint64_t u = splitmix64_mix(key) >> 11;
  return (double)u * 0x1.0p-53 - 0.5;
}

void orc_gen_uniform(uint64_t seed, long row0, long col0, long m, long n,
                     double *A, long lda) {
#pragma omp parallel for schedule(static) if (m * n >= (1L << 22))
  for (long i = 0; i < m; ++i)
    for (long j = 0; j < n; ++j)
      A[i * lda + j] = orc_gen_value(seed, row0 + i, col0 + j);
}

void orc_gen_diagdom(uint64_t seed, long n, double *A, long lda) {
#pragma omp parallel for schedule(static) if (n * n >= (1L << 22))
  for (long i = 0; i < n; ++i) {
    double s = 0.0;
    for (long j = 0; j < n; ++j) {
      double v = orc_gen_value(seed, i, j);
      A[i * lda + j] = v;
      s += fabs(v);
    }
    A[i * lda + i] = s;
  }
}

/* ------------------------------------------------------------------------ */
/* BLAS surface (platformBlas.h), row-major CBLAS semantics                  */
/* ------------------------------------------------------------------------ */

enum { CB_ROW = 101, CB_NOTRANS = 111, CB_UPPER = 121, CB_LOWER = 122,
       CB_NONUNIT = 131, CB_UNIT = 132, CB_LEFT = 141 };
typedef void (*cblas_dgemm_t)(int, int, int, int, int, int, double,
                              const double *, int, const double *, int, double,
                              double *, int);
typedef void (*cblas_dtrsm_t)(int, int, int, int, int, int, int, double,
                              const double *, int, double *, int);
typedef void (*cblas_dger_t)(int, int, int, double, const double *, int,
                             const double *, int, double *, int);
typedef void (*cblas_dgemv_t)(int, int, int, int, double, const double *, int,
                              const double *, int, double, double *, int);
static cblas_dgemm_t p_dgemm;
static cblas_dtrsm_t p_dtrsm;
static cblas_dger_t p_dger;
static cblas_dgemv_t p_dgemv;

int orc_use_cblas(const char *libpath) {
  void *h = dlopen(libpath, RTLD_NOW | RTLD_LOCAL);
  if (!h) return -1;
  /* one sequential BLAS per worker thread, as one Charm++ PE per core */
  int (*set_layer)(int) = (int (*)(int))dlsym(h, "MKL_Set_Threading_Layer");
  if (set_layer) set_layer(1 /* MKL_THREADING_SEQUENTIAL */);
  cblas_dgemm_t g = (cblas_dgemm_t)dlsym(h, "cblas_dgemm");
  cblas_dtrsm_t t = (cblas_dtrsm_t)dlsym(h, "cblas_dtrsm");
  cblas_dger_t r = (cblas_dger_t)dlsym(h, "cblas_dger");
  cblas_dgemv_t v = (cblas_dgemv_t)dlsym(h, "cblas_dgemv");
  if (!g || !t || !r || !v) return -2;
  p_dgemm = g; p_dtrsm = t; p_dger = r; p_dgemv = v;
  return 0;
}
int orc_using_cblas(void) { return p_dgemm != 0; }

/* cblas_dgemm(RowMajor, NoTrans, NoTrans, ...) as called at lu.C:36-41. */
void orc_dgemm(int M, int N, int K, double alpha, const double *A, int lda,
               const double *B, int ldb, double beta, double *C, int ldc) {
  if (p_dgemm) {
    p_dgemm(CB_ROW, CB_NOTRANS, CB_NOTRANS, M, N, K, alpha, A, lda, B, ldb,
            beta, C, ldc);
    return;
  }
  double *acc = (double *)malloc(sizeof(double) * (N > 0 ? N : 1));
  for (int i = 0; i < M; ++i) {
    for (int j = 0; j < N; ++j) acc[j] = 0.0;
    for (int k = 0; k < K; ++k) {
      double a = A[(long)i * lda + k];
      const double *Bk = B + (long)k * ldb;
      for (int j = 0; j < N; ++j) acc[j] += a * Bk[j];
    }
    double *Ci = C + (long)i * ldc;
    for (int j = 0; j < N; ++j)
      Ci[j] = (beta == 0.0 ? 0.0 : beta * Ci[j]) + alpha * acc[j];
  }
  free(acc);
}

/* cblas_dtrsm(RowMajor, Left, Lower, NoTrans, Unit, ...) (lu.C:187):
 * row i of the result = B_i - sum_{k<i} L[i][k] * X_k. */
void orc_dtrsm_llnu(int M, int N, const double *L, int ldl, double *B, int ldb) {
  if (p_dtrsm) {
    p_dtrsm(CB_ROW, CB_LEFT, CB_LOWER, CB_NOTRANS, CB_UNIT, M, N, 1.0, L, ldl,
            B, ldb);
    return;
  }
  for (int i = 1; i < M; ++i) {
    double *Bi = B + (long)i * ldb;
    for (int k = 0; k < i; ++k) {
      double l = L[(long)i * ldl + k];
      const double *Bk = B + (long)k * ldb;
      for (int j = 0; j < N; ++j) Bi[j] -= l * Bk[j];
    }
  }
}

/* cblas_dger(RowMajor, ...) (lu.C:247-252); the Accelerate branch at
 * lu.C:242-245 is the same loop written out. */
void orc_dger(int M, int N, double alpha, const double *x, int incx,
              const double *y, int incy, double *A, int lda) {
  if (p_dger) {
    p_dger(CB_ROW, M, N, alpha, x, incx, y, incy, A, lda);
    return;
  }
  for (int i = 0; i < M; ++i) {
    double xi = alpha * x[(long)i * incx];
    double *Ai = A + (long)i * lda;
    for (int j = 0; j < N; ++j) Ai[j] += xi * y[(long)j * incy];
  }
}

/* cblas_dgemv(RowMajor, NoTrans, ...) (benchmark.C:494-496). */
void orc_dgemv(int M, int N, double alpha, const double *A, int lda,
               const double *x, double beta, double *y) {
  if (p_dgemv) {
    p_dgemv(CB_ROW, CB_NOTRANS, M, N, alpha, A, lda, x, 1, beta, y, 1);
    return;
  }
  for (int i = 0; i < M; ++i) {
    double s = 0.0;
    for (int j = 0; j < N; ++j) s += A[(long)i * lda + j] * x[j];
    y[i] = (beta == 0.0 ? 0.0 : beta * y[i]) + alpha * s;
  }
}

/* ------------------------------------------------------------------------ */
/* pivot records -> cycles -> row moves                                      */
/* ------------------------------------------------------------------------ */

/* recordPivot (lu.C:342-352) keeps position -> row-now-living-there; a swap
 * exchanges the two entries.  announceAgglomeratedPivots (lu.C:305-339)
 * walks the map from its smallest key, following pos -> row, to emit one
 * cycle per chain.  The std::map is restated as a sorted array. */
typedef struct { int pos, row; } rec_t;

static int rec_find(rec_t *r, int cnt, int pos) {
  for (int i = 0; i < cnt; ++i)
    if (r[i].pos == pos) return i;
  return -1;
}

static int rec_insert(rec_t *r, int *cnt, int pos) {
  int i = rec_find(r, *cnt, pos);
  if (i >= 0) return i;
  r[*cnt].pos = pos;
  r[*cnt].row = pos;
  return (*cnt)++;
}

static int rec_cmp(const void *a, const void *b) {
  return ((const rec_t *)a)->pos - ((const rec_t *)b)->pos;
}

int orc_swaps_to_cycles(int count, const int *pos, const int *row,
                        int *seq_index, int *seq) {
  rec_t *r = (rec_t *)malloc(sizeof(rec_t) * (2 * count + 1));
  int cnt = 0;
  for (int k = 0; k < count; ++k) {
    if (pos[k] == row[k]) continue; /* lu.C:345 */
    int i1 = rec_insert(r, &cnt, pos[k]);
    int i2 = rec_insert(r, &cnt, row[k]);
    int t = r[i1].row;
    r[i1].row = r[i2].row;
    r[i2].row = t;
  }
  qsort(r, cnt, sizeof(rec_t), rec_cmp);
  char *used = (char *)calloc(cnt + 1, 1);
  int nseq = 0, i = 0;
  for (int s = 0; s < cnt; ++s) {
    if (used[s]) continue;
    seq_index[nseq++] = i;
    int start = r[s].pos;
    seq[i++] = start;
    used[s] = 1;
    int cur = s;
    while (r[cur].row != start) {
      int nxt = r[cur].row;
      seq[i++] = nxt;
      cur = rec_find(r, cnt, nxt);
      used[cur] = 1;
    }
  }
  seq_index[nseq] = i;
  free(used);
  free(r);
  return nseq;
}

/* sendPendingPivots (lu.C:46-163) + applySwap (lu.C:166-169): position
 * seq[i] receives the old row seq[i+1], circularly; an all-local cycle goes
 * through a temporary row (lu.C:104-115, 152-154). */
void orc_apply_cycles(int ncycles, const int *seq_index, const int *seq,
                      double *A, long lda, int c0, int c1) {
  int w = c1 - c0;
  if (w <= 0) return;
  double *tmp = (double *)malloc(sizeof(double) * w);
  for (int c = 0; c < ncycles; ++c) {
    const int *f = seq + seq_index[c], *e = seq + seq_index[c + 1];
    memcpy(tmp, A + (long)f[0] * lda + c0, sizeof(double) * w);
    for (const int *to = f; to + 1 < e; ++to)
      memcpy(A + (long)to[0] * lda + c0, A + (long)to[1] * lda + c0,
             sizeof(double) * w);
    memcpy(A + (long)e[-1] * lda + c0, tmp, sizeof(double) * w);
  }
  free(tmp);
}

/* ------------------------------------------------------------------------ */
/* factorization                                                             */
/* ------------------------------------------------------------------------ */

#define AT(i, j) A[(long)(i) * lda + (j)]

/* Optional near-tie log of orc_getrf: per global column j, |value| of the
 * chosen pivot, and the row and |value| of the best other candidate (rows
 * >= j, the same values the argmax saw; -1 / 0 if none).  A device pivot
 * that differs from the oracle's is a near-tie when it is that row and the
 * two magnitudes lie within a few ulp. */
static int *g_tie_row = NULL;
static double *g_tie_mag = NULL, *g_piv_mag = NULL;

void orc_set_tie_log(int *row2, double *mag2, double *mag1) {
  g_tie_row = row2;
  g_tie_mag = mag2;
  g_piv_mag = mag1;
}

/* argmax step with the runner-up: strict '>' keeps the first maximum
 * (MaxElm, lu.C:197-218); an equal later value becomes the runner-up */
static inline void cand2(double v, int i, double *cval, int *cloc, double *c2, int *c2loc) {
  const double a = fabs(v);
  if (a > fabs(*cval)) {
    *c2 = fabs(*cval);
    *c2loc = *cloc;
    *cval = v;
    *cloc = i;
  } else if (a > *c2) {
    *c2 = a;
    *c2loc = i;
  }
}

/* Emit one pivot batch and apply it to every column block but the active
 * panel (lu.ci:92-123 on the receiving blocks; the rhs copies were already
 * permuted inside the panel, lu.ci:290-323). */
static void flush_batch(int n, int b, int t, int count, const int *pos,
                        const int *row, double *A, long lda, int nthreads) {
  if (count == 0) return;
  int *seq_index = (int *)malloc(sizeof(int) * (count + 1));
  int *seq = (int *)malloc(sizeof(int) * (2 * count));
  int ncyc = orc_swaps_to_cycles(count, pos, row, seq_index, seq);
  int nb = n / b;
  (void)nthreads;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
  for (int y = 0; y < nb; ++y) {
    if (y == t) continue;
    orc_apply_cycles(ncyc, seq_index, seq, A, lda, y * b, (y + 1) * b);
  }
  free(seq_index);
  free(seq);
}

/* Number of block steps orc_getrf runs (0: all): the CPU baseline times the
 * first steps of a large factorization as a bounded sample of it. */
static int g_max_steps = 0;
void orc_set_max_steps(int k) { g_max_steps = k > 0 ? k : 0; }

int orc_getrf(int n, int b, int batch, double *A, long lda, double *rhs,
              int *ipiv, int nthreads) {
  if (n < 0 || b <= 0) return -1;
  if (n == 0) return 0;
  if (n % b != 0) return -1; /* benchmark.C:143-147 */
  if (batch <= 0) batch = b / 4 > 0 ? b / 4 : 1; /* benchmark.C:95-98 */
  if (nthreads < 1) nthreads = 1;
  const int nb = n / b;
  const double tiny = 100 * DBL_EPSILON; /* lu.ci:339-341 */
  int info = 0;
  int *bpos = (int *)malloc(sizeof(int) * b);
  int *brow = (int *)malloc(sizeof(int) * b);

  const int steps = g_max_steps > 0 && g_max_steps < nb ? g_max_steps : nb;
  for (int t = 0; t < steps; ++t) {
    const int r0 = t * b, r1 = r0 + b;
    /* ---- panel: factorDiagBlock (lu.ci:256-370) + factorBelowDiagBlock
     *      (lu.ci:186-251) together over rows [r0, n) ---- */
    /* findMaxElm(0, 0) on every panel block (lu.C:197-205), reduced with
     * MaxElm_max (lu.C:207-218): strict '>' from {0.0, -1}. */
    double cval = 0.0, c2 = 0.0;
    int cloc = -1, c2loc = -1;
    for (int i = r0; i < n; ++i) cand2(AT(i, r0), i, &cval, &cloc, &c2, &c2loc);
    int nrec = 0;
    for (int c = 0; c < b; ++c) {
      const int j = r0 + c;
      int p = cloc < 0 ? j : cloc; /* all-zero column: no row exchange */
      ipiv[j] = p;
      if (g_tie_row) {
        g_tie_row[j] = c2loc;
        g_tie_mag[j] = c2;
        g_piv_mag[j] = fabs(cval);
      }
      /* swapLocal / pivotRowSwap + applySwap: full panel-width rows and
       * the rhs entry (lu.C:166-177, lu.ci:290-323) */
      if (p != j) {
        for (int q = r0; q < r1; ++q) {
          double tq = AT(j, q); AT(j, q) = AT(p, q); AT(p, q) = tq;
        }
        if (rhs) { double tb = rhs[j]; rhs[j] = rhs[p]; rhs[p] = tb; }
      }
      /* recordPivot + announceAgglomeratedPivots (lu.ci:327-333) */
      bpos[nrec] = j; brow[nrec] = p; ++nrec;
      if (nrec >= batch || c == b - 1) {
        flush_batch(n, b, t, nrec, bpos, brow, A, lda, nthreads);
        nrec = 0;
      }
      const double piv = AT(j, j);
      if (fabs(piv) <= tiny && info == 0) info = j + 1;
      cval = 0.0;
      cloc = -1;
      c2 = 0.0;
      c2loc = -1;
      if (piv == 0.0) continue; /* nothing to eliminate with */
      /* computeMultipliersAndFindColMax (lu.C:259-288): division, eager
       * update of column j+1, candidate for the next column. */
      if (c < b - 1) {
        const double u1 = AT(j, j + 1);
        for (int i = j + 1; i < n; ++i) {
          AT(i, j) = AT(i, j) / piv;
          AT(i, j + 1) -= AT(i, j) * u1;
          cand2(AT(i, j + 1), i, &cval, &cloc, &c2, &c2loc);
        }
      } else {
        for (int i = j + 1; i < n; ++i) AT(i, j) = AT(i, j) / piv;
      }
      /* updateLsubBlock(activeCol, U, offset 2) (lu.C:230-254) */
      if (j + 2 < r1)
        orc_dger(n - j - 1, r1 - j - 2, -1.0, &AT(j + 1, j), (int)lda,
                 &AT(j, j + 2), 1, &AT(j + 1, j + 2), (int)lda);
    }
    /* ---- U row: computeU (lu.C:180-189) for every block (t, y>t) ---- */
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int y = t + 1; y < nb; ++y)
      orc_dtrsm_llnu(b, b, &AT(r0, r0), (int)lda, &AT(r0, y * b), (int)lda);
    /* ---- trailing: updateMatrix (lu.C:26-43) for every block (x>t, y>t) */
    const int m = nb - t - 1;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (long xy = 0; xy < (long)m * m; ++xy) {
      int x = t + 1 + (int)(xy / m), y = t + 1 + (int)(xy % m);
      orc_dgemm(b, b, b, -1.0, &AT(x * b, r0), (int)lda, &AT(r0, y * b),
                (int)lda, 1.0, &AT(x * b, y * b), (int)lda);
    }
  }
  free(bpos);
  free(brow);
  return info;
}

int orc_getf2(int n, double *A, long lda, int *ipiv) {
  int info = 0;
  for (int j = 0; j < n; ++j) {
    int p = j;
    double best = 0.0;
    int found = 0;
    for (int i = j; i < n; ++i)
      if (fabs(AT(i, j)) > best) { best = fabs(AT(i, j)); p = i; found = 1; }
    if (!found) p = j;
    ipiv[j] = p;
    if (p != j)
      for (int q = 0; q < n; ++q) {
        double tq = AT(j, q); AT(j, q) = AT(p, q); AT(p, q) = tq;
      }
    const double piv = AT(j, j);
    if (fabs(piv) <= 100 * DBL_EPSILON && info == 0) info = j + 1;
    if (piv == 0.0) continue;
    for (int i = j + 1; i < n; ++i) {
      AT(i, j) = AT(i, j) / piv;
      for (int q = j + 1; q < n; ++q) AT(i, q) -= AT(i, j) * AT(j, q);
    }
  }
  return info;
}

void orc_apply_ipiv(int n, const int *ipiv, double *v) {
  for (int j = 0; j < n; ++j) {
    int p = ipiv[j];
    if (p != j) { double t = v[j]; v[j] = v[p]; v[p] = t; }
  }
}

/* forwardSolve / backwardSolve / offDiagSolve (lu.ci:387-458).  The
 * off-diagonal contributions are row-wise dot products per block
 * (lu.ci:443-447) summed over blocks; the reference sums them in run-time
 * order (sum_double), here in ascending block order. */
void orc_solve(int n, int b, const double *LU, long lda, double *rhs) {
  const double *A = LU;
  int nb = n / b;
  double *pre = (double *)malloc(sizeof(double) * b);
  double *xv = (double *)malloc(sizeof(double) * b);
  for (int x = 0; x < nb; ++x) {
    double *bx = rhs + (long)x * b;
    if (x > 0) {
      for (int i = 0; i < b; ++i) pre[i] = 0.0;
      for (int y = 0; y < x; ++y) {
        for (int i = 0; i < b; ++i) {
          xv[i] = 0.0;
          for (int j = 0; j < b; ++j) xv[i] += AT(x * b + i, y * b + j) * rhs[(long)y * b + j];
        }
        for (int i = 0; i < b; ++i) pre[i] += xv[i];
      }
      for (int i = 0; i < b; ++i) bx[i] -= pre[i];
    }
    for (int i = 0; i < b; ++i)
      for (int j = 0; j < i; ++j) bx[i] -= AT(x * b + i, x * b + j) * bx[j];
  }
  for (int x = nb - 1; x >= 0; --x) {
    double *bx = rhs + (long)x * b;
    if (x < nb - 1) {
      for (int i = 0; i < b; ++i) pre[i] = 0.0;
      for (int y = x + 1; y < nb; ++y) {
        for (int i = 0; i < b; ++i) {
          xv[i] = 0.0;
          for (int j = 0; j < b; ++j) xv[i] += AT(x * b + i, y * b + j) * rhs[(long)y * b + j];
        }
        for (int i = 0; i < b; ++i) pre[i] += xv[i];
      }
      for (int i = 0; i < b; ++i) bx[i] -= pre[i];
    }
    for (int i = b - 1; i >= 0; --i) {
      for (int j = i + 1; j < b; ++j) bx[i] -= AT(x * b + i, x * b + j) * bx[j];
      bx[i] /= AT(x * b + i, x * b + i);
    }
  }
  free(pre);
  free(xv);
}

/* calcScaledResidual (benchmark.C:385-400) over the maxima gathered by
 * recvXvec/sumBvec/calcResiduals (benchmark.C:485-573); ||A|| there is the
 * largest |a_ij| (infNorm over the block's entries, benchmark.C:476-483). */
double orc_residual(int n, const double *A, long lda, const double *x,
                    const double *b, double *out4, double *ns) {
  double amax = 0, bmax = 0, xmax = 0, rmax = 0, arow = 0;
  for (int i = 0; i < n; ++i) {
    double s = 0.0, rs = 0.0;
    for (int j = 0; j < n; ++j) {
      double a = AT(i, j);
      s += a * x[j];
      rs += fabs(a);
      if (fabs(a) > amax) amax = fabs(a);
    }
    double r = fabs(b[i] - s);
    if (r > rmax) rmax = r;
    if (fabs(b[i]) > bmax) bmax = fabs(b[i]);
    if (fabs(x[i]) > xmax) xmax = fabs(x[i]);
    if (rs > arow) arow = rs;
  }
  if (out4) { out4[0] = amax; out4[1] = bmax; out4[2] = xmax; out4[3] = rmax; }
  if (ns) *ns = rmax / (arow * xmax * n * DBL_EPSILON);
  return rmax / ((amax * xmax + bmax) * n * DBL_EPSILON);
}

/* The library LU the CPU baseline also reports (not the reference's
 * algorithm): the vendor LAPACK's dgetrf + dgetrs on all given threads.
 * The row-major A is A^T in column-major storage, so dgetrf factors A^T
 * in place and dgetrs with trans = 'T' solves A x = b from those factors
 * (no transposed copy).  Returns LAPACK's info (or -100 if the library or
 * a symbol is missing); *seconds = factor + solve wall time. */
typedef void (*dgetrf_t)(const int *, const int *, double *, const int *, int *, int *);
typedef void (*dgetrs_t)(const char *, const int *, const int *, const double *, const int *,
                         const int *, double *, const int *, int *, size_t);
int orc_lapack_gesv(const char *libpath, int n, double *A, long lda, double *b, int *ipiv,
                    int nthreads, double *seconds) {
  void *h = dlopen(libpath, RTLD_NOW | RTLD_LOCAL);
  if (!h) return -100;
  int (*set_layer)(int) = (int (*)(int))dlsym(h, "MKL_Set_Threading_Layer");
  if (set_layer) set_layer(3 /* MKL_THREADING_GNU: the oracle's own OpenMP runtime */);
  void (*set_threads)(int) = (void (*)(int))dlsym(h, "MKL_Set_Num_Threads");
  if (set_threads) set_threads(nthreads);
  dgetrf_t getrf = (dgetrf_t)dlsym(h, "dgetrf_");
  dgetrs_t getrs = (dgetrs_t)dlsym(h, "dgetrs_");
  if (!getrf || !getrs) return -100;
  const int ni = n, ld = (int)lda, one = 1;
  int info = 0;
  const double t0 = orc_wtime();
  getrf(&ni, &ni, A, &ld, ipiv, &info);
  if (info == 0) getrs("T", &ni, &one, A, &ld, ipiv, b, &ni, &info, 1);
  *seconds = orc_wtime() - t0;
  return info;
}

double orc_wtime(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}
