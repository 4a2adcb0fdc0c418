// trsm.hip -- triangular solves on a block of right-hand-side columns.
//
// trsm_llnu replaces the reference's U-row solve
//   cblas_dtrsm(RowMajor, Left, Lower, NoTrans, Unit, b, b, 1.0, L, b, A, b)
// (computeU, lu.C:180-189), applied to the whole b x N_t row panel of a step
// in one launch.  trsm_lunn is the diagonal-block step of the backward solve
// (backwardSolve, lu.ci:408-436); llnu also serves the forward solve
// (forwardSolve, lu.ci:387-407).  The transposed forms serve A^T x = b.
//
// The right-hand sides are solved in 32-column slabs held in LDS, 16 rows
// of the triangle at a time (see trsm_kernel).
//
// gemv_sub is the off-diagonal part of the block solve (offDiagSolve,
// lu.ci:439-456): y -= A x for a tall block column, 16 lanes per row.
#include "common.h"
#include "kernels.h"

namespace clu {

// op(F)[i][k] = TR ? F[k][i] : F[i][k]; OPLOW: op(F) is lower triangular
// (forward substitution), else upper (backward).  UNIT: unit diagonal.
//
// One 256-thread workgroup owns NC = 32 right-hand-side columns, kept in
// LDS for the whole solve (M x 32 values, M <= 256).  The rows are taken in
// blocks of 16 in solve order; for each block the 16 rows of op(F) are
// staged in LDS, the contributions of every row solved so far are
// subtracted (8 row groups x 32 columns of threads, 2 rows each), and the
// 16 x 16 diagonal block is solved by one thread per column.
template <typename T, bool OPLOW, bool TR, bool UNIT>
__global__ __launch_bounds__(256) void trsm_kernel(int M, int N, T alpha, const T *__restrict__ F,
                                                   long ldf, T *__restrict__ B, long ldb,
                                                   long fstride, long bstride) {
  constexpr int NC = 32, BS = 16;
  // blockIdx.y selects one of a batch of independent solves
  F += blockIdx.y * fstride;
  B += blockIdx.y * bstride;
  extern __shared__ unsigned char smem_raw[];
  T *X = reinterpret_cast<T *>(smem_raw);  // [M][NC]
  T *Fb = X + (size_t)M * NC;              // [BS][M]
  const int t = threadIdx.x;
  const int c = t % NC, rg = t / NC;  // rg in 0..7: rows 2*rg, 2*rg+1 of a block
  const int col0 = blockIdx.x * NC;
  for (int e = t; e < M * NC; e += 256) {
    const int i = e / NC, cc = e % NC, gc = col0 + cc;
    X[e] = gc < N ? alpha * B[(long)i * ldb + gc] : (T)0;
  }
  const int nblk = (M + BS - 1) / BS;
  for (int sidx = 0; sidx < nblk; ++sidx) {
    const int blk = OPLOW ? sidx : nblk - 1 - sidx;
    const int r0 = blk * BS;
    const int rows = M - r0 < BS ? M - r0 : BS;
    // columns of op(F) this block needs: everything solved before plus
    // the diagonal block itself
    const int k_lo = OPLOW ? 0 : r0;
    const int k_hi = OPLOW ? r0 + rows : M;
    const int kn = k_hi - k_lo;
    __syncthreads();  // X loads / previous block's solve done before Fb is rewritten
    if (!TR) {
      for (int e = t; e < BS * kn; e += 256) {
        const int rr = e / kn, k = k_lo + e % kn;
        Fb[rr * M + k] = rr < rows ? F[(long)(r0 + rr) * ldf + k] : (T)0;
      }
    } else {
      for (int e = t; e < BS * kn; e += 256) {
        const int k = k_lo + e / BS, rr = e % BS;
        Fb[rr * M + k] = rr < rows ? F[(long)k * ldf + r0 + rr] : (T)0;
      }
    }
    __syncthreads();
    // contributions of the rows solved so far
    const int u_lo = OPLOW ? 0 : r0 + rows;
    const int u_hi = OPLOW ? r0 : M;
    const int ra = 2 * rg, rb = 2 * rg + 1;
    T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int k = u_lo;
    for (; k + 1 < u_hi; k += 2) {
      const T x0 = X[k * NC + c], x1 = X[(k + 1) * NC + c];
      s0 += Fb[ra * M + k] * x0;
      s1 += Fb[rb * M + k] * x0;
      s2 += Fb[ra * M + k + 1] * x1;
      s3 += Fb[rb * M + k + 1] * x1;
    }
    if (k < u_hi) {
      const T x0 = X[k * NC + c];
      s0 += Fb[ra * M + k] * x0;
      s1 += Fb[rb * M + k] * x0;
    }
    if (ra < rows) X[(r0 + ra) * NC + c] -= s0 + s2;
    if (rb < rows) X[(r0 + rb) * NC + c] -= s1 + s3;
    __syncthreads();
    // the diagonal block, one thread per column, its 16 values in registers
    // (read back from LDS after each row, the substitution was a chain of
    // 120 dependent LDS round trips per block: 16 us for a 32 x 32 solve)
    if (t < NC) {
      T xr[BS];
#pragma clang loop unroll(full)
      for (int i = 0; i < BS; ++i) xr[i] = i < rows ? X[(r0 + i) * NC + t] : (T)0;
#pragma clang loop unroll(full)
      for (int ii = 0; ii < BS; ++ii) {
        const int i = OPLOW ? ii : BS - 1 - ii;
        if (i < rows) {
          T x = xr[i];
          if (OPLOW) {
#pragma clang loop unroll(full)
            for (int kk = 0; kk < i; ++kk) x -= Fb[i * M + r0 + kk] * xr[kk];
          } else {
#pragma clang loop unroll(full)
            for (int kk = i + 1; kk < BS; ++kk)
              if (kk < rows) x -= Fb[i * M + r0 + kk] * xr[kk];
          }
          if (!UNIT) x = x / Fb[i * M + r0 + i];
          xr[i] = x;
        }
      }
#pragma clang loop unroll(full)
      for (int i = 0; i < BS; ++i)
        if (i < rows) X[(r0 + i) * NC + t] = xr[i];
    }
  }
  __syncthreads();
  for (int e = t; e < M * NC; e += 256) {
    const int i = e / NC, cc = e % NC, gc = col0 + cc;
    if (gc < N) B[(long)i * ldb + gc] = X[e];
  }
}

// One right-hand side (the reference's solve has exactly one, lu.ci:387-458):
// 256 threads as 16 rows x 16 lanes.  Per 16-row block, each row's dot
// product with the solved part of x is split over its 16 lanes and reduced
// by shuffles; the 16 x 16 diagonal block is then solved by wave 0 with one
// shuffle broadcast per row.
template <typename T, bool OPLOW, bool TR, bool UNIT>
__global__ __launch_bounds__(256) void trsv_kernel(int M, T alpha, const T *__restrict__ F,
                                                   long ldf, T *__restrict__ B, long ldb) {
  __shared__ T x[256];
  __shared__ T rs[16];  // row sums of the current block
  const int t = threadIdx.x, r = t >> 4, kl = t & 15;
  auto f = [&](int i, int k) -> T { return TR ? F[(long)k * ldf + i] : F[(long)i * ldf + k]; };
  for (int i = t; i < M; i += 256) x[i] = alpha * B[(long)i * ldb];
  __syncthreads();
  const int nblk = (M + 15) / 16;
  for (int sidx = 0; sidx < nblk; ++sidx) {
    const int blk = OPLOW ? sidx : nblk - 1 - sidx;
    const int r0 = blk * 16;
    const int rows = M - r0 < 16 ? M - r0 : 16;
    const int u_lo = OPLOW ? 0 : r0 + rows;
    const int u_hi = OPLOW ? r0 : M;
    T acc0 = 0, acc1 = 0;
    if (r < rows) {
      int k = u_lo + kl;
      for (; k + 16 < u_hi; k += 32) {
        acc0 += f(r0 + r, k) * x[k];
        acc1 += f(r0 + r, k + 16) * x[k + 16];
      }
      if (k < u_hi) acc0 += f(r0 + r, k) * x[k];
    }
    T acc = acc0 + acc1;
#pragma unroll
    for (int off = 8; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 16);
    if (kl == 0 && r < rows) rs[r] = acc;
    __syncthreads();
    if (t < 64) {
      // wave 0: lane i < rows owns row r0 + i of the diagonal block
      const int i = t;
      T xi = (i < rows) ? x[r0 + i] - rs[i] : (T)0;
      T fi[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) fi[j] = (i < rows && j < rows) ? f(r0 + i, r0 + j) : (T)0;
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) {
        const int j = OPLOW ? jj : 15 - jj;
        if (j < rows) {
          if (!UNIT && i == j) xi = xi / fi[j];
          const T xj = __shfl(xi, j, 64);
          const bool later = OPLOW ? (i > j) : (i < j);
          if (later && i < rows) xi -= fi[j] * xj;
        }
      }
      if (i < rows) x[r0 + i] = xi;
    }
    __syncthreads();
  }
  for (int i = t; i < M; i += 256) B[(long)i * ldb] = x[i];
}

template <typename T, bool OPLOW, bool TR, bool UNIT>
static int trsm_launch(int M, int N, T alpha, const T *F, long ldf, T *B, long ldb,
                       hipStream_t s) {
  if (M <= 0 || N <= 0) return CLU_OK;
  if (M > 256) return -1;
  if (N == 1) {
    hipLaunchKernelGGL((trsv_kernel<T, OPLOW, TR, UNIT>), dim3(1), dim3(256), 0, s, M, alpha, F,
                       ldf, B, ldb);
    CLU_LAUNCHED("trsv_kernel");
    return CLU_OK;
  }
  const size_t shm = (size_t)M * 32 * sizeof(T) + (size_t)16 * M * sizeof(T);
  hipLaunchKernelGGL((trsm_kernel<T, OPLOW, TR, UNIT>), dim3((N + 31) / 32), dim3(256), shm, s,
                     M, N, alpha, F, ldf, B, ldb, 0L, 0L);
  CLU_LAUNCHED("trsm_kernel");
  return CLU_OK;
}

// ---------------------------------------------------------------------------
// Full inverse of a triangle by recursive doubling on the matrix cores.
// The 16 x 16 diagonal blocks are inverted first, one wave each (lane j
// forms column j of the block inverse in registers, reading the triangle
// from LDS as broadcasts).  Then, for s = 32, 64, ..., pairs of adjacent
// inverted diagonal blocks of size s are joined into one of size 2s:
//   lower:  X21 = -X22 (L21 X11),   upper:  X12 = -X11 (U12 X22),
// two strided-batched GEMMs per level over all pairs at once.
// ---------------------------------------------------------------------------
constexpr int TIB = 16;  // base block

// Block k of every triangle (grid (nblocks, ntri)): invert the diagonal
// block, write it into X and zero the part of its block row of X that lies
// on the zero side of the triangle.
template <typename T, bool UPPER, bool UNIT>
__global__ __launch_bounds__(64) void trinv_base_kernel(int n, const T *__restrict__ A, long lda,
                                                        long sa, T *__restrict__ X, long ldx,
                                                        long sx) {
  __shared__ T Ls[TIB][TIB + 1];
  A += blockIdx.y * sa;
  X += blockIdx.y * sx;
  const int k0 = blockIdx.x * TIB;
  const int bs = n - k0 < TIB ? n - k0 : TIB;
  const int lane = threadIdx.x;
  // all 16 loads of a lane in flight before the first LDS store
  constexpr int RPP = 64 / TIB;  // rows per pass of the wave
  T v[TIB / RPP];
#pragma unroll
  for (int r = 0; r < TIB / RPP; ++r) {
    const int i = r * RPP + lane / TIB, j = lane % TIB;
    v[r] = (i < bs && j < bs) ? A[(long)(k0 + i) * lda + k0 + j] : (T)0;
  }
#pragma unroll
  for (int r = 0; r < TIB / RPP; ++r) Ls[r * RPP + lane / TIB][lane % TIB] = v[r];
  __syncthreads();
  const int j = lane;  // lanes 0..TIB-1: column j of the block inverse
  T x[TIB];
  if (!UPPER) {
#pragma unroll
    for (int i = 0; i < TIB; ++i) {
      T acc = (T)0;
#pragma unroll
      for (int q = 0; q < i; ++q) acc += Ls[i][q] * x[q];
      x[i] = (i == j) ? (T)1 : -acc;
      if (!UNIT) x[i] = x[i] / Ls[i][i];
    }
  } else {
#pragma unroll
    for (int i = TIB - 1; i >= 0; --i) {
      T acc = (T)0;
#pragma unroll
      for (int q = i + 1; q < TIB; ++q) acc += Ls[i][q] * x[q];
      // rows and columns past the block's end are zero: x there stays 0
      T v = (i == j) ? (T)1 : -acc;
      if (!UNIT) v = (i < bs) ? v / Ls[i][i] : (T)0;
      x[i] = v;
    }
  }
  if (j < bs) {
#pragma unroll
    for (int i = 0; i < TIB; ++i)
      if (i < bs) X[(long)(k0 + i) * ldx + k0 + j] = x[i];
  }
  // zero side of this block row: columns right of the block (lower) or
  // left of it (upper)
  const int c_lo = UPPER ? 0 : k0 + bs, c_hi = UPPER ? k0 : n;
  for (int i = 0; i < bs; ++i)
    for (int c = c_lo + lane; c < c_hi; c += 64) X[(long)(k0 + i) * ldx + c] = (T)0;
}

template <typename T>
int gemm_batched(int M, int N, int K, T alpha, const T *A, long lda, const T *B, long ldb,
                 T beta, T *C, long ldc, int ny, int nz, long a_y, long a_z, long b_y, long b_z,
                 long c_y, long c_z, hipStream_t s);

template <typename T>
int tri_inv(bool upper, bool unit, int n, int ntri, const T *A, long lda, long sa, T *X, long ldx,
            long sx, T *tbuf, hipStream_t s) {
  if (n <= 0 || ntri <= 0) return CLU_OK;
  if (ntri > 65535) return -1;
  const dim3 g((n + TIB - 1) / TIB, ntri);
#define CLU_TRINV(U, UN) \
  if (upper == U && unit == UN) \
    hipLaunchKernelGGL((trinv_base_kernel<T, U, UN>), g, dim3(64), 0, s, n, A, lda, sa, X, ldx, sx);
  CLU_TRINV(false, true)
  CLU_TRINV(false, false)
  CLU_TRINV(true, true)
  CLU_TRINV(true, false)
#undef CLU_TRINV
  CLU_LAUNCHED("trinv_base_kernel");
  // tbuf holds one (at most s x s) product per pair: n x n per triangle
  const long tstride = (long)n * n;
  for (int sz = TIB; sz < n; sz *= 2) {
    const int full = n / (2 * sz), rem = n - full * 2 * sz;
    // the full pairs (grid rows), then the partial last pair if it has a right part
    for (int part = 0; part < 2; ++part) {
      const int np = part == 0 ? full : (rem > sz ? 1 : 0);
      if (np == 0) continue;
      const long o = part == 0 ? 0 : (long)full * 2 * sz;  // first pair's diagonal offset
      const int s2 = part == 0 ? sz : rem - sz;
      const long pa = 2L * sz * (lda + 1), px = 2L * sz * (ldx + 1), pt = (long)sz * sz;
      const T *Ad = A + o * (lda + 1);
      T *Xd = X + o * (ldx + 1);
      T *Tb = tbuf + (part == 0 ? 0 : (long)full * sz * sz);
      int rc;
      if (!upper) {
        // T = L21 (s2 x sz) X11 (sz x sz);  X21 = -X22 (s2 x s2) T
        rc = gemm_batched<T>(s2, sz, sz, (T)1, Ad + (long)sz * lda, lda, Xd, ldx, (T)0, Tb, sz,
                             np, ntri, pa, sa, px, sx, pt, tstride, s);
        if (rc) return rc;
        rc = gemm_batched<T>(s2, sz, s2, (T)-1, Xd + (long)sz * ldx + sz, ldx, Tb, sz, (T)0,
                             Xd + (long)sz * ldx, ldx, np, ntri, px, sx, pt, tstride, px, sx, s);
      } else {
        // T = U12 (sz x s2) X22 (s2 x s2);  X12 = -X11 (sz x sz) T
        rc = gemm_batched<T>(sz, s2, s2, (T)1, Ad + sz, lda, Xd + (long)sz * ldx + sz, ldx, (T)0,
                             Tb, s2, np, ntri, pa, sa, px, sx, pt, tstride, s);
        if (rc) return rc;
        rc = gemm_batched<T>(sz, s2, sz, (T)-1, Xd, ldx, Tb, s2, (T)0, Xd + sz, ldx, np, ntri,
                             px, sx, pt, tstride, px, sx, s);
      }
      if (rc) return rc;
    }
  }
  return CLU_OK;
}
template int tri_inv<double>(bool, bool, int, int, const double *, long, long, double *, long,
                             long, double *, hipStream_t);
template int tri_inv<float>(bool, bool, int, int, const float *, long, long, float *, long, long,
                            float *, hipStream_t);

// In-place transpose of nblk contiguous S x S blocks (S a multiple of 32):
// each workgroup swaps one pair of 32 x 32 tiles (i, j) and (j, i), i <= j,
// through LDS.  The transposed solve keeps its diagonal-block inverses
// transposed, so that applying them reads rows.
template <typename T>
__global__ __launch_bounds__(256) void blk_transpose_kernel(int S, T *__restrict__ G) {
  const int ti = blockIdx.x, tj = blockIdx.y;
  if (ti > tj) return;
  __shared__ T a[32][33], b[32][33];
  T *B = G + (size_t)blockIdx.z * S * S;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    a[r][tx] = B[(long)(ti * 32 + r) * S + tj * 32 + tx];
    b[r][tx] = B[(long)(tj * 32 + r) * S + ti * 32 + tx];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    B[(long)(ti * 32 + r) * S + tj * 32 + tx] = b[tx][r];
    B[(long)(tj * 32 + r) * S + ti * 32 + tx] = a[tx][r];
  }
}

template <typename T>
int blk_transpose(int S, int nblk, T *G, hipStream_t s) {
  if (S <= 0 || nblk <= 0) return CLU_OK;
  if (S % 32 != 0 || nblk > 65535) return -1;
  hipLaunchKernelGGL((blk_transpose_kernel<T>), dim3(S / 32, S / 32, nblk), dim3(256), 0, s, S,
                     G);
  CLU_LAUNCHED("blk_transpose_kernel");
  return CLU_OK;
}
template int blk_transpose<double>(int, int, double *, hipStream_t);
template int blk_transpose<float>(int, int, float *, hipStream_t);

// y = op(G) x for one m x m diagonal-block inverse G (m <= 256, ld ldg).
// No transpose: one wave per row (all of a lane's loads issued before the
// shuffle reduction); transpose: one lane per output, rows split over the 4
// waves of a workgroup.
template <typename T, bool TR>
__global__ __launch_bounds__(256) void blk_gemv_kernel(int m, const T *__restrict__ G, long ldg,
                                                       const T *__restrict__ x,
                                                       T *__restrict__ y) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (!TR) {
    const int i = blockIdx.x * 4 + wave;
    if (i >= m) return;
    double acc = 0.0;
    // 4 values per lane per pass, all loads of a pass before its products
    for (int c0 = 0; c0 < m; c0 += 256) {
      T g[4], xv[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int c = c0 + lane + 64 * k;
        g[k] = c < m ? G[(long)i * ldg + c] : (T)0;
        xv[k] = c < m ? x[c] : (T)0;
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) acc += (double)g[k] * (double)xv[k];
    }
    acc = wave_sum(acc);
    if (lane == 0) y[i] = (T)acc;
  } else {
    __shared__ double part[4][64];
    const int j = blockIdx.x * 64 + lane;
    double acc = 0.0;
    if (j < m)
      for (int k = wave; k < m; k += 4) acc += (double)G[(long)k * ldg + j] * (double)x[k];
    part[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && j < m)
      y[j] = (T)((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]));
  }
}

template <typename T>
int blk_gemv(bool trans, int m, const T *G, long ldg, const T *x, T *y, hipStream_t s) {
  if (m <= 0) return CLU_OK;
  if (m > 1024) return -1;
  if (!trans)
    hipLaunchKernelGGL((blk_gemv_kernel<T, false>), dim3((m + 3) / 4), dim3(256), 0, s, m, G, ldg,
                       x, y);
  else
    hipLaunchKernelGGL((blk_gemv_kernel<T, true>), dim3((m + 63) / 64), dim3(256), 0, s, m, G,
                       ldg, x, y);
  CLU_LAUNCHED("blk_gemv_kernel");
  return CLU_OK;
}
template int blk_gemv<double>(bool, int, const double *, long, const double *, double *,
                              hipStream_t);
template int blk_gemv<float>(bool, int, const float *, long, const float *, float *,
                             hipStream_t);

template <typename T>
int trsm_small(bool upper, bool trans, bool unit, int M, int N, T alpha, const T *F, long ldf,
               T *B, long ldb, hipStream_t s) {
  const bool oplow = (!upper) != trans;
#define CLU_TRSM_CASE(L, TR, U)                                                         \
  if (oplow == L && trans == TR && unit == U)                                           \
    return trsm_launch<T, L, TR, U>(M, N, alpha, F, ldf, B, ldb, s);
  CLU_TRSM_CASE(true, false, true)
  CLU_TRSM_CASE(true, false, false)
  CLU_TRSM_CASE(true, true, true)
  CLU_TRSM_CASE(true, true, false)
  CLU_TRSM_CASE(false, false, true)
  CLU_TRSM_CASE(false, false, false)
  CLU_TRSM_CASE(false, true, true)
  CLU_TRSM_CASE(false, true, false)
#undef CLU_TRSM_CASE
  return -1;
}
template int trsm_small<double>(bool, bool, bool, int, int, double, const double *, long,
                                double *, long, hipStream_t);
template int trsm_small<float>(bool, bool, bool, int, int, float, const float *, long, float *,
                               long, hipStream_t);

template <typename T>
int trsm_llnu(int M, int N, const T *L, long ldl, T *B, long ldb, hipStream_t s) {
  return trsm_launch<T, true, false, true>(M, N, (T)1, L, ldl, B, ldb, s);
}
template <typename T>
int trsm_lunn(int M, int N, const T *U, long ldu, T *B, long ldb, hipStream_t s) {
  return trsm_launch<T, false, false, false>(M, N, (T)1, U, ldu, B, ldb, s);
}
template <typename T>
int trsm_lltu(int M, int N, const T *L, long ldl, T *B, long ldb, hipStream_t s) {
  return trsm_launch<T, false, true, true>(M, N, (T)1, L, ldl, B, ldb, s);
}
template <typename T>
int trsm_lutn(int M, int N, const T *U, long ldu, T *B, long ldb, hipStream_t s) {
  return trsm_launch<T, true, true, false>(M, N, (T)1, U, ldu, B, ldb, s);
}

#define CLU_INST_TRSM(T)                                                          \
  template int trsm_llnu<T>(int, int, const T *, long, T *, long, hipStream_t);  \
  template int trsm_lunn<T>(int, int, const T *, long, T *, long, hipStream_t);  \
  template int trsm_lltu<T>(int, int, const T *, long, T *, long, hipStream_t);  \
  template int trsm_lutn<T>(int, int, const T *, long, T *, long, hipStream_t);
CLU_INST_TRSM(double)
CLU_INST_TRSM(float)

// y[i] -= sum_k A[i][k] x[k]; 16 lanes per row, 16 rows per 256-thread block
template <typename T>
__global__ __launch_bounds__(256) void gemv_sub_n_kernel(int M, int K, const T *__restrict__ A,
                                                         long lda, const T *__restrict__ x,
                                                         T *__restrict__ y) {
  const int row = blockIdx.x * 16 + (threadIdx.x >> 4);
  const int sub = threadIdx.x & 15;
  double s = 0.0;
  if (row < M)
    for (int k = sub; k < K; k += 16) s += (double)A[(long)row * lda + k] * (double)x[k];
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) s += __shfl_xor(s, off, 16);
  if (row < M && sub == 0) y[row] = (T)((double)y[row] - s);
}

// The same with 16-byte loads (rows 16-byte aligned, K a multiple of the
// vector width): each 16-lane row group reads 256 contiguous bytes per step.
template <typename T>
__global__ __launch_bounds__(256) void gemv_sub_nv_kernel(int M, int K, const T *__restrict__ A,
                                                          long lda, const T *__restrict__ x,
                                                          T *__restrict__ y) {
  constexpr int V = 16 / sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(V)));
  const int row = blockIdx.x * 16 + (threadIdx.x >> 4);
  const int sub = threadIdx.x & 15;
  double s = 0.0;
  if (row < M) {
    // two loads in flight and two independent sums per lane
    const vec_t *a = reinterpret_cast<const vec_t *>(A + (long)row * lda);
    const vec_t *xv = reinterpret_cast<const vec_t *>(x);
    const int KV = K / V;
    double s1 = 0.0;
    int k = sub;
    for (; k + 16 < KV; k += 32) {
      const vec_t a0 = a[k], a1 = a[k + 16], x0 = xv[k], x1 = xv[k + 16];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        s += (double)a0[v] * (double)x0[v];
        s1 += (double)a1[v] * (double)x1[v];
      }
    }
    if (k < KV) {
      const vec_t a0 = a[k], x0 = xv[k];
#pragma unroll
      for (int v = 0; v < V; ++v) s += (double)a0[v] * (double)x0[v];
    }
    s += s1;
  }
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) s += __shfl_xor(s, off, 16);
  if (row < M && sub == 0) y[row] = (T)((double)y[row] - s);
}

// y[j] -= sum_k A[k][j] x[k] (j < M, k < K).  128 columns per workgroup,
// two per lane (one 2-element load per row); the 16 waves take every 16th
// row, 8 rows in flight per lane, and their partial sums are added in a
// fixed order through LDS.  (One lane per column with one 4-byte load in
// flight read a 1024-row strip of FP32 factors at about 1 TB/s.)
template <typename T, bool VEC>
__global__ __launch_bounds__(1024) void gemv_sub_t_kernel(int M, int K, const T *__restrict__ A,
                                                          long lda, const T *__restrict__ x,
                                                          T *__restrict__ y) {
  typedef T v2_t __attribute__((ext_vector_type(2)));
  __shared__ double part[16][2][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c0 = blockIdx.x * 128 + 2 * lane;
  double s0[8], s1[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) s0[u] = s1[u] = 0.0;
  if (c0 < M) {
    int k = wave;
    if (VEC) {
      for (; k + 112 < K; k += 128) {
        v2_t a[8];
        T xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          a[u] = *reinterpret_cast<const v2_t *>(A + (long)(k + 16 * u) * lda + c0);
          xv[u] = x[k + 16 * u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          s0[u] += (double)a[u][0] * (double)xv[u];
          s1[u] += (double)a[u][1] * (double)xv[u];
        }
      }
      for (; k < K; k += 16) {
        const v2_t a0 = *reinterpret_cast<const v2_t *>(A + (long)k * lda + c0);
        s0[0] += (double)a0[0] * (double)x[k];
        s1[0] += (double)a0[1] * (double)x[k];
      }
    } else {
      for (; k < K; k += 16) {
        const double xk = (double)x[k];
        s0[0] += (double)A[(long)k * lda + c0] * xk;
        if (c0 + 1 < M) s1[0] += (double)A[(long)k * lda + c0 + 1] * xk;
      }
    }
  }
  part[wave][0][lane] = ((s0[0] + s0[1]) + (s0[2] + s0[3])) + ((s0[4] + s0[5]) + (s0[6] + s0[7]));
  part[wave][1][lane] = ((s1[0] + s1[1]) + (s1[2] + s1[3])) + ((s1[4] + s1[5]) + (s1[6] + s1[7]));
  __syncthreads();
  if (wave < 2) {
    const int c = c0 + wave;
    if (c < M) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < 16; ++w) t += part[w][wave][lane];
      y[c] = (T)((double)y[c] - t);
    }
  }
}

template <typename T>
int gemv_sub(bool trans, int M, int K, const T *A, long lda, const T *x, T *y, hipStream_t s) {
  if (M <= 0 || K <= 0) return CLU_OK;
  constexpr int V = 16 / sizeof(T);
  const bool aligned = lda % V == 0 && K % V == 0 && ((uintptr_t)A & 15) == 0 &&
                       ((uintptr_t)x & 15) == 0;
  if (!trans && aligned) {
    hipLaunchKernelGGL((gemv_sub_nv_kernel<T>), dim3((M + 15) / 16), dim3(256), 0, s, M, K, A,
                       lda, x, y);
  } else if (!trans) {
    hipLaunchKernelGGL((gemv_sub_n_kernel<T>), dim3((M + 15) / 16), dim3(256), 0, s, M, K, A,
                       lda, x, y);
  } else if (lda % 2 == 0 && M % 2 == 0 && ((uintptr_t)A % (2 * sizeof(T))) == 0) {
    hipLaunchKernelGGL((gemv_sub_t_kernel<T, true>), dim3((M + 127) / 128), dim3(1024), 0, s, M,
                       K, A, lda, x, y);
  } else {
    hipLaunchKernelGGL((gemv_sub_t_kernel<T, false>), dim3((M + 127) / 128), dim3(1024), 0, s, M,
                       K, A, lda, x, y);
  }
  CLU_LAUNCHED("gemv_sub_kernel");
  return CLU_OK;
}
template int gemv_sub<double>(bool, int, int, const double *, long, const double *, double *,
                              hipStream_t);
template int gemv_sub<float>(bool, int, int, const float *, long, const float *, float *,
                             hipStream_t);

// ---------------------------------------------------------------------------
// trsv_tiles: one triangle of the block solve of one right-hand side in one
// launch (forwardSolve / backwardSolve with their offDiagSolve updates,
// lu.ci:387-458, for a single vector).
//
// The launch-per-block form (blk_gemv, then gemv_sub, for every diagonal
// block) spends most of a large solve in the gaps and ramps of its 2 n / S
// dependent launches.  Here workgroup q owns a tile of TR = 16, 32 or 64
// rows (WR = TR / 4 per wave; S / TR tiles per diagonal block of S rows),
// counted from the top for the forward
// (unit lower) triangle and from the bottom for the backward (upper) one.
// It streams its rows of every off-diagonal block column in solve order, each
// as soon as the solution of that block is published, keeping FP64 partial
// sums per lane; then it publishes t = rhs - sum for its rows, reads the t of
// its whole block once every tile of the block has published, and computes
// its rows of x = inv(D) t from the stored inverse of the diagonal block D.
//
// Published values travel as 64-bit words, (sequence number << 32) | 32 bits
// of payload (two words for a double), stored and loaded with relaxed
// agent-scope atomics: a reader accepts a block only when every word carries
// the call's sequence number, so no flags, fences or clearing are needed
// between calls.  A workgroup waits only for tiles of earlier blocks (lower
// workgroup indices) and for the other tiles of its own block (adjacent
// indices); workgroups are dispatched in index order, so every wait ends once
// S / TR workgroups fit on the device at once.  Each wait is still bounded:
// after TRSV_SPIN rounds the workgroup stores the sequence number into the
// abort word and writes NaN to its rows (the tiles that need them then stop
// at their first round), so a lost value shows up as NaN and as the abort
// word, never as a hang.
constexpr int TRSV_SPIN = 1 << 22;

template <typename T>
CLU_DEV void tag_store(unsigned long long *tag, long i, T v, unsigned seq) {
  constexpr int VW = sizeof(T) / 4;
  unsigned u[VW];
  __builtin_memcpy(u, &v, sizeof(T));
  const unsigned long long hi = (unsigned long long)seq << 32;
#pragma unroll
  for (int k = 0; k < VW; ++k) st_agent(tag + i * VW + k, hi | u[k]);
}

// Read the S values published at tag into dst (LDS).  Polls one value per
// tile until each carries seq, then reads all S and checks every word, again
// until all do.  Returns false (the same in every thread) when it gave up.
template <typename T, int S, int TR>
CLU_DEV bool tag_wait(const unsigned long long *tag, unsigned seq, T *dst, unsigned *abortw) {
  constexpr int VW = sizeof(T) / 4;
  constexpr int NTILE = S / TR;
  static_assert(NTILE <= 256, "one polling thread per tile");
  const int t = threadIdx.x;
  bool full = false;
  for (int spin = 0;; ++spin) {
    bool ok = true;
    if (!full) {
      if (t < NTILE) {
        const long i = (long)t * TR + TR - 1;
#pragma unroll
        for (int k = 0; k < VW; ++k) ok = ok && (ld_agent(tag + i * VW + k) >> 32) == seq;
      }
    } else {
      for (int i = t; i < S; i += blockDim.x) {
        unsigned u[VW];
#pragma unroll
        for (int k = 0; k < VW; ++k) {
          const unsigned long long w = ld_agent(tag + (long)i * VW + k);
          ok = ok && (w >> 32) == seq;
          u[k] = (unsigned)w;
        }
        T v;
        __builtin_memcpy(&v, u, sizeof(T));
        dst[i] = v;
      }
    }
    if (__syncthreads_and(ok)) {
      if (full) return true;
      full = true;
      continue;
    }
    const bool quit = spin >= TRSV_SPIN || (t == 0 && ld_agent(abortw) == seq);
    if (__syncthreads_or(quit)) {
      if (t == 0) st_agent(abortw, seq);
      return false;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

template <typename T, int S, bool UPPER, bool TRANS, int WR>
__global__ __launch_bounds__(256) void trsv_tiles_kernel(
    int n, const T *__restrict__ F, long lda, const T *__restrict__ Inv,
    const T *__restrict__ rhs, T *__restrict__ out, unsigned long long *tagx,
    unsigned long long *tagt, unsigned seq, unsigned *abortw) {
  constexpr int VW = sizeof(T) / 4;
  constexpr int VE = 16 / sizeof(T);          // elements per 16-byte load
  constexpr int NV = S / (64 * VE);           // loads per lane per row
  constexpr int TR = 4 * WR;                  // rows per tile (WR per wave)
  constexpr int RB0 = NV >= 16 ? 1 : 16 / NV;
  constexpr int RB = RB0 < WR ? RB0 : WR;     // rows per batch: up to 16 loads in flight
  static_assert(NV >= 1 && WR % RB == 0 && S % TR == 0, "tile shape");
  typedef T vec_t __attribute__((ext_vector_type(VE)));
  __shared__ __attribute__((aligned(16))) T xs[S];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ntile = n / TR, nblk = n / S;
  const int q = UPPER ? ntile - 1 - (int)blockIdx.x : (int)blockIdx.x;
  const int r0 = q * TR, blk = r0 / S;
  const int rw = r0 + wave * WR;  // this wave's WR rows
  double acc[WR];
#pragma unroll
  for (int r = 0; r < WR; ++r) acc[r] = 0.0;

  // acc[r] += sum over this lane's columns c of M[rw + r][c] * xs[c], the
  // columns of M starting at the block's first column (row stride ldm)
  auto dots = [&](const T *M, long ldm) {
    vec_t xv[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) xv[v] = *reinterpret_cast<const vec_t *>(&xs[(v * 64 + lane) * VE]);
#pragma unroll
    for (int b = 0; b < WR; b += RB) {
      vec_t a[RB][NV];
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int v = 0; v < NV; ++v)
          a[r][v] = *reinterpret_cast<const vec_t *>(M + (long)(rw + b + r) * ldm +
                                                     (v * 64 + lane) * VE);
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
          for (int e = 0; e < VE; ++e) acc[b + r] += (double)a[r][v][e] * (double)xv[v][e];
    }
  };
  // TRANS: the tile's rows of op(F) are columns r0 .. r0 + TR - 1 of F; lane
  // group rg of CQ lanes reads one row of F's S-row block (CQ 16-byte pieces
  // of TR columns), NRG rows at a time, and keeps VE column sums; the row
  // groups' sums are added in LDS in a fixed order at the end.
  constexpr int CQ = TR / VE, NRG = 256 / CQ;
  static_assert(TR % VE == 0 && S % NRG == 0, "transposed tile shape");
  __shared__ double red[TRANS ? NRG : 1][TRANS ? TR : 1];
  const int cq = threadIdx.x % CQ, rg = threadIdx.x / CQ;
  double acct[VE];
#pragma unroll
  for (int e = 0; e < VE; ++e) acct[e] = 0.0;
  auto dots_t = [&](const T *M) {  // M: row 0 of the block, column r0
    constexpr int U = 8;            // rows in flight per lane
    static_assert((S / NRG) % U == 0, "transposed unroll");
    for (int k0 = rg; k0 < S; k0 += U * NRG) {
      vec_t a[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        a[u] = *reinterpret_cast<const vec_t *>(M + (long)(k0 + u * NRG) * lda + cq * VE);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const double z = (double)xs[k0 + u * NRG];
#pragma unroll
        for (int e = 0; e < VE; ++e) acct[e] += (double)a[u][e] * z;
      }
    }
  };
  // lane r (< WR) of each wave ends with the sum of row rw + r
  auto reduce_t = [&]() {
#pragma unroll
    for (int e = 0; e < VE; ++e) red[rg][cq * VE + e] = acct[e];
    __syncthreads();
    double mine = 0.0;
    if (lane < WR)
      for (int g2 = 0; g2 < NRG; ++g2) mine += red[g2][wave * WR + lane];
    return mine;
  };
  // lane r (< WR) ends with the sum of row rw + r
  auto reduce = [&]() {
    double mine = 0.0;
#pragma unroll
    for (int r = 0; r < WR; ++r) {
      const double s = wave_sum(acc[r]);
      if (lane == r) mine = s;
      acc[r] = 0.0;
    }
    return mine;
  };

  bool ok = true;
  const int nj = UPPER ? nblk - 1 - blk : blk;
  for (int jj = 0; jj < nj && ok; ++jj) {
    const int j = UPPER ? nblk - 1 - jj : jj;
    ok = tag_wait<T, S, TR>(tagx + (long)j * S * VW, seq, xs, abortw);
    if (ok) {
      if constexpr (TRANS) dots_t(F + (long)j * S * lda + r0);
      else dots(F + (long)j * S, lda);
    }
    __syncthreads();  // every lane has its x values before xs is refilled
  }
  const T nan = __builtin_nan("");
  double s;
  if constexpr (TRANS) s = reduce_t();
  else s = reduce();
  if (lane < WR) {
    const int r = rw + lane;
    tag_store<T>(tagt, r, ok ? (T)((double)rhs[r] - s) : nan, seq);
  }
  if (ok) ok = tag_wait<T, S, TR>(tagt + (long)blk * S * VW, seq, xs, abortw);
  // row r of block blk's inverse is at Inv + r * S (blocks of S x S, ld S)
  if (ok) dots(Inv, S);
  s = reduce();
  if (lane < WR) {
    const int r = rw + lane;
    const T xr = ok ? (T)s : nan;
    out[r] = xr;
    tag_store<T>(tagx, r, xr, seq);
  }
}

template <typename T>
int trsv_tiles(bool upper, bool trans, int n, int S, const T *F, long lda, const T *Inv,
               const T *rhs, T *out, unsigned long long *tagx, unsigned long long *tagt,
               unsigned seq, unsigned *abortw, hipStream_t s) {
  constexpr int VE = 16 / sizeof(T);
  constexpr int WR = 8;  // 32-row tiles (16 rows measured the same, 64 slower)
  if (n <= 0 || (S != 256 && S != 1024) || n % S || lda % VE ||
      ((uintptr_t)F & 15) || ((uintptr_t)Inv & 15))
    return -1;
  const dim3 grid(n / (4 * WR)), block(256);
#define CLU_TRSV(SS, UP, TT)                                                                    \
  hipLaunchKernelGGL((trsv_tiles_kernel<T, SS, UP, TT, WR>), grid, block, 0, s, n, F, lda, Inv, \
                     rhs, out, tagx, tagt, seq, abortw)
#define CLU_TRSV_T(SS, UP)   \
  do {                       \
    if (trans) CLU_TRSV(SS, UP, true); \
    else CLU_TRSV(SS, UP, false);      \
  } while (0)
  if (S == 256) {
    if (upper) CLU_TRSV_T(256, true);
    else CLU_TRSV_T(256, false);
  } else {
    if (upper) CLU_TRSV_T(1024, true);
    else CLU_TRSV_T(1024, false);
  }
#undef CLU_TRSV_T
#undef CLU_TRSV
  CLU_LAUNCHED("trsv_tiles_kernel");
  return CLU_OK;
}
template int trsv_tiles<double>(bool, bool, int, int, const double *, long, const double *,
                                const double *, double *, unsigned long long *,
                                unsigned long long *, unsigned, unsigned *, hipStream_t);
template int trsv_tiles<float>(bool, bool, int, int, const float *, long, const float *, const float *,
                               float *, unsigned long long *, unsigned long long *, unsigned,
                               unsigned *, hipStream_t);

}  // namespace clu
