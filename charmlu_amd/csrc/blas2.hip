// blas2.hip -- the rank-1 update and matrix-vector product of the
// platformBlas.h surface on device pointers.
//
// dger replaces cblas_dger(RowMajor, M, N, alpha, x, incx, y, incy, A, lda)
// (updateLsubBlock, lu.C:247-252): one thread per element of A, rows
// contiguous across lanes.  dgemv replaces cblas_dgemv(RowMajor, NoTrans /
// Trans, ...) (recvXvec, benchmark.C:494-496): 16 lanes per row for the
// row-major no-transpose product, two columns per lane for the transpose.
// Negative increments follow the CBLAS convention (the vector is walked from
// its last stored element).
#include "common.h"
#include "kernels.h"

namespace clu {

CLU_DEV long vidx(long i, long n, long inc) { return inc > 0 ? i * inc : (n - 1 - i) * (-inc); }

__global__ void dger_kernel(int M, int N, double alpha, const double *__restrict__ x, long incx,
                            const double *__restrict__ y, long incy, double *__restrict__ A,
                            long lda) {
  const long total = (long)M * N;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const long i = e / N, j = e % N;
    A[i * lda + j] += (alpha * x[vidx(i, M, incx)]) * y[vidx(j, N, incy)];
  }
}

int dger(int M, int N, double alpha, const double *x, long incx, const double *y, long incy,
         double *A, long lda, hipStream_t s) {
  if (M <= 0 || N <= 0 || alpha == 0.0) return CLU_OK;
  long blocks = ((long)M * N + 255) / 256;
  if (blocks > 65536L * 16) blocks = 65536L * 16;
  hipLaunchKernelGGL(dger_kernel, dim3((unsigned)blocks), dim3(256), 0, s, M, N, alpha, x, incx,
                     y, incy, A, lda);
  CLU_LAUNCHED("dger_kernel");
  return CLU_OK;
}

// y = alpha * A x + beta * y, A is M x N row-major
__global__ __launch_bounds__(256) void dgemv_n_kernel(int M, int N, double alpha,
                                                      const double *__restrict__ A, long lda,
                                                      const double *__restrict__ x, long incx,
                                                      double beta, double *__restrict__ y,
                                                      long incy) {
  const int row = blockIdx.x * 16 + (threadIdx.x >> 4);
  const int sub = threadIdx.x & 15;
  double s = 0.0;
  // unit-stride x and 16-byte aligned rows: 16-byte loads and four partial
  // sums, so each lane keeps several loads in flight (the refinement's
  // residual streams all of A: 5.0 TB/s with one 8-byte load chain)
  const bool vec = incx == 1 && ((lda | N) & 1) == 0 &&
                   (((unsigned long)A | (unsigned long)x) & 15) == 0;
  if (row < M && vec) {
    const double2 *a2 = reinterpret_cast<const double2 *>(A + (long)row * lda);
    const double2 *x2 = reinterpret_cast<const double2 *>(x);
    const int N2 = N / 2;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int k = sub;
    for (; k + 48 < N2; k += 64) {
      const double2 p0 = a2[k], p1 = a2[k + 16], p2 = a2[k + 32], p3 = a2[k + 48];
      const double2 q0 = x2[k], q1 = x2[k + 16], q2 = x2[k + 32], q3 = x2[k + 48];
      s0 += p0.x * q0.x + p0.y * q0.y;
      s1 += p1.x * q1.x + p1.y * q1.y;
      s2 += p2.x * q2.x + p2.y * q2.y;
      s3 += p3.x * q3.x + p3.y * q3.y;
    }
    for (; k < N2; k += 16) {
      const double2 p0 = a2[k], q0 = x2[k];
      s0 += p0.x * q0.x + p0.y * q0.y;
    }
    s = (s0 + s1) + (s2 + s3);
  } else if (row < M) {
    for (int k = sub; k < N; k += 16) s += A[(long)row * lda + k] * x[vidx(k, N, incx)];
  }
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) s += __shfl_xor(s, off, 16);
  if (row < M && sub == 0) {
    double *yp = y + vidx(row, M, incy);
    *yp = beta == 0.0 ? alpha * s : alpha * s + beta * *yp;
  }
}

// y = alpha * A^T x + beta * y, A is M x N row-major, y has N entries.
// 128 columns per workgroup, two per lane (16-byte loads when the rows are
// 16-byte aligned); the 8 waves take every 8th row, four rows in flight per
// lane, and their partial sums are added in a fixed order through LDS (one
// lane per column with one load in flight streamed A at a fraction of HBM
// speed: the adjoint's transposed products dominated its refinement).
__global__ __launch_bounds__(512) void dgemv_t_kernel(int M, int N, double alpha,
                                                      const double *__restrict__ A, long lda,
                                                      const double *__restrict__ x, long incx,
                                                      double beta, double *__restrict__ y,
                                                      long incy) {
  __shared__ double part[8][2][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c0 = blockIdx.x * 128 + 2 * lane;
  const bool vec = ((lda | N) & 1) == 0 && ((unsigned long)A & 15) == 0;
  double sx[4] = {0.0, 0.0, 0.0, 0.0}, sy[4] = {0.0, 0.0, 0.0, 0.0};
  if (c0 < N) {
    if (vec) {
      int k = wave;
      for (; k + 24 < M; k += 32) {
        double2 a[4];
        double xv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          a[u] = *reinterpret_cast<const double2 *>(A + (long)(k + 8 * u) * lda + c0);
          xv[u] = x[vidx(k + 8 * u, M, incx)];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          sx[u] += a[u].x * xv[u];
          sy[u] += a[u].y * xv[u];
        }
      }
      for (; k < M; k += 8) {
        const double2 a0 = *reinterpret_cast<const double2 *>(A + (long)k * lda + c0);
        const double x0 = x[vidx(k, M, incx)];
        sx[0] += a0.x * x0;
        sy[0] += a0.y * x0;
      }
    } else {
      for (int k = wave; k < M; k += 8) {
        const double x0 = x[vidx(k, M, incx)];
        sx[0] += A[(long)k * lda + c0] * x0;
        if (c0 + 1 < N) sy[0] += A[(long)k * lda + c0 + 1] * x0;
      }
    }
  }
  part[wave][0][lane] = (sx[0] + sx[1]) + (sx[2] + sx[3]);
  part[wave][1][lane] = (sy[0] + sy[1]) + (sy[2] + sy[3]);
  __syncthreads();
  if (wave < 2) {
    const int c = c0 + wave;
    if (c < N) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < 8; ++w) t += part[w][wave][lane];
      double *yp = y + vidx(c, N, incy);
      *yp = beta == 0.0 ? alpha * t : alpha * t + beta * *yp;
    }
  }
}

int dgemv(bool trans, int M, int N, double alpha, const double *A, long lda, const double *x,
          long incx, double beta, double *y, long incy, hipStream_t s) {
  if (M <= 0 || N <= 0) return CLU_OK;
  if (!trans)
    hipLaunchKernelGGL(dgemv_n_kernel, dim3((M + 15) / 16), dim3(256), 0, s, M, N, alpha, A, lda,
                       x, incx, beta, y, incy);
  else
    hipLaunchKernelGGL(dgemv_t_kernel, dim3((N + 127) / 128), dim3(512), 0, s, M, N, alpha, A, lda,
                       x, incx, beta, y, incy);
  CLU_LAUNCHED("dgemv_kernel");
  return CLU_OK;
}

}  // namespace clu
