// ir.hip -- FP64 vector kernels of the GMRES-based iterative refinement.
//
// The mixed-precision solve (charmlu_dgesv_ir) keeps x, the residual
// r = b - A x, the Krylov basis and every inner product in FP64 and uses the
// low-precision LU factors only to apply the preconditioner.  The basis V is
// stored as k rows of n doubles (row-major), so V^T w and V^T y below are
// row dot products and axpy sweeps that read each basis row once.
#include "common.h"
#include "kernels.h"

namespace clu {

// h[i] = sum_j V[i][j] * w[j] for i < k (one workgroup per row of V)
__global__ __launch_bounds__(256) void basis_dot_kernel(int k, long n,
                                                        const double *__restrict__ V, long ldv,
                                                        const double *__restrict__ w,
                                                        double *__restrict__ h) {
  __shared__ double part[4];
  const int i = blockIdx.x;
  double s = 0.0;
  for (long j = threadIdx.x; j < n; j += 256) s += V[(long)i * ldv + j] * w[j];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) h[i] = (part[0] + part[1]) + (part[2] + part[3]);
}

// w[j] += alpha * sum_i h[i] * V[i][j]
__global__ void basis_axpy_kernel(int k, long n, double alpha, const double *__restrict__ V,
                                  long ldv, const double *__restrict__ h, double *__restrict__ w) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double s = 0.0;
  for (int i = 0; i < k; ++i) s += h[i] * V[(long)i * ldv + j];
  w[j] += alpha * s;
}

// out[0] = sum x^2, out[1] = max |x|: one workgroup, a fixed summation
// order, so the result does not depend on scheduling
__global__ __launch_bounds__(1024) void norms_kernel(long n, const double *__restrict__ x,
                                                     double *__restrict__ out) {
  __shared__ double ps[16], pm[16];
  double s = 0.0, m = 0.0;
  for (long j = threadIdx.x; j < n; j += 1024) {
    const double v = x[j];
    s += v * v;
    m = fmax(m, fabs(v));
  }
  s = wave_sum(s);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) { ps[threadIdx.x >> 6] = s; pm[threadIdx.x >> 6] = m; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ts = 0.0, tm = 0.0;
    for (int k = 0; k < 16; ++k) { ts += ps[k]; tm = fmax(tm, pm[k]); }
    out[0] = ts;
    out[1] = tm;
  }
}

// y = a * x (y may alias x)
__global__ void dscal_copy_kernel(long n, double a, const double *__restrict__ x, double *y) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) y[j] = a * x[j];
}

// y32 = (float)(a * x)
__global__ void d2s_kernel(long n, double a, const double *__restrict__ x, float *__restrict__ y) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) y[j] = (float)(a * x[j]);
}
// y = a * (double)x32
__global__ void s2d_kernel(long n, double a, const float *__restrict__ x, double *__restrict__ y) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) y[j] = a * (double)x[j];
}

static unsigned nblocks(long n) { return (unsigned)((n + 255) / 256); }

int basis_dot(int k, long n, const double *V, long ldv, const double *w, double *h,
              hipStream_t s) {
  if (k <= 0) return CLU_OK;
  hipLaunchKernelGGL(basis_dot_kernel, dim3(k), dim3(256), 0, s, k, n, V, ldv, w, h);
  CLU_LAUNCHED("basis_dot_kernel");
  return CLU_OK;
}
int basis_axpy(int k, long n, double alpha, const double *V, long ldv, const double *h,
               double *w, hipStream_t s) {
  if (k <= 0 || n <= 0) return CLU_OK;
  hipLaunchKernelGGL(basis_axpy_kernel, dim3(nblocks(n)), dim3(256), 0, s, k, n, alpha, V, ldv,
                     h, w);
  CLU_LAUNCHED("basis_axpy_kernel");
  return CLU_OK;
}
int vec_norms(long n, const double *x, double *out, hipStream_t s) {
  hipLaunchKernelGGL(norms_kernel, dim3(1), dim3(1024), 0, s, n, x, out);
  CLU_LAUNCHED("norms_kernel");
  return CLU_OK;
}
int vec_scale(long n, double a, const double *x, double *y, hipStream_t s) {
  if (n <= 0) return CLU_OK;
  hipLaunchKernelGGL(dscal_copy_kernel, dim3(nblocks(n)), dim3(256), 0, s, n, a, x, y);
  CLU_LAUNCHED("dscal_copy_kernel");
  return CLU_OK;
}
int vec_d2s(long n, double a, const double *x, float *y, hipStream_t s) {
  if (n <= 0) return CLU_OK;
  hipLaunchKernelGGL(d2s_kernel, dim3(nblocks(n)), dim3(256), 0, s, n, a, x, y);
  CLU_LAUNCHED("d2s_kernel");
  return CLU_OK;
}
int vec_s2d(long n, double a, const float *x, double *y, hipStream_t s) {
  if (n <= 0) return CLU_OK;
  hipLaunchKernelGGL(s2d_kernel, dim3(nblocks(n)), dim3(256), 0, s, n, a, x, y);
  CLU_LAUNCHED("s2d_kernel");
  return CLU_OK;
}

}  // namespace clu
