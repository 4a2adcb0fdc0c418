// trsm.hip -- triangular solves on a block of right-hand-side columns.
//
// trsm_llnu replaces the reference's U-row solve
//   cblas_dtrsm(RowMajor, Left, Lower, NoTrans, Unit, b, b, 1.0, L, b, A, b)
// (computeU, lu.C:180-189), applied to the whole b x N_t row panel of a step
// in one launch.  trsm_lunn is the diagonal-block step of the backward solve
// (backwardSolve, lu.ci:408-436); llnu also serves the forward solve
// (forwardSolve, lu.ci:387-407).  The transposed forms serve A^T x = b.
//
// One wave64 owns 64 right-hand-side columns and keeps them in LDS
// (M x 64 values, M <= 256); each lane runs the substitution for its own
// column, reading the triangular factor with wave-uniform (scalar) loads.
// Each dot product is split over four independent partial sums so the FMA
// latency overlaps; they are combined in a fixed order.
//
// gemv_sub is the off-diagonal part of the block solve (offDiagSolve,
// lu.ci:439-456): y -= A x for a tall block column, 16 lanes per row.
#include "common.h"
#include "kernels.h"

namespace clu {

// op(T)[i][k] = TR ? T[k][i] : T[i][k]; OPLOW: op(T) is lower triangular
// (forward substitution), else upper (backward).  UNIT: unit diagonal.
template <typename T, bool OPLOW, bool TR, bool UNIT>
__global__ __launch_bounds__(64) void trsm_kernel(int M, int N, T alpha, const T *__restrict__ F,
                                                  long ldf, T *__restrict__ B, long ldb) {
  extern __shared__ unsigned char smem_raw[];
  T *X = reinterpret_cast<T *>(smem_raw);  // [M][64]
  const int lane = threadIdx.x;
  const int col = blockIdx.x * 64 + lane;
  const bool ok = col < N;
  for (int i = 0; i < M; ++i) X[i * 64 + lane] = ok ? alpha * B[(long)i * ldb + col] : (T)0;
  __syncthreads();
  for (int ii = 0; ii < M; ++ii) {
    const int i = OPLOW ? ii : M - 1 - ii;
    T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    const int k_lo = OPLOW ? 0 : i + 1;
    const int k_hi = OPLOW ? i : M;
    auto f = [&](int kk) -> T { return TR ? F[(long)kk * ldf + i] : F[(long)i * ldf + kk]; };
    int k = k_lo;
    for (; k + 3 < k_hi; k += 4) {
      s0 += f(k) * X[k * 64 + lane];
      s1 += f(k + 1) * X[(k + 1) * 64 + lane];
      s2 += f(k + 2) * X[(k + 2) * 64 + lane];
      s3 += f(k + 3) * X[(k + 3) * 64 + lane];
    }
    for (; k < k_hi; ++k) s0 += f(k) * X[k * 64 + lane];
    T x = X[i * 64 + lane] - ((s0 + s1) + (s2 + s3));
    if (!UNIT) x = x / F[(long)i * ldf + i];
    X[i * 64 + lane] = x;
  }
  if (ok)
    for (int i = 0; i < M; ++i) B[(long)i * ldb + col] = X[i * 64 + lane];
}

template <typename T, bool OPLOW, bool TR, bool UNIT>
static int trsm_launch(int M, int N, T alpha, const T *F, long ldf, T *B, long ldb,
                       hipStream_t s) {
  if (M <= 0 || N <= 0) return CLU_OK;
  if (M > 256) return -1;
  const size_t shm = (size_t)M * 64 * sizeof(T);
  hipLaunchKernelGGL((trsm_kernel<T, OPLOW, TR, UNIT>), dim3((N + 63) / 64), dim3(64), shm, s,
                     M, N, alpha, F, ldf, B, ldb);
  CLU_LAUNCHED("trsm_kernel");
  return CLU_OK;
}

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

// y[j] -= sum_k A[k][j] x[k]; one lane per column j, K split over 4 waves
template <typename T>
__global__ __launch_bounds__(256) void gemv_sub_t_kernel(int M, int K, const T *__restrict__ A,
                                                         long lda, const T *__restrict__ x,
                                                         T *__restrict__ y) {
  __shared__ double part[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + lane;
  double s = 0.0;
  if (j < M)
    for (int k = wave; k < K; k += 4) s += (double)A[(long)k * lda + j] * (double)x[k];
  part[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && j < M)
    y[j] = (T)((double)y[j] - ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane])));
}

template <typename T>
int gemv_sub(bool trans, int M, int K, const T *A, long lda, const T *x, T *y, hipStream_t s) {
  if (M <= 0 || K <= 0) return CLU_OK;
  if (!trans) {
    hipLaunchKernelGGL((gemv_sub_n_kernel<T>), dim3((M + 15) / 16), dim3(256), 0, s, M, K, A,
                       lda, x, y);
  } else {
    hipLaunchKernelGGL((gemv_sub_t_kernel<T>), dim3((M + 63) / 64), dim3(256), 0, s, M, K, A,
                       lda, x, y);
  }
  CLU_LAUNCHED("gemv_sub_kernel");
  return CLU_OK;
}
template int gemv_sub<double>(bool, int, int, const double *, long, const double *, double *,
                              hipStream_t);
template int gemv_sub<float>(bool, int, int, const float *, long, const float *, float *,
                             hipStream_t);

}  // namespace clu
