// gemm.hip -- FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4_f64).
//
// Replaces the reference's per-block trailing update
//   cblas_dgemm(RowMajor, NoTrans, NoTrans, b, b, b, -1.0, L, b, U, b, 1.0, C, b)
// (lu.C:26-43) with one launch over the whole trailing matrix of a step.
//
// Work decomposition: 256-thread workgroups (4 wave64s) own a BM x BN tile
// of C; each wave owns a (BM/WM) x (BN/WN) sub-tile made of 16x16 MFMA
// blocks held in registers.  K is walked in chunks of 16 that are staged
// global -> registers -> LDS with two LDS buffers, so the loads of chunk
// k+1 are in flight while chunk k feeds the matrix cores; one barrier per
// chunk.  LDS tiles are stored k-major (As[k][m], Bs[k][n]) so that the 16
// lanes that share an MFMA k index read 16 consecutive doubles.
//
// MFMA 16x16x4 f64 operand layout (per lane l, 0..63):
//   A[i][k] from lane i + 16k,  B[k][j] from lane j + 16k,
//   D[i][j] in lane j + 16*(i%4), element i/4 of the 4-double result
// (measured on gfx950; the FP32/FP16/BF16 forms use lane j + 16*(i/4),
// element i%4 instead).
#include "common.h"
#include "kernels.h"

namespace clu {

constexpr int GBK = 16;

// Map a workgroup id to a C tile.  Workgroups are handed to the 8 XCDs
// round-robin (id % 8); renumber so that each XCD receives a contiguous
// range of tiles, then walk that range in groups of GM tile rows so the
// tiles resident on one XCD share rows of A and columns of B in its L2.
CLU_DEV void tile_coords(int bid, int nblk, int tiles_m, int tiles_n, int &tm,
                         int &tn) {
  const int per = nblk / 8;
  int lid = bid;
  if (bid < per * 8) lid = (bid % 8) * per + bid / 8;
  constexpr int GM = 8;
  const int group_size = GM * tiles_n;
  const int group = lid / group_size;
  const int first_m = group * GM;
  const int gm = min(tiles_m - first_m, GM);
  const int in_group = lid - group * group_size;
  tm = first_m + in_group % gm;
  tn = in_group / gm;
}

template <int BM, int BN, int WM, int WN, bool TA, bool TB>
__global__ __launch_bounds__(256, 2) void dgemm_mfma_kernel(
    int M, int N, int K, double alpha, const double *__restrict__ A, long lda,
    const double *__restrict__ B, long ldb, double beta, double *__restrict__ C,
    long ldc, int tiles_m, int tiles_n) {
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  constexpr int A_PER = BM * GBK / 256, B_PER = BN * GBK / 256;
  constexpr int ALD = BM + 1, BLD = BN + 1;
  static_assert(WM * WN == 4, "four waves");
  static_assert(TM >= 1 && TN >= 1, "wave tile");
  static_assert(A_PER >= 1 && B_PER >= 1, "load split");

  __shared__ double As[2][GBK][ALD];
  __shared__ double Bs[2][GBK][BLD];

  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm0 = (wave / WN) * WTM, wn0 = (wave % WN) * WTN;
  const int lr = lane & 15, lq = lane >> 4;

  double ra[A_PER], rb[B_PER];
  auto load = [&](int k0) {
#pragma unroll
    for (int r = 0; r < A_PER; ++r) {
      const int e = t + 256 * r;
      int m, k;
      if (!TA) { m = e / GBK; k = e % GBK; } else { k = e / BM; m = e % BM; }
      const int gm = m0 + m, gk = k0 + k;
      ra[r] = (gm < M && gk < K)
                  ? (!TA ? A[(long)gm * lda + gk] : A[(long)gk * lda + gm])
                  : 0.0;
    }
#pragma unroll
    for (int r = 0; r < B_PER; ++r) {
      const int e = t + 256 * r;
      int n, k;
      if (!TB) { k = e / BN; n = e % BN; } else { n = e / GBK; k = e % GBK; }
      const int gn = n0 + n, gk = k0 + k;
      rb[r] = (gn < N && gk < K)
                  ? (!TB ? B[(long)gk * ldb + gn] : B[(long)gn * ldb + gk])
                  : 0.0;
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int r = 0; r < A_PER; ++r) {
      const int e = t + 256 * r;
      int m, k;
      if (!TA) { m = e / GBK; k = e % GBK; } else { k = e / BM; m = e % BM; }
      As[buf][k][m] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < B_PER; ++r) {
      const int e = t + 256 * r;
      int n, k;
      if (!TB) { k = e / BN; n = e % BN; } else { n = e / GBK; k = e % GBK; }
      Bs[buf][k][n] = rb[r];
    }
  };

  double4_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};

  const int nk = (K + GBK - 1) / GBK;
  load(0);
  stash(0);
  __syncthreads();
  for (int kc = 0; kc < nk; ++kc) {
    const int buf = kc & 1;
    if (kc + 1 < nk) load((kc + 1) * GBK);
#pragma unroll
    for (int s = 0; s < GBK / 4; ++s) {
      double a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) a[i] = As[buf][4 * s + lq][wm0 + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < TN; ++j) b[j] = Bs[buf][4 * s + lq][wn0 + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (kc + 1 < nk) stash(buf ^ 1);
    __syncthreads();
  }

  // epilogue: C = alpha * acc + beta * C
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int col = n0 + wn0 + j * 16 + lr;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = m0 + wm0 + i * 16 + lq + 4 * v;
        if (row < M && col < N) {
          double *c = C + (long)row * ldc + col;
          const double prod = alpha * acc[i][j][v];
          *c = beta == 0.0 ? prod : prod + beta * *c;
        }
      }
    }
}

template <int BM, int BN, int WM, int WN, bool TA, bool TB>
static int launch(int M, int N, int K, double alpha, const double *A, long lda,
                  const double *B, long ldb, double beta, double *C, long ldc,
                  hipStream_t s) {
  const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  const long nblk = (long)tiles_m * tiles_n;
  if (nblk > 0x7fffffffL) return -1;
  hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, WM, WN, TA, TB>), dim3((unsigned)nblk),
                     dim3(256), 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                     tiles_m, tiles_n);
  CLU_LAUNCHED("dgemm_mfma_kernel");
  return CLU_OK;
}

template <bool TA, bool TB>
static int dispatch(int M, int N, int K, double alpha, const double *A, long lda,
                    const double *B, long ldb, double beta, double *C, long ldc,
                    hipStream_t s) {
  // narrow right-hand sides (panel-internal updates, multi-RHS solves)
  if (N <= 32)
    return launch<256, 32, 4, 1, TA, TB>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  if (N <= 64)
    return launch<64, 64, 2, 2, TA, TB>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  if (M <= 32)
    return launch<32, 256, 1, 4, TA, TB>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  // small problems: more, smaller tiles to fill 256 CUs
  if ((long)((M + 127) / 128) * ((N + 127) / 128) < 512)
    return launch<64, 64, 2, 2, TA, TB>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  return launch<128, 128, 2, 2, TA, TB>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

int dgemm_variant(int M, int N) {
  if (N <= 32) return 1;
  if (N <= 64) return 4;
  if (M <= 32) return 3;
  if ((long)((M + 127) / 128) * ((N + 127) / 128) < 512) return 4;
  return 0;
}

int dgemm(bool ta, bool tb, int M, int N, int K, double alpha, const double *A,
          long lda, const double *B, long ldb, double beta, double *C, long ldc,
          hipStream_t s) {
  if (M <= 0 || N <= 0) return CLU_OK;
  if (K <= 0 || alpha == 0.0) {
    // C = beta * C: run the kernel with K = 0 (acc stays zero)
    K = 0;
  }
  if (!ta && !tb) return dispatch<false, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  if (ta && !tb) return dispatch<true, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  if (!ta && tb) return dispatch<false, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  return dispatch<true, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

}  // namespace clu
