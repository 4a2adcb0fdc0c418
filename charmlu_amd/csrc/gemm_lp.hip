// gemm_lp.hip -- the low-precision trailing update of the mixed-precision LU.
//
// The FP32 working matrix is updated as C -= L21 * U12 with both operands
// rounded to BF16 or FP16 and products accumulated in FP32 on the matrix
// cores (v_mfma_f32_16x16x32_bf16 / _f16): the reference's
// updateMatrix dgemm (lu.C:26-43) at reduced input precision.  Accuracy is
// recovered afterwards by FP64 iterative refinement.
//
// Operands are packed once per step into 16-bit buffers whose rows hold K
// contiguous values (L21 as is, U12 transposed), so each lane's MFMA operand
// -- 8 consecutive k of one row -- is a single 16-byte LDS read.  K is
// zero-padded to a multiple of 32.
//
// hgemm_nt_kernel: 128x128 C tiles, 4 waves of 64x64, K in stages of 32
// loaded with global_load_lds_dwordx4 straight into LDS, 4 stages in flight.
// LDS rows are 64 bytes (4 chunks of 16 bytes); the chunk index is XORed with
// (row >> 2) & 3 so the 16 lanes reading one operand column hit distinct
// banks.
//
// MFMA 16x16x32 (BF16/FP16) layout, per lane l: A[i][k] in lane i + 16*(k/8),
// element k%8; B[k][j] in lane j + 16*(k/8), element k%8; D[i][j] in lane
// j + 16*(i/4), element i%4 (measured on gfx950).
//
// fgemm_nt_kernel (FP8 e4m3): the same staging, with gfx950's block-scaled
// v_mfma_scale_f32_32x32x64_f8f6f4, whose rate (4.9 PFLOP/s measured in a
// register-only loop) is twice that of v_mfma_f32_16x16x32_fp8_fp8 (2.5).
// Lane l holds 32 consecutive k of row l % 32 (k block l / 32), one 64-byte
// stage row is one instruction's K; D[i][j] in lane j + 32 * ((i / 4) % 2),
// element 4 * (i / 8) + i % 4 (checked on gfx950 against a brute-force
// product, with scales).  The per-tile power-of-two scales of the operands
// are the instruction's E8M0 block scales (2^(s - 127)), so products come
// out unscaled and C is used as it is.
#include <hip/hip_bf16.h>
#include <stdlib.h>

#include "common.h"
#include "kernels.h"

namespace clu {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(3))) void *lds_ptr_t;
typedef __attribute__((address_space(1))) void *gbl_ptr_t;

template <typename T16> struct Lp;
template <> struct Lp<__bf16> {
  typedef bf16x8_t vec_t;
  static CLU_DEV float4_t mfma(vec_t a, vec_t b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  static CLU_DEV __bf16 from(float x) { return (__bf16)x; }
};
template <> struct Lp<_Float16> {
  typedef f16x8_t vec_t;
  static CLU_DEV float4_t mfma(vec_t a, vec_t b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
  static CLU_DEV _Float16 from(float x) { return (_Float16)x; }
};
typedef int int8v_t __attribute__((ext_vector_type(8)));
typedef int int4v_t __attribute__((ext_vector_type(4)));

// s_waitcnt vmcnt(N): on gfx9 the counter's low 4 bits are bits 3:0 of the
// immediate and its high 2 bits are bits 15:14 (expcnt and lgkmcnt left at
// their maxima, i.e. not waited for)
template <int N>
CLU_DEV void wait_vm_lp() {
  static_assert(N >= 0 && N < 64, "vmcnt");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

template <typename T16, int BM, int BN, int WM, int WN, int NST, int OCC, bool CINIT, bool VEC,
          bool EDGE = true, int GMT = 0>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64, OCC) void hgemm_nt_kernel(
    int M, int N, int K, float alpha, const T16 *__restrict__ A, long lda,
    const T16 *__restrict__ B, long ldb, float beta, float *__restrict__ C, long ldc,
    int tiles_m, int tiles_n) {
  // BM x BN C tile, one wave per WM x WN block of it (WM / 16 x WN / 16
  // blocks of 16 x 16).  A stage holds 64 bytes of K per row: 32 16-bit
  // values.  EDGE = false: every tile lies inside C.
  constexpr int NT = (BM / WM) * (BN / WN) * 64;  // threads
  constexpr int EPC = 16 / sizeof(T16);  // elements per 16-byte chunk
  constexpr int BK = 4 * EPC;
  constexpr int WGN = BN / WN;
  constexpr int BI = WN / 16, BJ = WM / 16;  // 16 x 16 blocks per wave (columns, rows)
  constexpr int HALF = BM * 64;          // bytes of the A operand's stage
  constexpr int STAGE = (BM + BN) * 64;
  constexpr int CA = BM * 4 / NT, CB = BN * 4 / NT;  // 16-byte chunks per thread
  static_assert(BM % WM == 0 && BN % WN == 0 && WM % 16 == 0 && WN % 16 == 0, "tile");
  static_assert(CA >= 1 && CB >= 1 && CA * NT == BM * 4 && CB * NT == BN * 4, "staging");
  __shared__ __attribute__((aligned(1024))) char smem[NST * STAGE];
  typedef typename Lp<T16>::vec_t vec_t;

  // XCD-aware tile order: consecutive workgroups go to the 8 XCDs in turn,
  // so each XCD gets a contiguous range of tiles, walked in groups of GM
  // tile rows (512 rows of A; for 128-row tiles GM = 4 measured best: 325
  // TFLOP/s at 65024^2 x 512 BF16 vs 315 with 8 and 295 with 16; without the
  // XCD remap 293 and 1.6x the fetched bytes; at K = 1024, 548 with 4 or 8,
  // 530 with 2, 534 with 16)
  int tm, tn;
  {
    const int bid = blockIdx.x, nblk = gridDim.x, per = nblk / 8;
    int lid = bid;
    if (bid < per * 8) lid = (bid % 8) * per + bid / 8;
    constexpr int GM = GMT > 0 ? GMT : 512 / BM;
    const int group_size = GM * tiles_n, group = lid / group_size, first_m = group * GM;
    const int gm = min(tiles_m - first_m, GM), in_group = lid - group * group_size;
    tm = first_m + in_group % gm;
    tn = in_group / gm;
  }
  const int m0 = tm * BM, n0 = tn * BN;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm0 = (wave / WGN) * WM, wn0 = (wave % WGN) * WN;
  const int lr = lane & 15, lq = lane >> 4;
  const bool full = !EDGE || (m0 + BM <= M && n0 + BN <= N);

  // The MFMA's first operand is the B tile (16 C columns), its second the A
  // tile (16 C rows), so D[r][c] of block (i, j) is C[row][col] with
  // row = m0 + wm0 + 16 j + lr and col = n0 + wn0 + 16 i + 4 lq + r: each
  // lane owns 4 consecutive columns of one row (one 16-byte access).
  float4_t acc[BI][BJ];
  auto crow = [&](int j) { return m0 + wm0 + 16 * j + lr; };
  auto ccol = [&](int i) { return n0 + wn0 + 16 * i + 4 * lq; };
  if (CINIT) {
    // C += A B^T: C itself is the initial accumulator, requested before the
    // first operand stage so its latency overlaps the K loop
#pragma unroll
    for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < BJ; ++j) {
        const int row = crow(j), col = ccol(i);
        if (VEC && full) {
          const float4_t *cp4 = reinterpret_cast<const float4_t *>(C + (long)row * ldc + col);
          acc[i][j] = *cp4;
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            acc[i][j][r] = (row < M && col + r < N) ? C[(long)row * ldc + col + r] : 0.f;
        }
      }
  } else {
#pragma unroll
    for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < BJ; ++j) acc[i][j] = (float4_t){0.f, 0.f, 0.f, 0.f};
  }

  // this thread's chunks of A and of B per stage (rows past M or N are
  // clamped to the last one; their products are never stored)
  const T16 *ga[CA], *gb[CB];
  int loffa[CA], loffb[CB];
#pragma unroll
  for (int r = 0; r < CA; ++r) {
    const int c = wave * 64 + NT * r + lane;  // chunk slot in the A stage
    const int row = c >> 2, kq = (c & 3) ^ lds_swz(row);
    ga[r] = A + (long)min(m0 + row, M - 1) * lda + EPC * kq;
    loffa[r] = (wave * 64 + NT * r) * 16;
  }
#pragma unroll
  for (int r = 0; r < CB; ++r) {
    const int c = wave * 64 + NT * r + lane;  // chunk slot in the B stage
    const int row = c >> 2, kq = (c & 3) ^ lds_swz(row);
    gb[r] = B + (long)min(n0 + row, N - 1) * ldb + EPC * kq;
    loffb[r] = HALF + (wave * 64 + NT * r) * 16;
  }
  const int nk = K / BK;
  auto issue = [&](int kc) {
    const int kk = kc * BK;
    char *st = smem + (kc % NST) * STAGE;
#pragma unroll
    for (int r = 0; r < CA; ++r)
      __builtin_amdgcn_global_load_lds((gbl_ptr_t)(ga[r] + kk), (lds_ptr_t)(st + loffa[r]), 16, 0,
                                       0);
#pragma unroll
    for (int r = 0; r < CB; ++r)
      __builtin_amdgcn_global_load_lds((gbl_ptr_t)(gb[r] + kk), (lds_ptr_t)(st + loffb[r]), 16, 0,
                                       0);
  };
  int aoff[BJ], boff[BI];
#pragma unroll
  for (int j = 0; j < BJ; ++j) {
    const int m = wm0 + 16 * j + lr;
    aoff[j] = m * 64 + ((lq ^ lds_swz(m)) << 4);
  }
#pragma unroll
  for (int i = 0; i < BI; ++i) {
    const int n = wn0 + 16 * i + lr;
    boff[i] = HALF + n * 64 + ((lq ^ lds_swz(n)) << 4);
  }

  static_assert(NST >= 2 && NST <= 5 && (CA + CB) * (NST - 2) < 64, "stages");
#pragma unroll
  for (int st = 0; st < NST - 1; ++st)
    if (st < nk) issue(st);
  for (int kc = 0; kc < nk; ++kc) {
    // stage kc has landed once at most the CA + CB loads of each later stage
    // already issued (min(NST - 2, nk - 1 - kc) of them; no stage past the
    // end is loaded) are outstanding; the C prefetch (CINIT) was issued
    // before the stages, so the first wait covers it too
    const int ahead = nk - 1 - kc;
    if (NST >= 5 && ahead >= 3) wait_vm_lp<(CA + CB) * (NST >= 5 ? 3 : 0)>();
    else if (NST >= 4 && ahead >= 2) wait_vm_lp<(CA + CB) * (NST >= 4 ? 2 : 0)>();
    else if (NST >= 3 && ahead >= 1) wait_vm_lp<(CA + CB) * (NST >= 3 ? 1 : 0)>();
    else wait_vm_lp<0>();
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
    if (kc + NST - 1 < nk) issue(kc + NST - 1);
    const char *st = smem + (kc % NST) * STAGE;
    vec_t a[BJ], b[BI];
#pragma unroll
    for (int j = 0; j < BJ; ++j) a[j] = *reinterpret_cast<const vec_t *>(st + aoff[j]);
#pragma unroll
    for (int i = 0; i < BI; ++i) b[i] = *reinterpret_cast<const vec_t *>(st + boff[i]);
#pragma unroll
    for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < BJ; ++j) acc[i][j] = Lp<T16>::mfma(b[i], a[j], acc[i][j]);
  }
  wait_vm_lp<0>();  // no direct-to-LDS write may be in flight when the workgroup exits

  // epilogue: CINIT stores the accumulators; otherwise C = alpha acc + beta C
#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int j = 0; j < BJ; ++j) {
      const int row = crow(j), col = ccol(i);
      float *cp = C + (long)row * ldc + col;
      if (VEC && full) {
        float4_t v = acc[i][j];
        if (!CINIT) {
          const float4_t old = beta != 0.f ? *reinterpret_cast<const float4_t *>(cp)
                                           : (float4_t){0.f, 0.f, 0.f, 0.f};
          v = alpha * v + beta * old;
        }
        *reinterpret_cast<float4_t *>(cp) = v;
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (row < M && col + r < N) {
            float v = acc[i][j][r];
            if (!CINIT) v = alpha * v + (beta != 0.f ? beta * cp[r] : 0.f);
            cp[r] = v;
          }
      }
    }
}

template <typename T16, int BMv, int BNv, int NS, int OC, bool CI, bool V, int WMv = 64,
          int WNv = 64, bool E = true, int GMv = 0>
static void hgemm_launch(int M, int N, int K, float alpha, const T16 *A, long lda, const T16 *B,
                         long ldb, float beta, float *C, long ldc, hipStream_t s) {
  constexpr int NT = (BMv / WMv) * (BNv / WNv) * 64;
  const int tm = E ? (M + BMv - 1) / BMv : M / BMv, tn = E ? (N + BNv - 1) / BNv : N / BNv;
  if (tm <= 0 || tn <= 0) return;
  hipLaunchKernelGGL((hgemm_nt_kernel<T16, BMv, BNv, WMv, WNv, NS, OC, CI, V, E, GMv>),
                     dim3(tm * tn), dim3(NT), 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                     tm, tn);
}

// Large accumulate-into-C updates take 256 x 256 tiles over the whole tiles
// of C, walked in groups of 4 tile rows, and 128 x 128 tiles for the rest
// (FP8: four waves of 128 x 128 per tile, the accumulators in the AGPRs;
// 16-bit: 16 waves of 64 x 64) unless CHARMLU_LP_WIDE=0 (read at each call)
// selects one launch of 256 x 256 (16-bit) or 128 x 256 (FP8) tiles of
// 64 x 64 per wave walked in groups of 2.  Each element of C receives the
// same products in the same order from either, so the choice does not
// change C.
static bool lp_wide() {
  const char *e = getenv("CHARMLU_LP_WIDE");
  return !(e && *e == '0');
}
// 128 x 128 tiles, 2 stages, 4 workgroups per CU (measured at 65280^2 x 256
// BF16: 212 TFLOP/s against 188 with 4 stages and 2 per CU, 198 with 64 x
// 256 tiles; non-temporal C loads or stores were slower)
template <typename T16>
static void hgemm_launch2(bool cinit, bool vec, int M, int N, int K, float alpha, const T16 *A,
                          long lda, const T16 *B, long ldb, float beta, float *C, long ldc,
                          hipStream_t s) {
  // Smaller grids keep 128 x 128 tiles: K >= 1024 on 3 stages, 3
  // workgroups per CU (546 TFLOP/s at 64512^2 x 1024 BF16 against 526 with
  // 2 stages and 4 per CU, 522 with 4 and 2, before the bank-conflict-free
  // swizzle).
  // Large accumulate-into-C updates (the mixed LU's trailing update at
  // block 1024): bigger tiles read fewer operand bytes per flop.  At
  // 64512^2 x 1024 the BF16 update ran at 615 TFLOP/s standalone with
  // 256 x 256 tiles (4 stages, one workgroup of 16 waves per CU) against
  // 572 with 128 x 128 (3 stages, 3 per CU) and 607 with 128 x 256, and at
  // 507 against 433 inside the LU at n = 65536.  Only while the grid keeps
  // every CU busy with them (at least 2^25 C elements).
  const bool large = (long)M * N >= (1L << 25);
  if (cinit && vec && K >= 1024 && large && lp_wide()) {
    // whole 256 x 256 tiles walked in groups of 4 tile rows per XCD, then
    // the rows below and the columns beside them with 128 x 128 tiles (the
    // same per-element products in the same order).  Kernel time at
    // 53248^2 x 12288 BF16 (the update of a group of 12 steps of 1024):
    // 1010 TFLOP/s against 980 with groups of 2 tile rows over all tiles;
    // four waves of 128 x 128 instead of 16 of 64 x 64: 980 (4 stages) and
    // 968 (3 stages); at 61440^2 x 4096 928 against 886, at 64512^2 x 1024
    // 619 against 596.
    const int Mw = M / 256 * 256, Nw = N / 256 * 256;
    hgemm_launch<T16, 256, 256, 4, 1, true, true, 64, 64, false, 4>(Mw, Nw, K, alpha, A, lda, B,
                                                                    ldb, beta, C, ldc, s);
    if (M > Mw)
      hgemm_launch<T16, 128, 128, 3, 3, true, true>(M - Mw, N, K, alpha, A + (long)Mw * lda, lda,
                                                    B, ldb, beta, C + (long)Mw * ldc, ldc, s);
    if (N > Nw && Mw > 0)
      hgemm_launch<T16, 128, 128, 3, 3, true, true>(Mw, N - Nw, K, alpha, A, lda,
                                                    B + (long)Nw * ldb, ldb, beta, C + Nw, ldc, s);
    return;
  }
  if (cinit && vec && K >= 1024 && large) {
    hgemm_launch<T16, 256, 256, 4, 1, true, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
    return;
  }
  if (cinit && vec && K >= 1024) {
    hgemm_launch<T16, 128, 128, 3, 3, true, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
    return;
  }
  if (cinit && vec)
    hgemm_launch<T16, 128, 128, 2, 4, true, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  else if (cinit)
    hgemm_launch<T16, 128, 128, 2, 4, true, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  else if (vec)
    hgemm_launch<T16, 128, 128, 2, 4, false, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  else
    hgemm_launch<T16, 128, 128, 2, 4, false, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                                                    s);
}

template <typename T16>
int hgemm_nt(int M, int N, int K, float alpha, const T16 *A, long lda, const T16 *B, long ldb,
             float beta, float *C, long ldc, hipStream_t s) {
  if (M <= 0 || N <= 0) return CLU_OK;
  if (K % 32 != 0 || K <= 0 || lda % 8 != 0 || ldb % 8 != 0 || ((uintptr_t)A & 15) ||
      ((uintptr_t)B & 15))
    return -1;
  const bool vec = ldc % 4 == 0 && ((uintptr_t)C & 15) == 0;
  hgemm_launch2<T16>(alpha == 1.f && beta == 1.f, vec, M, N, K, alpha, A, lda, B, ldb, beta, C,
                     ldc, s);
  CLU_LAUNCHED("hgemm_nt_kernel");
  return CLU_OK;
}

// C = alpha A B^T + beta C (or C += A B^T with CINIT) for FP8 e4m3 operands
// quantized with power-of-two scales per 128-row tile (escA[g]: rows 128 g ..
// of A were multiplied by 2^escA[g] before rounding; escB likewise per 128
// rows of B, i.e. per 128 columns of C).  BM x BN C tile, one wave per
// WM x WN block of it (WM / 32 x WN / 32 blocks of 32 x 32), K in stages of
// 64 bytes per row staged as in hgemm_nt_kernel.  EDGE = false: every tile
// lies inside C (the caller launches only whole tiles).
template <int BM, int BN, int WM, int WN, int NST, int OCC, bool CINIT, bool VEC, bool EDGE,
          int GMT = 0>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64, OCC) void fgemm_nt_kernel(
    int M, int N, int K, float alpha, const uint8_t *__restrict__ A, long lda,
    const uint8_t *__restrict__ B, long ldb, float beta, float *__restrict__ C, long ldc,
    int tiles_m, int tiles_n, const int *__restrict__ escA, const int *__restrict__ escB, int ks,
    int nseg, long estride) {
  constexpr int NT = (BM / WM) * (BN / WN) * 64;
  constexpr int BK = 64;  // K per stage (bytes per row)
  constexpr int WGN = BN / WN;
  constexpr int BI = WN / 32, BJ = WM / 32;  // 32 x 32 blocks per wave (columns, rows)
  constexpr int HALF = BM * 64;
  constexpr int STAGE = (BM + BN) * 64;
  constexpr int CA = BM * 4 / NT, CB = BN * 4 / NT;
  static_assert(BM % WM == 0 && BN % WN == 0 && WM % 32 == 0 && WN % 32 == 0, "tile");
  static_assert(128 % WM == 0 && 128 % WN == 0, "a wave's rows and columns lie in one scale tile");
  static_assert(CA >= 1 && CB >= 1 && CA * NT == BM * 4 && CB * NT == BN * 4, "staging");
  __shared__ __attribute__((aligned(1024))) char smem[NST * STAGE];

  int tm, tn;  // XCD-aware tile order, as in hgemm_nt_kernel
  {
    const int bid = blockIdx.x, nblk = gridDim.x, per = nblk / 8;
    int lid = bid;
    if (bid < per * 8) lid = (bid % 8) * per + bid / 8;
    constexpr int GM = GMT > 0 ? GMT : 512 / BM;
    const int group_size = GM * tiles_n, group = lid / group_size, first_m = group * GM;
    const int gm = min(tiles_m - first_m, GM), in_group = lid - group * group_size;
    tm = first_m + in_group % gm;
    tn = in_group / gm;
  }
  const int m0 = tm * BM, n0 = tn * BN;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm0 = (wave / WGN) * WM, wn0 = (wave % WGN) * WN;
  const int lr = lane & 31, lh = lane >> 5;
  const bool full = !EDGE || (m0 + BM <= M && n0 + BN <= N);
  // E8M0 block scales of the two operands (2^(s - 127) undoes 2^esc); a
  // wave's rows and columns each lie in one 128-row scale tile.  K runs
  // over nseg segments of ks stages, products of their own scaling (up to
  // 16 updates of C summed in one pass over it), whose exponents lie
  // estride apart; the current segment's pair is reloaded at each boundary
  const int ta = (m0 + wm0) >> 7, tb = (n0 + wn0) >> 7;
  int sa = 127 - escA[ta], sb = 127 - escB[tb];
  int seg = 0, seg_end = ks;

  // The matrix-core op's first operand is the B tile (32 C columns), its
  // second the A tile (32 C rows): element r of block (i, j) is C[row][col]
  // with row = m0 + wm0 + 32 j + lr and col = n0 + wn0 + 32 i + 8 (r / 4) +
  // 4 lh + r % 4, so elements 4 s .. 4 s + 3 are 4 consecutive columns.
  float16_t acc[BI][BJ];
  auto crow = [&](int j) { return m0 + wm0 + 32 * j + lr; };
  auto ccol = [&](int i, int s4) { return n0 + wn0 + 32 * i + 8 * s4 + 4 * lh; };
#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int j = 0; j < BJ; ++j)
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        float4_t v = {0.f, 0.f, 0.f, 0.f};
        if (CINIT) {
          const int row = crow(j), col = ccol(i, s4);
          if (VEC && full) {
            v = *reinterpret_cast<const float4_t *>(C + (long)row * ldc + col);
          } else {
#pragma unroll
            for (int r = 0; r < 4; ++r)
              v[r] = (row < M && col + r < N) ? C[(long)row * ldc + col + r] : 0.f;
          }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[i][j][4 * s4 + r] = v[r];
      }

  // this thread's chunks of A and of B per stage (rows past M or N are
  // clamped to the last one; their products are never stored)
  const uint8_t *ga[CA], *gb[CB];
  int loffa[CA], loffb[CB];
#pragma unroll
  for (int r = 0; r < CA; ++r) {
    const int c = wave * 64 + NT * r + lane;
    const int row = c >> 2, kq = (c & 3) ^ lds_swz(row);
    ga[r] = A + (long)min(m0 + row, M - 1) * lda + 16 * kq;
    loffa[r] = (wave * 64 + NT * r) * 16;
  }
#pragma unroll
  for (int r = 0; r < CB; ++r) {
    const int c = wave * 64 + NT * r + lane;
    const int row = c >> 2, kq = (c & 3) ^ lds_swz(row);
    gb[r] = B + (long)min(n0 + row, N - 1) * ldb + 16 * kq;
    loffb[r] = HALF + (wave * 64 + NT * r) * 16;
  }
  const int nk = K / BK;
  auto issue = [&](int kc) {
    const int kk = kc * BK;
    char *st = smem + (kc % NST) * STAGE;
#pragma unroll
    for (int r = 0; r < CA; ++r)
      __builtin_amdgcn_global_load_lds((gbl_ptr_t)(ga[r] + kk), (lds_ptr_t)(st + loffa[r]), 16, 0,
                                       0);
#pragma unroll
    for (int r = 0; r < CB; ++r)
      __builtin_amdgcn_global_load_lds((gbl_ptr_t)(gb[r] + kk), (lds_ptr_t)(st + loffb[r]), 16, 0,
                                       0);
  };
  // a lane's operand: the two 16-byte chunks 2 lh and 2 lh + 1 of its row
  int aoff[BJ][2], boff[BI][2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
#pragma unroll
    for (int j = 0; j < BJ; ++j) {
      const int m = wm0 + 32 * j + lr;
      aoff[j][h] = m * 64 + (((2 * lh + h) ^ lds_swz(m)) << 4);
    }
#pragma unroll
    for (int i = 0; i < BI; ++i) {
      const int n = wn0 + 32 * i + lr;
      boff[i][h] = HALF + n * 64 + (((2 * lh + h) ^ lds_swz(n)) << 4);
    }
  }

  static_assert(NST >= 2 && NST <= 5 && (CA + CB) * (NST - 2) < 64, "stages");
#pragma unroll
  for (int st = 0; st < NST - 1; ++st)
    if (st < nk) issue(st);
  for (int kc = 0; kc < nk; ++kc) {
    const int ahead = nk - 1 - kc;
    if (NST >= 5 && ahead >= 3) wait_vm_lp<(CA + CB) * (NST >= 5 ? 3 : 0)>();
    else if (NST >= 4 && ahead >= 2) wait_vm_lp<(CA + CB) * (NST >= 4 ? 2 : 0)>();
    else if (NST >= 3 && ahead >= 1) wait_vm_lp<(CA + CB) * (NST >= 3 ? 1 : 0)>();
    else wait_vm_lp<0>();
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
    if (kc + NST - 1 < nk) issue(kc + NST - 1);
    const char *st = smem + (kc % NST) * STAGE;
    if (kc == seg_end && seg + 1 < nseg) {
      ++seg;
      seg_end += ks;
      sa = 127 - escA[seg * estride + ta];
      sb = 127 - escB[seg * estride + tb];
    }
    int8v_t a[BJ], b[BI];
#pragma unroll
    for (int j = 0; j < BJ; ++j) {
      const int4v_t a0 = *reinterpret_cast<const int4v_t *>(st + aoff[j][0]);
      const int4v_t a1 = *reinterpret_cast<const int4v_t *>(st + aoff[j][1]);
      a[j] = (int8v_t){a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
    }
#pragma unroll
    for (int i = 0; i < BI; ++i) {
      const int4v_t b0 = *reinterpret_cast<const int4v_t *>(st + boff[i][0]);
      const int4v_t b1 = *reinterpret_cast<const int4v_t *>(st + boff[i][1]);
      b[i] = (int8v_t){b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
    }
#pragma unroll
    for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < BJ; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(b[i], a[j], acc[i][j], 0, 0,
                                                                    0, sb, 0, sa);
  }
  wait_vm_lp<0>();  // no direct-to-LDS write may be in flight when the workgroup exits

#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int j = 0; j < BJ; ++j)
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        const int row = crow(j), col = ccol(i, s4);
        float *cp = C + (long)row * ldc + col;
        float4_t v = {acc[i][j][4 * s4], acc[i][j][4 * s4 + 1], acc[i][j][4 * s4 + 2],
                      acc[i][j][4 * s4 + 3]};
        if (VEC && full) {
          if (!CINIT) {
            const float4_t old = beta != 0.f ? *reinterpret_cast<const float4_t *>(cp)
                                             : (float4_t){0.f, 0.f, 0.f, 0.f};
            v = alpha * v + beta * old;
          }
          *reinterpret_cast<float4_t *>(cp) = v;
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (row < M && col + r < N) {
              float x = v[r];
              if (!CINIT) x = alpha * x + (beta != 0.f ? beta * cp[r] : 0.f);
              cp[r] = x;
            }
        }
      }
}

template <int BMv, int BNv, int WMv, int WNv, int NS, int OC, bool CI, bool V, bool E = true,
          int GMv = 0>
static void fgemm_launch(int M, int N, int K, float alpha, const uint8_t *A, long lda,
                         const uint8_t *B, long ldb, float beta, float *C, long ldc,
                         hipStream_t s, const int *escA, const int *escB, int ks, int nseg,
                         long estride) {
  constexpr int NT = (BMv / WMv) * (BNv / WNv) * 64;
  const int tm = E ? (M + BMv - 1) / BMv : M / BMv, tn = E ? (N + BNv - 1) / BNv : N / BNv;
  if (tm <= 0 || tn <= 0) return;
  hipLaunchKernelGGL((fgemm_nt_kernel<BMv, BNv, WMv, WNv, NS, OC, CI, V, E, GMv>), dim3(tm * tn),
                     dim3(NT), 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn, escA,
                     escB, ks, nseg, estride);
}


// FP8 operands with per-tile power-of-two scales.  The large accumulate-
// into-C updates (the mixed LU's trailing update) take the wide tiles on
// 3 stages below K = 4096 and 4 from there, walked in groups of 4 tile rows
// (1024 rows of A) per XCD; kernel time at 57344^2 x 16384 (the update of a
// group of 16 steps of 1024): 2007 TFLOP/s against 1770 with 128 x 256
// tiles of 64 x 64 per wave (3 stages, 2 workgroups per CU), 1984 with 3
// stages, 1937 with groups of 2 tile rows; at 61440^2 x 4096 1615 against
// 1512; at 64512^2 x 1024 842 against 833 (4 stages: 763).  The other
// calls take 128 x 128 tiles on 2 stages with 4 per CU.  Before the
// block-scaled instruction: 749 TFLOP/s at 64512^2 x 1024 with two
// v_mfma_f32_16x16x32_fp8_fp8 per 16-byte operand read and the scales
// applied to C before and after the K loop; persistent workgroups, two per
// CU with or without a start offset between them, measured slower (678-689
// against 781 for 128 x 256 tiles at the time).
int fgemm_nt(int M, int N, int K, float alpha, const uint8_t *A, long lda, const int *escA,
             const uint8_t *B, long ldb, const int *escB, float beta, float *C, long ldc,
             hipStream_t s, int kseg, int nseg, long estride) {
  if (M <= 0 || N <= 0) return CLU_OK;
  if (K % 64 != 0 || K <= 0 || lda % 16 != 0 || ldb % 16 != 0 || ((uintptr_t)A & 15) ||
      ((uintptr_t)B & 15) || !escA || !escB)
    return -1;
  // nseg segments of kseg K indices each, with their own scales
  if (nseg < 1 || nseg > 16 || (nseg > 1 && (kseg <= 0 || kseg % 64 != 0 || kseg * nseg != K)))
    return -1;
  const int ks = nseg > 1 ? kseg / 64 : K / 64;
  const bool vec = ldc % 4 == 0 && ((uintptr_t)C & 15) == 0;
  const bool cinit = alpha == 1.f && beta == 1.f;
  const bool large = (long)M * N >= (1L << 25);
  if (cinit && vec && K >= 1024 && large && lp_wide()) {
    // whole 256 x 256 tiles, then the rows below them and the columns beside
    // them with 128 x 128 tiles (an element's products are accumulated in
    // the same order by either kernel, so the split does not change C);
    // the scale tiles of the edges start 2 Mw / 256 (2 Nw / 256) tiles in
    const int Mw = M / 256 * 256, Nw = N / 256 * 256;
    if (K >= 4096)
      fgemm_launch<256, 256, 128, 128, 4, 1, true, true, false, 4>(
          Mw, Nw, K, alpha, A, lda, B, ldb, beta, C, ldc, s, escA, escB, ks, nseg, estride);
    else
      fgemm_launch<256, 256, 128, 128, 3, 1, true, true, false, 4>(
          Mw, Nw, K, alpha, A, lda, B, ldb, beta, C, ldc, s, escA, escB, ks, nseg, estride);
    if (M > Mw)
      fgemm_launch<128, 128, 64, 64, 2, 4, true, true>(
          M - Mw, N, K, alpha, A + (long)Mw * lda, lda, B, ldb, beta, C + (long)Mw * ldc, ldc, s,
          escA + Mw / 128, escB, ks, nseg, estride);
    if (N > Nw && Mw > 0)
      fgemm_launch<128, 128, 64, 64, 2, 4, true, true>(
          Mw, N - Nw, K, alpha, A, lda, B + (long)Nw * ldb, ldb, beta, C + Nw, ldc, s, escA,
          escB + Nw / 128, ks, nseg, estride);
  } else if (cinit && vec && K >= 1024 && large)
    fgemm_launch<128, 256, 64, 64, 3, 4, true, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                                                     s, escA, escB, ks, nseg, estride);
  else if (cinit && vec)
    fgemm_launch<128, 128, 64, 64, 2, 4, true, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                                                     s, escA, escB, ks, nseg, estride);
  else if (cinit)
    fgemm_launch<128, 128, 64, 64, 2, 4, true, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                                                      s, escA, escB, ks, nseg, estride);
  else if (vec)
    fgemm_launch<128, 128, 64, 64, 2, 4, false, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                                                      s, escA, escB, ks, nseg, estride);
  else
    fgemm_launch<128, 128, 64, 64, 2, 4, false, false>(M, N, K, alpha, A, lda, B, ldb, beta, C,
                                                       ldc, s, escA, escB, ks, nseg, estride);
  CLU_LAUNCHED("fgemm_nt_kernel");
  return CLU_OK;
}

// ---- FP8 quantization with per-tile power-of-two scales --------------------
// A tile's scale 2^e is the largest power of two with 2^e max|v| <= 448 (the
// largest finite e4m3 value), so no value saturates; v 2^e is rounded to
// nearest-even e4m3 by the hardware conversion.

CLU_DEV float block_max_256(float v, float *red) {
  for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  const int t = threadIdx.x;
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  const float m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  __syncthreads();
  return m;
}

CLU_DEV int tile_exponent(float mx) {
  if (!(mx > 0.f) || !isfinite(mx)) return 0;
  int e = (int)floorf(log2f(448.f / mx));
  e = max(-120, min(120, e));
  while (e > -120 && ldexpf(mx, e) > 448.f) --e;  // guard log2f rounding
  return e;
}

CLU_DEV unsigned quant4(float a, float b, float c, float d) {
  auto cl = [](float x) { return fminf(448.f, fmaxf(-448.f, x)); };
  int w = __builtin_amdgcn_cvt_pk_fp8_f32(cl(a), cl(b), 0, false);
  w = __builtin_amdgcn_cvt_pk_fp8_f32(cl(c), cl(d), w, true);
  return (unsigned)w;
}

// Each 128-row (128-column) scale tile is split over QSPLIT workgroups: a
// first kernel writes the tile's QSPLIT partial maxima of |v|, a second
// converts, every workgroup of a tile deriving the same exponent from the
// same partials (the maximum does not depend on the order), so the scales
// and bytes equal those of one workgroup per tile.  part holds QSPLIT
// floats per tile.  (One workgroup per tile, reading its tile twice,
// streamed 0.8 TB/s at n = 65536: 45 ms of each FP8 LU.)
constexpr int QSPLIT = FP8_PART_PER_TILE;

// part[g * QSPLIT + s] = max |src| over rows 128 g + 16 s .. + 15 (K columns)
__global__ __launch_bounds__(256) void absmax_rows_kernel(int M, int K,
                                                          const float *__restrict__ src, long lds,
                                                          float *__restrict__ part, int vec) {
  __shared__ float red[4];
  const int g = blockIdx.x, sl = blockIdx.y, t = threadIdx.x;
  const int r0 = g * 128 + sl * (128 / QSPLIT);
  const int rows = max(0, min(128 / QSPLIT, M - r0));
  float mx = 0.f;
  for (int i = 0; i < rows; ++i) {
    const float *sp = src + (long)(r0 + i) * lds;
    for (int k = 4 * t; k < K; k += 1024) {
      if (vec && k + 3 < K) {
        const float4_t f = *reinterpret_cast<const float4_t *>(sp + k);
        mx = fmaxf(mx, fmaxf(fmaxf(fabsf(f[0]), fabsf(f[1])), fmaxf(fabsf(f[2]), fabsf(f[3]))));
      } else {
        for (int q = 0; q < 4 && k + q < K; ++q) mx = fmaxf(mx, fabsf(sp[k + q]));
      }
    }
  }
  mx = block_max_256(mx, red);
  if (t == 0) part[g * QSPLIT + sl] = mx;
}

// part[g * QSPLIT + s] = max |src| over columns 128 g .. + 127 of rows
// s, s + QSPLIT, s + 2 QSPLIT, ... (< K)
__global__ __launch_bounds__(256) void absmax_cols_kernel(int K, int N,
                                                          const float *__restrict__ src, long lds,
                                                          float *__restrict__ part) {
  __shared__ float red[4];
  const int g = blockIdx.x, sl = blockIdx.y, t = threadIdx.x;
  const int c = g * 128 + (t & 127);
  float mx = 0.f;
  if (c < N)
    for (int k = sl + QSPLIT * (t >> 7); k < K; k += 2 * QSPLIT)
      mx = fmaxf(mx, fabsf(src[(long)k * lds + c]));
  mx = block_max_256(mx, red);
  if (t == 0) part[g * QSPLIT + sl] = mx;
}

CLU_DEV int tile_exponent_of(const float *part, int g) {
  float mx = part[g * QSPLIT];
#pragma unroll
  for (int q = 1; q < QSPLIT; ++q) mx = fmaxf(mx, part[g * QSPLIT + q]);
  return tile_exponent(mx);
}

// dst (M x Kp bytes, rows ldd apart) = fp8(sign * src * 2^e_g) for rows
// 128 g + 16 s .. + 15 of 128-row tile g, zero for k >= K; esc[g] = e_g
__global__ __launch_bounds__(256) void quant_rows_fp8_kernel(int M, int K, int Kp, float sign,
                                                             const float *__restrict__ src,
                                                             long lds, uint8_t *__restrict__ dst,
                                                             const float *__restrict__ part,
                                                             int *__restrict__ esc, int vec,
                                                             float *__restrict__ cpy, long ldc,
                                                             long ldd) {
  const int g = blockIdx.x, sl = blockIdx.y, t = threadIdx.x;
  const int r0 = g * 128 + sl * (128 / QSPLIT);
  const int rows = max(0, min(128 / QSPLIT, M - r0));
  const int ex = tile_exponent_of(part, g);
  if (sl == 0 && t == 0) esc[g] = ex;
  for (int i = 0; i < rows; ++i) {
    const float *sp = src + (long)(r0 + i) * lds;
    for (int k = 4 * t; k < Kp; k += 1024) {
      float v[4];
      if (vec && k + 3 < K) {
        const float4_t f = *reinterpret_cast<const float4_t *>(sp + k);
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = f[q];
        if (cpy) *reinterpret_cast<float4_t *>(cpy + (long)(r0 + i) * ldc + k) = f;
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = k + q < K ? sp[k + q] : 0.f;
        if (cpy)
          for (int q = 0; q < 4 && k + q < K; ++q) cpy[(long)(r0 + i) * ldc + k + q] = v[q];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = k + q < K ? sign * ldexpf(v[q], ex) : 0.f;
      *reinterpret_cast<unsigned *>(dst + (long)(r0 + i) * ldd + k) =
          quant4(v[0], v[1], v[2], v[3]);
    }
  }
}

// dst (N x Kp bytes, rows ldd apart) = fp8(src^T * 2^e_g) for the columns of
// 128-column tile g of src (K x N), zero for k >= K; esc[g] = e_g.  Workgroup (g, j)
// transposes the 64-row slab j through LDS.
__global__ __launch_bounds__(256) void quant_cols_fp8_kernel(int K, int N, int Kp,
                                                             const float *__restrict__ src,
                                                             long lds, uint8_t *__restrict__ dst,
                                                             const float *__restrict__ part,
                                                             int *__restrict__ esc,
                                                             float *__restrict__ cpy, long ldc,
                                                             long ldd) {
  __shared__ float tile[64][129];
  const int g = blockIdx.x, t = threadIdx.x, c0 = g * 128, k0 = blockIdx.y * 64;
  const int cols = min(128, N - c0);
  const int ex = tile_exponent_of(part, g);
  if (blockIdx.y == 0 && t == 0) esc[g] = ex;
  const int c = t & 127, h = t >> 7;  // column, half of the 64-row slab
  for (int e = t; e < 64 * 128; e += 256) {
    const int kk = e >> 7, cc = e & 127;
    const bool in = k0 + kk < K && cc < cols;
    const float v = in ? src[(long)(k0 + kk) * lds + c0 + cc] : 0.f;
    tile[kk][cc] = v;
    if (cpy && in) cpy[(long)(k0 + kk) * ldc + c0 + cc] = v;
  }
  __syncthreads();
  if (c < cols) {
    unsigned w[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int kk = 32 * h + 4 * q;
      w[q] = quant4(ldexpf(tile[kk][c], ex), ldexpf(tile[kk + 1][c], ex),
                    ldexpf(tile[kk + 2][c], ex), ldexpf(tile[kk + 3][c], ex));
    }
    uint4 *dp = reinterpret_cast<uint4 *>(dst + (long)(c0 + c) * ldd + k0 + 32 * h);
    dp[0] = make_uint4(w[0], w[1], w[2], w[3]);
    dp[1] = make_uint4(w[4], w[5], w[6], w[7]);
  }
}

int quant_rows_fp8(int M, int K, int Kp, float sign, const float *src, long lds, uint8_t *dst,
                   int *esc, float *part, hipStream_t s, float *cpy, long ldc, long ldd) {
  if (M <= 0) return CLU_OK;
  if (ldd == 0) ldd = Kp;
  if (Kp % 64 != 0 || Kp < K || ldd < Kp || ldd % 4 != 0 || ((uintptr_t)dst & 3)) return -1;
  // 16-byte source reads (and copy writes) when both are 16-byte aligned
  const int vec = lds % 4 == 0 && ((uintptr_t)src & 15) == 0 &&
                  (!cpy || (ldc % 4 == 0 && ((uintptr_t)cpy & 15) == 0));
  const dim3 grid((M + 127) / 128, QSPLIT);
  hipLaunchKernelGGL(absmax_rows_kernel, grid, dim3(256), 0, s, M, K, src, lds, part, vec);
  CLU_LAUNCHED("absmax_rows_kernel");
  hipLaunchKernelGGL(quant_rows_fp8_kernel, grid, dim3(256), 0, s, M, K, Kp, sign, src, lds, dst,
                     part, esc, vec, cpy, ldc, ldd);
  CLU_LAUNCHED("quant_rows_fp8_kernel");
  return CLU_OK;
}
int quant_cols_fp8(int K, int N, int Kp, const float *src, long lds, uint8_t *dst, int *esc,
                   float *part, hipStream_t s, float *cpy, long ldc, long ldd) {
  if (N <= 0) return CLU_OK;
  if (ldd == 0) ldd = Kp;
  if (Kp % 64 != 0 || Kp < K || ldd < Kp || ldd % 16 != 0 || ((uintptr_t)dst & 15)) return -1;
  const int tiles = (N + 127) / 128;
  hipLaunchKernelGGL(absmax_cols_kernel, dim3(tiles, QSPLIT), dim3(256), 0, s, K, N, src, lds,
                     part);
  CLU_LAUNCHED("absmax_cols_kernel");
  hipLaunchKernelGGL(quant_cols_fp8_kernel, dim3(tiles, Kp / 64), dim3(256), 0, s, K, N, Kp, src,
                     lds, dst, part, esc, cpy, ldc, ldd);
  CLU_LAUNCHED("quant_cols_fp8_kernel");
  return CLU_OK;
}

template int hgemm_nt<__bf16>(int, int, int, float, const __bf16 *, long, const __bf16 *, long,
                              float, float *, long, hipStream_t);
template int hgemm_nt<_Float16>(int, int, int, float, const _Float16 *, long, const _Float16 *,
                                long, float, float *, long, hipStream_t);

// dst (M x Kp, rows ldd apart) = scale * src (M x K) rounded to 16 bits, zero
// for k >= K (scale = -1 folds the update's minus sign into the operand)
template <typename T16>
__global__ __launch_bounds__(256) void pack_rows_kernel(int M, int K, int Kp, float scale,
                                                        const float *__restrict__ src, long lds,
                                                        T16 *__restrict__ dst, int vec,
                                                        float *__restrict__ cpy, long ldc,
                                                        long ldd) {
  // 4 consecutive k per thread (one 8-byte store; a 16-byte load when the
  // source rows are 16-byte aligned), 1024 k per workgroup, rows strided
  // over grid rows: no index division per element
  const int k0 = (blockIdx.x * 256 + threadIdx.x) * 4;
  if (k0 >= Kp) return;
  for (int i = blockIdx.y; i < M; i += gridDim.y) {
    const float *sr = src + (long)i * lds;
    float v[4];
    if (vec && k0 + 3 < K) {
      const float4_t f = *reinterpret_cast<const float4_t *>(sr + k0);
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = f[q];
      if (cpy) *reinterpret_cast<float4_t *>(cpy + (long)i * ldc + k0) = f;
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = k0 + q < K ? sr[k0 + q] : 0.f;
      if (cpy)
        for (int q = 0; q < 4 && k0 + q < K; ++q) cpy[(long)i * ldc + k0 + q] = v[q];
    }
    union { T16 h[4]; uint2 u; } o;
#pragma unroll
    for (int q = 0; q < 4; ++q) o.h[q] = Lp<T16>::from(scale * v[q]);
    *reinterpret_cast<uint2 *>(dst + (long)i * ldd + k0) = o.u;
  }
}

// dst (N x Kp, rows ldd apart) = src (K x N)^T rounded to 16 bits, zero for
// k >= K; 32 x 32 tiles through LDS so both sides are read and written along
// rows
template <typename T16>
__global__ __launch_bounds__(256) void pack_cols_kernel(int K, int N, int Kp,
                                                        const float *__restrict__ src, long lds,
                                                        T16 *__restrict__ dst,
                                                        float *__restrict__ cpy, long ldc,
                                                        long ldd) {
  __shared__ float tile[32][33];
  const int k0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int k = k0 + r, n = n0 + tx;
    const bool in = k < K && n < N;
    const float v = in ? src[(long)k * lds + n] : 0.f;
    tile[r][tx] = v;
    if (cpy && in) cpy[(long)k * ldc + n] = v;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int n = n0 + r, k = k0 + tx;
    if (n < N && k < Kp) dst[(long)n * ldd + k] = Lp<T16>::from(tile[tx][r]);
  }
}

template <typename T16>
int pack_rows(int M, int K, int Kp, float scale, const float *src, long lds, T16 *dst,
              hipStream_t s, float *cpy, long ldc, long ldd) {
  if (M <= 0 || Kp <= 0) return CLU_OK;
  if (ldd == 0) ldd = Kp;
  if (Kp % 4 != 0 || ldd < Kp || ldd % 4 != 0 || ((uintptr_t)dst & 7)) return -1;
  // 16-byte source reads (and copy writes) when both are 16-byte aligned
  const int vec = lds % 4 == 0 && ((uintptr_t)src & 15) == 0 &&
                  (!cpy || (ldc % 4 == 0 && ((uintptr_t)cpy & 15) == 0));
  const dim3 grid((Kp + 1023) / 1024, M < 8192 ? M : 8192);
  hipLaunchKernelGGL((pack_rows_kernel<T16>), grid, dim3(256), 0, s, M, K, Kp, scale, src, lds,
                     dst, vec, cpy, ldc, ldd);
  CLU_LAUNCHED("pack_rows_kernel");
  return CLU_OK;
}
template <typename T16>
int pack_cols(int K, int N, int Kp, const float *src, long lds, T16 *dst, hipStream_t s,
              float *cpy, long ldc, long ldd) {
  if (N <= 0 || Kp <= 0) return CLU_OK;
  if (ldd == 0) ldd = Kp;
  if (ldd < Kp) return -1;
  dim3 grid((N + 31) / 32, (Kp + 31) / 32);
  hipLaunchKernelGGL((pack_cols_kernel<T16>), grid, dim3(256), 0, s, K, N, Kp, src, lds, dst, cpy,
                     ldc, ldd);
  CLU_LAUNCHED("pack_cols_kernel");
  return CLU_OK;
}
template int pack_rows<__bf16>(int, int, int, float, const float *, long, __bf16 *, hipStream_t,
                               float *, long, long);
template int pack_rows<_Float16>(int, int, int, float, const float *, long, _Float16 *,
                                 hipStream_t, float *, long, long);
template int pack_cols<__bf16>(int, int, int, const float *, long, __bf16 *, hipStream_t, float *,
                               long, long);
template int pack_cols<_Float16>(int, int, int, const float *, long, _Float16 *, hipStream_t,
                                 float *, long, long);

}  // namespace clu
