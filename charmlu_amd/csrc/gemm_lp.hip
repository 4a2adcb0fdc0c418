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
#include <hip/hip_bf16.h>

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
// FP8 e4m3 (OCP, largest finite 448): a lane's 16-byte LDS read holds 16 k
// values, consumed by two v_mfma_f32_16x16x32_fp8_fp8 (8 bytes each; lane
// layout as for the 16-bit types, measured on gfx950)
typedef long long2_t __attribute__((ext_vector_type(2)));
template <> struct Lp<uint8_t> {
  typedef long2_t vec_t;
  static CLU_DEV float4_t mfma(vec_t a, vec_t b, float4_t c) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a.x, b.x, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a.y, b.y, c, 0, 0, 0);
  }
};

template <int N>
CLU_DEV void wait_vm_lp() {
  static_assert(N >= 0 && N < 16, "vmcnt");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | (0x7 << 4) | (0xF << 8));
}

template <typename T16, int BM, int BN, int NST, int OCC, bool CINIT, bool VEC>
__global__ __launch_bounds__((BM / 64) * (BN / 64) * 64, OCC) void hgemm_nt_kernel(
    int M, int N, int K, float alpha, const T16 *__restrict__ A, long lda,
    const T16 *__restrict__ B, long ldb, float beta, float *__restrict__ C, long ldc,
    int tiles_m, int tiles_n, const int *__restrict__ escA, const int *__restrict__ escB) {
  // BM x BN C tile, one wave per 64 x 64 block of it (BM / 64 x BN / 64
  // waves).  A stage holds 64 bytes of K per row: 32 16-bit or 64 FP8 values.
  // FP8 operands carry power-of-two scales per 128-row tile (escA: A * 2^e
  // was quantized; escB likewise per 128 rows of B, i.e. columns of C), so
  // a wave's products are 2^(escA[..] + escB[..]) too large.
  constexpr int NT = (BM / 64) * (BN / 64) * 64;  // threads
  constexpr int EPC = 16 / sizeof(T16);  // elements per 16-byte chunk
  constexpr int BK = 4 * EPC;
  constexpr int WGN = BN / 64;
  constexpr int HALF = BM * 64;          // bytes of the A operand's stage
  constexpr int STAGE = (BM + BN) * 64;
  constexpr int CA = BM * 4 / NT, CB = BN * 4 / NT;  // 16-byte chunks per thread
  static_assert(BM % 64 == 0 && BN % 64 == 0 && CA >= 1 && CB >= 1, "tile");
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
    constexpr int GM = 512 / BM;
    const int group_size = GM * tiles_n, group = lid / group_size, first_m = group * GM;
    const int gm = min(tiles_m - first_m, GM), in_group = lid - group * group_size;
    tm = first_m + in_group % gm;
    tn = in_group / gm;
  }
  const int m0 = tm * BM, n0 = tn * BN;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm0 = (wave / WGN) * 64, wn0 = (wave % WGN) * 64;
  const int lr = lane & 15, lq = lane >> 4;
  const bool full = m0 + BM <= M && n0 + BN <= N;
  // products of this wave's block carry 2^escale (FP8 only)
  const int escale = escA ? escA[(m0 + wm0) >> 7] + escB[(n0 + wn0) >> 7] : 0;

  // The MFMA's first operand is the B tile (16 C columns), its second the A
  // tile (16 C rows), so D[r][c] of block (i, j) is C[row][col] with
  // row = m0 + wm0 + 16 j + lr and col = n0 + wn0 + 16 i + 4 lq + r: each
  // lane owns 4 consecutive columns of one row (one 16-byte access).
  float4_t acc[4][4];
  auto crow = [&](int j) { return m0 + wm0 + 16 * j + lr; };
  auto ccol = [&](int i) { return n0 + wn0 + 16 * i + 4 * lq; };
  if (CINIT) {
    // C += A B^T: C itself is the initial accumulator, requested before the
    // first operand stage so its latency overlaps the K loop
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = crow(j), col = ccol(i);
        if (VEC && full) {
          const float4_t *cp4 = reinterpret_cast<const float4_t *>(C + (long)row * ldc + col);
          acc[i][j] = *cp4;
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            acc[i][j][r] = (row < M && col + r < N) ? C[(long)row * ldc + col + r] : 0.f;
        }
        if (escA) {
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[i][j][r] = ldexpf(acc[i][j][r], escale);
        }
      }
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = (float4_t){0.f, 0.f, 0.f, 0.f};
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
  int aoff[4], boff[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = wm0 + 16 * i + lr, n = wn0 + 16 * i + lr;
    aoff[i] = m * 64 + ((lq ^ lds_swz(m)) << 4);
    boff[i] = HALF + n * 64 + ((lq ^ lds_swz(n)) << 4);
  }

  static_assert(NST >= 2 && NST <= 5, "stages");
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
    vec_t a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = *reinterpret_cast<const vec_t *>(st + aoff[i]);
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = *reinterpret_cast<const vec_t *>(st + boff[j]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = Lp<T16>::mfma(b[i], a[j], acc[i][j]);
  }
  wait_vm_lp<0>();  // no direct-to-LDS write may be in flight when the workgroup exits

  // epilogue: CINIT stores the accumulators; otherwise C = alpha acc + beta C
  if (escA) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[i][j][r] = ldexpf(acc[i][j][r], -escale);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
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

template <typename T16, int BMv, int BNv, int NS, int OC, bool CI, bool V>
static void hgemm_launch(int M, int N, int K, float alpha, const T16 *A, long lda, const T16 *B,
                         long ldb, float beta, float *C, long ldc, hipStream_t s,
                         const int *escA = nullptr, const int *escB = nullptr) {
  constexpr int NT = (BMv / 64) * (BNv / 64) * 64;
  const int tm = (M + BMv - 1) / BMv, tn = (N + BNv - 1) / BNv;
  hipLaunchKernelGGL((hgemm_nt_kernel<T16, BMv, BNv, NS, OC, CI, V>), dim3(tm * tn), dim3(NT), 0,
                     s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn, escA, escB);
}
// 128 x 128 tiles, 2 stages, 4 workgroups per CU (measured at 65280^2 x 256
// BF16: 212 TFLOP/s against 188 with 4 stages and 2 per CU, 198 with 64 x
// 256 tiles; non-temporal C loads or stores were slower)
template <typename T16>
static void hgemm_launch2(bool cinit, bool vec, int M, int N, int K, float alpha, const T16 *A,
                          long lda, const T16 *B, long ldb, float beta, float *C, long ldc,
                          hipStream_t s, const int *escA = nullptr, const int *escB = nullptr) {
  // Smaller grids keep 128 x 128 tiles: 16-bit operands with K >= 1024 on
  // 3 stages, 3 workgroups per CU (546 TFLOP/s at 64512^2 x 1024 BF16
  // against 526 with 2 stages and 4 per CU, 522 with 4 and 2, before the
  // bank-conflict-free swizzle); FP8 on 2 stages and 4 per CU (its update at
  // n = 131072 took 3.5 s with 3 and 3 against 2.7 s).
  // Large accumulate-into-C updates (the mixed LU's trailing update at
  // block 1024): bigger tiles read fewer operand bytes per flop.  At
  // 64512^2 x 1024 the BF16 update ran at 615 TFLOP/s standalone with
  // 256 x 256 tiles (4 stages, one workgroup of 16 waves per CU) against
  // 572 with 128 x 128 (3 stages, 3 per CU) and 607 with 128 x 256, and at
  // 507 against 433 inside the LU at n = 65536; FP8 at 749 with 128 x 256
  // tiles (3 stages, 2 per CU) against 695 with 128 x 128 (2 stages, 4 per
  // CU) and 668 with 256 x 256 (639 against 562 inside the LU).  Only while
  // the grid keeps every CU busy with them (at least 2^25 C elements).
  const bool large = (long)M * N >= (1L << 25);
  if (cinit && vec && K >= 1024 && large) {
    if constexpr (sizeof(T16) == 2)
      hgemm_launch<T16, 256, 256, 4, 1, true, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                                                    s, escA, escB);
    else
      hgemm_launch<T16, 128, 256, 3, 2, true, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                                                    s, escA, escB);
    return;
  }
  if (cinit && vec && K >= 1024 && sizeof(T16) == 2) {
    hgemm_launch<T16, 128, 128, 3, 3, true, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s,
                                             escA, escB);
    return;
  }
  if (cinit && vec)
    hgemm_launch<T16, 128, 128, 2, 4, true, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s,
                                             escA, escB);
  else if (cinit)
    hgemm_launch<T16, 128, 128, 2, 4, true, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s,
                                              escA, escB);
  else if (vec)
    hgemm_launch<T16, 128, 128, 2, 4, false, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s,
                                              escA, escB);
  else
    hgemm_launch<T16, 128, 128, 2, 4, false, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s,
                                               escA, escB);
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

// FP8 operands with per-tile power-of-two scales (128 x 128 tiles only, so a
// C tile sees one scale of A and one of B)
int fgemm_nt(int M, int N, int K, float alpha, const uint8_t *A, long lda, const int *escA,
             const uint8_t *B, long ldb, const int *escB, float beta, float *C, long ldc,
             hipStream_t s) {
  if (M <= 0 || N <= 0) return CLU_OK;
  if (K % 64 != 0 || K <= 0 || lda % 16 != 0 || ldb % 16 != 0 || ((uintptr_t)A & 15) ||
      ((uintptr_t)B & 15) || !escA || !escB)
    return -1;
  const bool vec = ldc % 4 == 0 && ((uintptr_t)C & 15) == 0;
  hgemm_launch2<uint8_t>(alpha == 1.f && beta == 1.f, vec, M, N, K, alpha, A, lda, B, ldb, beta,
                         C, ldc, s, escA, escB);
  CLU_LAUNCHED("hgemm_nt_kernel (fp8)");
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

// dst (M x Kp bytes) = fp8(sign * src * 2^e_g) for the rows of 128-row tile
// g, zero for k >= K; esc[g] = e_g.  One workgroup per tile.
__global__ __launch_bounds__(256) void quant_rows_fp8_kernel(int M, int K, int Kp, float sign,
                                                             const float *__restrict__ src,
                                                             long lds, uint8_t *__restrict__ dst,
                                                             int *__restrict__ esc, int vec) {
  // rows one after another, 4 consecutive k per thread (a 16-byte load when
  // the source rows are 16-byte aligned): no index division per element
  __shared__ float red[4];
  const int g = blockIdx.x, t = threadIdx.x, r0 = g * 128;
  const int rows = min(128, M - r0);
  auto load4 = [&](const float *sp, int k, float (&v)[4]) {
    if (vec && k + 3 < K) {
      const float4_t f = *reinterpret_cast<const float4_t *>(sp + k);
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = f[q];
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = k + q < K ? sp[k + q] : 0.f;
    }
  };
  float mx = 0.f;
  for (int i = 0; i < rows; ++i) {
    const float *sp = src + (long)(r0 + i) * lds;
    for (int k = 4 * t; k < K; k += 1024) {
      float v[4];
      load4(sp, k, v);
#pragma unroll
      for (int q = 0; q < 4; ++q) mx = fmaxf(mx, fabsf(v[q]));
    }
  }
  mx = block_max_256(mx, red);
  const int ex = tile_exponent(mx);
  if (t == 0) esc[g] = ex;
  for (int i = 0; i < rows; ++i) {
    const float *sp = src + (long)(r0 + i) * lds;
    for (int k = 4 * t; k < Kp; k += 1024) {
      float v[4];
      load4(sp, k, v);
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = k + q < K ? sign * ldexpf(v[q], ex) : 0.f;
      *reinterpret_cast<unsigned *>(dst + (long)(r0 + i) * Kp + k) =
          quant4(v[0], v[1], v[2], v[3]);
    }
  }
}

// dst (N x Kp bytes) = fp8(src^T * 2^e_g) for the columns of 128-column tile
// g of src (K x N), zero for k >= K; esc[g] = e_g.  One workgroup per tile,
// K in slabs of 64 transposed through LDS.
__global__ __launch_bounds__(256) void quant_cols_fp8_kernel(int K, int N, int Kp,
                                                             const float *__restrict__ src,
                                                             long lds, uint8_t *__restrict__ dst,
                                                             int *__restrict__ esc) {
  __shared__ float red[4];
  __shared__ float tile[64][129];
  const int g = blockIdx.x, t = threadIdx.x, c0 = g * 128;
  const int cols = min(128, N - c0);
  float mx = 0.f;
  for (int e = t; e < K * 128; e += 256) {
    const int k = e / 128, c = e % 128;
    if (c < cols) mx = fmaxf(mx, fabsf(src[(long)k * lds + c0 + c]));
  }
  mx = block_max_256(mx, red);
  const int ex = tile_exponent(mx);
  if (t == 0) esc[g] = ex;
  const int c = t & 127, h = t >> 7;  // column, half of the 64-row slab
  for (int k0 = 0; k0 < Kp; k0 += 64) {
    for (int e = t; e < 64 * 128; e += 256) {
      const int kk = e / 128, cc = e % 128;
      tile[kk][cc] = (k0 + kk < K && cc < cols) ? src[(long)(k0 + kk) * lds + c0 + cc] : 0.f;
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
      uint4 *dp = reinterpret_cast<uint4 *>(dst + (long)(c0 + c) * Kp + k0 + 32 * h);
      dp[0] = make_uint4(w[0], w[1], w[2], w[3]);
      dp[1] = make_uint4(w[4], w[5], w[6], w[7]);
    }
    __syncthreads();
  }
}

int quant_rows_fp8(int M, int K, int Kp, float sign, const float *src, long lds, uint8_t *dst,
                   int *esc, hipStream_t s) {
  if (M <= 0) return CLU_OK;
  if (Kp % 64 != 0 || Kp < K) return -1;
  const int vec = lds % 4 == 0 && ((uintptr_t)src & 15) == 0;
  hipLaunchKernelGGL(quant_rows_fp8_kernel, dim3((M + 127) / 128), dim3(256), 0, s, M, K, Kp,
                     sign, src, lds, dst, esc, vec);
  CLU_LAUNCHED("quant_rows_fp8_kernel");
  return CLU_OK;
}
int quant_cols_fp8(int K, int N, int Kp, const float *src, long lds, uint8_t *dst, int *esc,
                   hipStream_t s) {
  if (N <= 0) return CLU_OK;
  if (Kp % 64 != 0 || Kp < K) return -1;
  hipLaunchKernelGGL(quant_cols_fp8_kernel, dim3((N + 127) / 128), dim3(256), 0, s, K, N, Kp, src,
                     lds, dst, esc);
  CLU_LAUNCHED("quant_cols_fp8_kernel");
  return CLU_OK;
}

template int hgemm_nt<__bf16>(int, int, int, float, const __bf16 *, long, const __bf16 *, long,
                              float, float *, long, hipStream_t);
template int hgemm_nt<_Float16>(int, int, int, float, const _Float16 *, long, const _Float16 *,
                                long, float, float *, long, hipStream_t);

// dst (M x Kp, row-major) = scale * src (M x K) rounded to 16 bits, zero for
// k >= K (scale = -1 folds the update's minus sign into the operand)
template <typename T16>
__global__ __launch_bounds__(256) void pack_rows_kernel(int M, int K, int Kp, float scale,
                                                        const float *__restrict__ src, long lds,
                                                        T16 *__restrict__ dst, int vec) {
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
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = k0 + q < K ? sr[k0 + q] : 0.f;
    }
    union { T16 h[4]; uint2 u; } o;
#pragma unroll
    for (int q = 0; q < 4; ++q) o.h[q] = Lp<T16>::from(scale * v[q]);
    *reinterpret_cast<uint2 *>(dst + (long)i * Kp + k0) = o.u;
  }
}

// dst (N x Kp, row-major) = src (K x N)^T rounded to 16 bits, zero for k >= K;
// 32 x 32 tiles through LDS so both sides are read and written along rows
template <typename T16>
__global__ __launch_bounds__(256) void pack_cols_kernel(int K, int N, int Kp,
                                                        const float *__restrict__ src, long lds,
                                                        T16 *__restrict__ dst) {
  __shared__ float tile[32][33];
  const int k0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int k = k0 + r, n = n0 + tx;
    tile[r][tx] = (k < K && n < N) ? src[(long)k * lds + n] : 0.f;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int n = n0 + r, k = k0 + tx;
    if (n < N && k < Kp) dst[(long)n * Kp + k] = Lp<T16>::from(tile[tx][r]);
  }
}

template <typename T16>
int pack_rows(int M, int K, int Kp, float scale, const float *src, long lds, T16 *dst,
              hipStream_t s) {
  if (M <= 0 || Kp <= 0) return CLU_OK;
  if (Kp % 4 != 0 || ((uintptr_t)dst & 7)) return -1;
  const int vec = lds % 4 == 0 && ((uintptr_t)src & 15) == 0;
  const dim3 grid((Kp + 1023) / 1024, M < 8192 ? M : 8192);
  hipLaunchKernelGGL((pack_rows_kernel<T16>), grid, dim3(256), 0, s, M, K, Kp, scale, src, lds,
                     dst, vec);
  CLU_LAUNCHED("pack_rows_kernel");
  return CLU_OK;
}
template <typename T16>
int pack_cols(int K, int N, int Kp, const float *src, long lds, T16 *dst, hipStream_t s) {
  if (N <= 0 || Kp <= 0) return CLU_OK;
  dim3 grid((N + 31) / 32, (Kp + 31) / 32);
  hipLaunchKernelGGL((pack_cols_kernel<T16>), grid, dim3(256), 0, s, K, N, Kp, src, lds, dst);
  CLU_LAUNCHED("pack_cols_kernel");
  return CLU_OK;
}
template int pack_rows<__bf16>(int, int, int, float, const float *, long, __bf16 *, hipStream_t);
template int pack_rows<_Float16>(int, int, int, float, const float *, long, _Float16 *,
                                 hipStream_t);
template int pack_cols<__bf16>(int, int, int, const float *, long, __bf16 *, hipStream_t);
template int pack_cols<_Float16>(int, int, int, const float *, long, _Float16 *, hipStream_t);

}  // namespace clu
