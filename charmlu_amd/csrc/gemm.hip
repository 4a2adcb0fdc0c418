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
// chunk.  FP64 LDS tiles are stored k-major (As[k][m], Bs[k][n]) so that
// the 16 lanes that share an MFMA k index read 16 consecutive doubles; FP32
// tiles k-minor with rows of 17 (bank-conflict-free stores and fragment
// reads); NN operands with aligned rows are fetched 16 bytes per lane (FP32: 94-99 -> 102-110 TFLOP/s at 1024 x 64512 x 1024, 65024 x 512 x 512
// and 8192^2 x 1024, 77 -> 90 at 64512 x 256 x 256).
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

// The 16x16x4 matrix-core step for each element type and where element v of
// a lane's 4-wide result sits: FP64 puts row i of the 16x16 block in lane
// group i%4, element i/4; FP32 in lane group i/4, element i%4 (measured on
// gfx950).
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef double4_t acc_t;
  static CLU_DEV acc_t step(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static CLU_DEV int row(int lq, int v) { return lq + 4 * v; }
};
template <> struct Mfma<float> {
  typedef float4_t acc_t;
  static CLU_DEV acc_t step(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static CLU_DEV int row(int lq, int v) { return 4 * lq + v; }
};

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

// C = alpha * acc + beta * C for one wave's TM x TN blocks of 16x16.  The
// beta test is uniform; with beta != 0 the C values of MFMA row block i+1
// are requested before block i is combined and stored, so the epilogue costs
// about one memory round trip instead of one per element.
template <typename T, int TM, int TN>
CLU_DEV void store_tile(const typename Mfma<T>::acc_t (&acc)[TM][TN], T *__restrict__ C, long ldc,
                        int r0, int c0, int lr, int lq, T alpha, T beta, int M, int N,
                        bool full) {
  auto inb = [&](int row, int col) { return full || (row < M && col < N); };
  if (beta == (T)0) {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = r0 + i * 16 + Mfma<T>::row(lq, v), col = c0 + j * 16 + lr;
          if (inb(row, col)) C[(long)row * ldc + col] = alpha * acc[i][j][v];
        }
    return;
  }
  T cur[TN][4], nxt[TN][4];
  auto fetch = [&](int i, T (&dst)[TN][4]) {
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = r0 + i * 16 + Mfma<T>::row(lq, v), col = c0 + j * 16 + lr;
        dst[j][v] = inb(row, col) ? C[(long)row * ldc + col] : (T)0;
      }
  };
  fetch(0, cur);
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    if (i + 1 < TM) fetch(i + 1, nxt);
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = r0 + i * 16 + Mfma<T>::row(lq, v), col = c0 + j * 16 + lr;
        if (inb(row, col)) C[(long)row * ldc + col] = alpha * acc[i][j][v] + beta * cur[j][v];
      }
    if (i + 1 < TM) {
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) cur[j][v] = nxt[j][v];
    }
  }
}

// Grid rows and layers (blockIdx.y, blockIdx.z) select independent problems
// of a strided batch: operand X of problem (y, z) starts at X + y*bs.x + z*bs.y.
struct BatchStrides {
  long a_y, a_z, b_y, b_z, c_y, c_z;
};

template <typename T, int BM, int BN, int WM, int WN, bool TA, bool TB>
__global__ __launch_bounds__(256, 2) void gemm_mfma_kernel(
    int M, int N, int K, T alpha, const T *__restrict__ A, long lda,
    const T *__restrict__ B, long ldb, T beta, T *__restrict__ C,
    long ldc, int tiles_m, int tiles_n, BatchStrides bs) {
  A += blockIdx.y * bs.a_y + blockIdx.z * bs.a_z;
  B += blockIdx.y * bs.b_y + blockIdx.z * bs.b_z;
  C += blockIdx.y * bs.c_y + blockIdx.z * bs.c_z;
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  constexpr int A_PER = BM * GBK / 256, B_PER = BN * GBK / 256;
  // LDS layout: FP32 operands k-minor ([m][k] and [n][k], rows of GBK + 1):
  // the stores (16 lanes along k, or along m / n) and the fragment reads (16
  // lanes along m / n times 4 along k) then fall in distinct banks.  FP64
  // keeps [k][m] and [k][n] (rows of BM + 1 and BN + 1).
  constexpr bool KMIN = sizeof(T) == 4;
  constexpr int ALD = BM + 1, BLD = BN + 1;
  static_assert(WM * WN == 4, "four waves");
  static_assert(TM >= 1 && TN >= 1, "wave tile");
  static_assert(A_PER >= 1 && B_PER >= 1, "load split");

  __shared__ T As[2][KMIN ? BM * (GBK + 1) : GBK * ALD];
  __shared__ T Bs[2][KMIN ? BN * (GBK + 1) : GBK * BLD];
  auto ai = [](int k, int m) { return KMIN ? m * (GBK + 1) + k : k * ALD + m; };
  auto bi = [](int k, int n) { return KMIN ? n * (GBK + 1) + k : k * BLD + n; };

  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm0 = (wave / WN) * WTM, wn0 = (wave % WN) * WTN;
  const int lr = lane & 15, lq = lane >> 4;

  T ra[A_PER], rb[B_PER];
  // NN with 16-byte aligned rows and K, N multiples of VE: each thread
  // loads VE consecutive values per instruction (4 floats or 2 doubles)
  constexpr int VE = 16 / (int)sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(VE)));
  // (FP64 128x128 tiles would spill: they keep the 8-byte loads)
  constexpr bool V4 = !TA && !TB && A_PER % VE == 0 && B_PER % VE == 0 &&
                      (sizeof(T) == 4 || BM * BN <= 64 * 64);
  const bool vec = V4 && ((lda | ldb | K | N) & (VE - 1)) == 0 &&
                   (((unsigned long)A | (unsigned long)B) & 15) == 0;
  auto load = [&](int k0) {
    if constexpr (V4) {
      if (vec) {
#pragma unroll
        for (int r = 0; r < A_PER / VE; ++r) {
          const int e = t + 256 * r, m = e / (GBK / VE), k = VE * (e % (GBK / VE));
          const int gm = m0 + m, gk = k0 + k;
          vec_t v = (vec_t)(T)0;
          if (gm < M && gk < K) v = *reinterpret_cast<const vec_t *>(&A[(long)gm * lda + gk]);
#pragma unroll
          for (int j = 0; j < VE; ++j) ra[VE * r + j] = v[j];
        }
#pragma unroll
        for (int r = 0; r < B_PER / VE; ++r) {
          const int e = t + 256 * r, k = e / (BN / VE), n = VE * (e % (BN / VE));
          const int gn = n0 + n, gk = k0 + k;
          vec_t v = (vec_t)(T)0;
          if (gn < N && gk < K) v = *reinterpret_cast<const vec_t *>(&B[(long)gk * ldb + gn]);
#pragma unroll
          for (int j = 0; j < VE; ++j) rb[VE * r + j] = v[j];
        }
        return;
      }
    }
#pragma unroll
    for (int r = 0; r < A_PER; ++r) {
      const int e = t + 256 * r;
      int m, k;
      if (!TA) { m = e / GBK; k = e % GBK; } else { k = e / BM; m = e % BM; }
      const int gm = m0 + m, gk = k0 + k;
      ra[r] = (gm < M && gk < K)
                  ? (!TA ? A[(long)gm * lda + gk] : A[(long)gk * lda + gm])
                  : (T)0;
    }
#pragma unroll
    for (int r = 0; r < B_PER; ++r) {
      const int e = t + 256 * r;
      int n, k;
      if (!TB) { k = e / BN; n = e % BN; } else { n = e / GBK; k = e % GBK; }
      const int gn = n0 + n, gk = k0 + k;
      rb[r] = (gn < N && gk < K)
                  ? (!TB ? B[(long)gk * ldb + gn] : B[(long)gn * ldb + gk])
                  : (T)0;
    }
  };
  auto stash = [&](int buf) {
    if constexpr (V4) {
      if (vec) {
#pragma unroll
        for (int r = 0; r < A_PER; ++r) {
          const int e = t + 256 * (r / VE), m = e / (GBK / VE), k = VE * (e % (GBK / VE)) + r % VE;
          As[buf][ai(k, m)] = ra[r];
        }
#pragma unroll
        for (int r = 0; r < B_PER; ++r) {
          const int e = t + 256 * (r / VE), k = e / (BN / VE), n = VE * (e % (BN / VE)) + r % VE;
          Bs[buf][bi(k, n)] = rb[r];
        }
        return;
      }
    }
#pragma unroll
    for (int r = 0; r < A_PER; ++r) {
      const int e = t + 256 * r;
      int m, k;
      if (!TA) { m = e / GBK; k = e % GBK; } else { k = e / BM; m = e % BM; }
      As[buf][ai(k, m)] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < B_PER; ++r) {
      const int e = t + 256 * r;
      int n, k;
      if (!TB) { k = e / BN; n = e % BN; } else { n = e / GBK; k = e % GBK; }
      Bs[buf][bi(k, n)] = rb[r];
    }
  };

  typename Mfma<T>::acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = (typename Mfma<T>::acc_t){0, 0, 0, 0};

  const int nk = (K + GBK - 1) / GBK;
  load(0);
  stash(0);
  __syncthreads();
  for (int kc = 0; kc < nk; ++kc) {
    const int buf = kc & 1;
    if (kc + 1 < nk) load((kc + 1) * GBK);
#pragma unroll
    for (int s = 0; s < GBK / 4; ++s) {
      T a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) a[i] = As[buf][ai(4 * s + lq, wm0 + i * 16 + lr)];
#pragma unroll
      for (int j = 0; j < TN; ++j) b[j] = Bs[buf][bi(4 * s + lq, wn0 + j * 16 + lr)];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = Mfma<T>::step(a[i], b[j], acc[i][j]);
    }
    if (kc + 1 < nk) stash(buf ^ 1);
    __syncthreads();
  }

  const bool full = m0 + BM <= M && n0 + BN <= N;
  store_tile<T, TM, TN>(acc, C, ldc, m0 + wm0, n0 + wn0, lr, lq, alpha, beta, M, N, full);
}

// ---------------------------------------------------------------------------
// Aligned full-tile NN kernel: 128x128 tiles, K in stages of 8 loaded with
// global_load_lds_dwordx4 straight into LDS (no register staging), NST
// stages in flight, one barrier per stage.  Used when M and N are multiples
// of 128, K of 8, and every row start is 16-byte aligned -- the case of the
// LU trailing update.
//
// Three workgroups per CU (three waves per SIMD: at most 168 VGPRs, three
// 48 KB stage rings): while one workgroup waits for its C tile at the start
// or drains its stores at the end, the other two keep the matrix cores busy.
// Two workgroups with four stages reached 67.8 % of the FP64 peak at
// 32512^2 x 256, three with three stages 75.7 %.  C is read and written
// with non-temporal accesses (each C element is touched once per launch), so
// it does not evict the A and B tiles from L2: 79.5 %; with the first stages
// requested before C, 82 %.  The global addresses
// are uniform bases plus one 32-bit lane offset each, so that the 32 C
// addresses of a lane are not kept live across the K loop (64 VGPRs).
//
// LDS layout per stage (16 KB): A as 128 rows x 4 chunks of 16 bytes, the
// chunk index XORed with lds_swz(row); B as 8 rows x 64 chunks, the chunk
// index XORed with (k & 1) << 3.  A direct-to-LDS load writes lane i's 16
// bytes at base + 16 i, so the swizzle is applied by choosing which global
// chunk each lane fetches.
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(3))) void *lds_ptr_t;
typedef __attribute__((address_space(1))) void *gbl_ptr_t;

CLU_DEV void load16_to_lds(const double *g, void *lds_base) {
  __builtin_amdgcn_global_load_lds((gbl_ptr_t)g, (lds_ptr_t)lds_base, 16, 0, 0);
}
typedef double dvec2_t __attribute__((ext_vector_type(2)));
CLU_DEV double2 ld_nt(const double2 *p) {
  const dvec2_t v = __builtin_nontemporal_load(reinterpret_cast<const dvec2_t *>(p));
  return double2{v.x, v.y};
}
CLU_DEV void st_nt(double2 *p, double x, double y) {
  __builtin_nontemporal_store((dvec2_t){x, y}, reinterpret_cast<dvec2_t *>(p));
}
// p advanced by a byte offset (a 32-bit lane offset added to a uniform base)
template <typename T>
CLU_DEV T *byte_off(T *p, unsigned off) {
  return reinterpret_cast<T *>(reinterpret_cast<char *>(const_cast<void *>(
                                   reinterpret_cast<const void *>(p))) + off);
}

// s_waitcnt vmcnt(N) (N < 16) leaving the other counters unconstrained
template <int N>
CLU_DEV void wait_vm() {
  static_assert(N >= 0 && N < 16, "vmcnt");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | (0x7 << 4) | (0xF << 8));
}

// One 128 x 128 tile of C = alpha A B + beta C.  CINIT (alpha = +-1 only):
// beta C is loaded into the accumulators before the K loop and alpha folded
// into the sign of the A fragments, so the epilogue is stores only and the
// C reads overlap the first stages' loads instead of following the last
// MFMA.  Otherwise the epilogue reads C once the products are done.
template <int NST, bool CINIT>
CLU_DEV void lds_tile(char *smem, int m0, int n0, int K, double alpha, const double *__restrict__ A,
                      long lda, const double *__restrict__ B, long ldb, double beta,
                      double *__restrict__ C, long ldc) {
  constexpr int BM = 128, BN = 128, BK = 8;
  constexpr int STAGE = (BM * BK + BK * BN) * 8;  // bytes
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wm0 = (wave >> 1) * 64, wn0 = (wave & 1) * 64;
  const int lr = lane & 15, lq = lane >> 4;

  // global source of this thread's chunks (two of A, two of B per stage):
  // A chunk slot c = 64 wave + 256 r + lane is row c / 4 = 16 wave + 64 r +
  // lane / 4, 16-byte chunk (c & 3) ^ lds_swz(row) (the swizzle depends on
  // lane / 4 only); B chunk c is k row c / 64 = wave + 4 r, chunk
  // (c & 63) ^ ((wave & 1) << 3).  Uniform row bases, one lane offset each.
  const double *ga[2], *gb[2];
  int la_off[2], lb_off[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    ga[r] = A + (long)(m0 + wave * 16 + 64 * r) * lda;
    gb[r] = B + (long)(wave + 4 * r) * ldb + n0;
    la_off[r] = (wave * 64 + 256 * r) * 16;  // wave-uniform slot base
    lb_off[r] = BM * BK * 8 + (wave * 64 + 256 * r) * 16;
  }
  const unsigned a_lo = (unsigned)(((lane >> 2) * (int)lda + 2 * ((lane & 3) ^ lds_swz(lane >> 2))) * 8);
  const unsigned b_lo = (unsigned)(2 * (lane ^ ((wave & 1) << 3)) * 8);
  const int nk = K / BK;
  auto issue = [&](int kc) {
    const int kk = kc * BK;
    char *st = smem + (kc % NST) * STAGE;
#pragma unroll
    for (int r = 0; r < 2; ++r) load16_to_lds(byte_off(ga[r] + kk, a_lo), st + la_off[r]);
#pragma unroll
    for (int r = 0; r < 2; ++r) load16_to_lds(byte_off(gb[r] + (long)kk * ldb, b_lo), st + lb_off[r]);
  };

  // per-lane fragment offsets inside a stage
  // The two matrix-core steps of a stage take k = 2 lq + s2 on lane group
  // lq (s2 = 0, 1), so a lane's A values of both steps are the two doubles
  // of one 16-byte LDS chunk (chunk lq of its row): one read per row block
  // The 16 columns of matrix-core block j are not contiguous: lane lr of
  // block pair p = j / 2 holds columns 32 p + 2 lr (block 2p) and
  // 32 p + 2 lr + 1 (block 2p + 1), so a lane's B values of both blocks of a
  // pair, and its C values, are 16-byte pairs
  int aoff[4], boff[2][2];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = wm0 + 16 * i + lr;
    aoff[i] = m * 64 + ((lq ^ lds_swz(m)) << 4);
  }
#pragma unroll
  for (int s2 = 0; s2 < 2; ++s2) {
    const int k = 2 * lq + s2;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int n = wn0 + 32 * p + 2 * lr;
      boff[s2][p] = BM * BK * 8 + k * 1024 + (((n >> 1) ^ ((k & 1) << 3)) << 4);
    }
  }

  // accumulator element (i, j, v) is C row 16 i + lq + 4 v, column
  // 32 (j / 2) + 2 lr + (j % 2) of this wave's 64 x 64 block: a uniform
  // base per (i, j / 2, v) plus the lane's offset c_lo (bytes)
  double *Cw = C + (long)(m0 + wm0) * ldc + n0 + wn0;
  unsigned c_lo = (unsigned)(lq * (int)ldc + 2 * lr) * 8u;
  auto cpair = [&](int i, int p, int v) {
    return byte_off(reinterpret_cast<double2 *>(Cw + (long)(i * 16 + 4 * v) * ldc + 32 * p), c_lo);
  };
  // The first operand stages are requested before the C tile: behind the
  // tile's 32 C loads per lane they would reach LDS later than the C values
  // reach the accumulators (79.5 -> 82 % of the FP64 peak at 32512^2 x 256).
  static_assert(NST >= 2 && NST <= 4, "stages");
#pragma unroll
  for (int st = 0; st < NST - 1; ++st)
    if (st < nk) issue(st);
  double4_t acc[4][4];
  if (CINIT && beta != 0.0) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const double2 c2 = ld_nt(cpair(i, p, v));
          acc[i][2 * p][v] = beta * c2.x;
          acc[i][2 * p + 1][v] = beta * c2.y;
        }
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
  }
  const bool neg = CINIT && alpha < 0.0;

  for (int kc = 0; kc < nk; ++kc) {
    // This thread's loads of stage kc have landed once at most the 4 loads
    // of each later stage already issued (min(NST - 2, nk - 1 - kc) of
    // them: no stage past the end is loaded) are outstanding (at kc = 0 the
    // C loads, issued after them, are waited for too); after the
    // barrier every thread's have, and every wave is done reading stage
    // kc-1, whose buffer the next issue overwrites.  A bare s_barrier:
    // __syncthreads() would also drain the direct-to-LDS loads of the later
    // stages.
    const int ahead = nk - 1 - kc;
    if (NST >= 4 && ahead >= 2) wait_vm<4 * (NST >= 4 ? 2 : 0)>();
    else if (NST >= 3 && ahead >= 1) wait_vm<4>();
    else wait_vm<0>();
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
    if (kc + NST - 1 < nk) issue(kc + NST - 1);
    const char *st = smem + (kc % NST) * STAGE;
    double2 a2[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a2[i] = *reinterpret_cast<const double2 *>(st + aoff[i]);
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = s2 ? a2[i].y : a2[i].x;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const double2 b2 = *reinterpret_cast<const double2 *>(st + boff[s2][p]);
        b[2 * p] = b2.x;
        b[2 * p + 1] = b2.y;
      }
      if (neg) {
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = -a[i];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
  wait_vm<0>();  // no direct-to-LDS write may be in flight past this tile
  // a new value as far as the compiler knows: the C addresses are formed
  // again here instead of being held in registers through the K loop
  __asm__ volatile("" : "+v"(c_lo));
  if (CINIT) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int v = 0; v < 4; ++v) st_nt(cpair(i, p, v), acc[i][2 * p][v], acc[i][2 * p + 1][v]);
  } else {
    // C = alpha acc + beta C, a row block's C pairs requested before the
    // previous one is combined and stored
    double2 cur[2][4], nxt[2][4];
    auto fetch = [&](int i, double2 (&d)[2][4]) {
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int v = 0; v < 4; ++v) d[p][v] = beta != 0.0 ? ld_nt(cpair(i, p, v)) : double2{0.0, 0.0};
    };
    fetch(0, cur);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i + 1 < 4) fetch(i + 1, nxt);
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          st_nt(cpair(i, p, v), alpha * acc[i][2 * p][v] + beta * cur[p][v].x,
                alpha * acc[i][2 * p + 1][v] + beta * cur[p][v].y);
      if (i + 1 < 4) {
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
          for (int v = 0; v < 4; ++v) cur[p][v] = nxt[p][v];
      }
    }
  }
}

// Aligned full-tile NN kernel, one tile per workgroup.  ROLE only names the
// launch site in profiles: 1 = the factorization's trailing update, 2 = its
// look-ahead block column, 0 = every other caller.  The general-alpha form
// (C read after the products) keeps two workgroups per CU: its epilogue
// holds a row block of C beside the accumulators.
template <int NST, int ROLE, bool CINIT>
__global__ __launch_bounds__(256, CINIT ? 3 : 2) void dgemm_nn_lds_kernel(
    int M, int N, int K, double alpha, const double *__restrict__ A, long lda,
    const double *__restrict__ B, long ldb, double beta, double *__restrict__ C, long ldc,
    int tiles_m, int tiles_n) {
  __shared__ __attribute__((aligned(1024))) char smem[NST * (128 * 8 + 8 * 128) * 8];
  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
  lds_tile<NST, CINIT>(smem, tm * 128, tn * 128, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

// ---------------------------------------------------------------------------
// The FP32 form of the direct-to-LDS NN kernel (the mixed LU's U row, its
// unpivoted L21 and the larger updates inside its panels): a 128 x 128 C
// tile, 4 waves of 64 x 64, K in stages of 16 (64-byte A rows as in the FP64
// kernel, with the same XOR swizzle of the 16-byte chunks), three stages and
// three workgroups per CU or four and two, v_mfma_f32_16x16x4f32.  Matrix-core step
// s of a stage takes k = 4 q + s on lane group q, so a lane's A values of
// the four steps are one 16-byte chunk.  The 16 columns of matrix-core
// block j are columns 4 r + j (r the lane in its group of 16) of the wave's
// 64, so one 16-byte read gives a lane its B values of the four blocks at
// one k, and its C values of one row are a 16-byte quad.
// FP32 operand layout (measured on gfx950): A[i][k] from lane i + 16 k,
// B[k][j] from lane j + 16 k, D[i][j] in lane j + 16 (i / 4), element i % 4.
// ---------------------------------------------------------------------------
CLU_DEV void load16_to_lds(const float *g, void *lds_base) {
  __builtin_amdgcn_global_load_lds((gbl_ptr_t)g, (lds_ptr_t)lds_base, 16, 0, 0);
}
CLU_DEV float4_t ld_nt4(const float4_t *p) { return __builtin_nontemporal_load(p); }
CLU_DEV void st_nt4(float4_t *p, float4_t v) { __builtin_nontemporal_store(v, p); }

template <int NST, bool CINIT>
CLU_DEV void lds_tile_f32(char *smem, int m0, int n0, int K, float alpha,
                          const float *__restrict__ A, long lda, const float *__restrict__ B,
                          long ldb, float beta, float *__restrict__ C, long ldc) {
  constexpr int BM = 128, BN = 128, BK = 16;
  constexpr int STAGE = (BM * BK + BK * BN) * 4;  // bytes
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wm0 = (wave >> 1) * 64, wn0 = (wave & 1) * 64;
  const int lr = lane & 15, lq = lane >> 4;

  // this thread's chunks per stage: A slot c = 64 wave + 256 r + lane is
  // row c / 4 = 16 wave + 64 r + lane / 4, chunk (c & 3) ^ lds_swz(row); B
  // slot c is k row c / 32 = 2 wave + 8 r + lane / 32, chunk c & 31 (B rows
  // of 128 floats, 512 bytes, unswizzled: a fragment read takes 16
  // consecutive chunks of one row per lane group)
  const float *ga[2], *gb[2];
  int la_off[2], lb_off[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    ga[r] = A + (long)(m0 + wave * 16 + 64 * r) * lda;
    gb[r] = B + (long)(2 * wave + 8 * r) * ldb + n0;
    la_off[r] = (wave * 64 + 256 * r) * 16;
    lb_off[r] = BM * BK * 4 + (wave * 64 + 256 * r) * 16;
  }
  const unsigned a_lo =
      (unsigned)(((lane >> 2) * (int)lda + 4 * ((lane & 3) ^ lds_swz(lane >> 2))) * 4);
  const unsigned b_lo = (unsigned)(((lane >> 5) * (int)ldb + 4 * (lane & 31)) * 4);
  const int nk = K / BK;
  auto issue = [&](int kc) {
    const int kk = kc * BK;
    char *st = smem + (kc % NST) * STAGE;
#pragma unroll
    for (int r = 0; r < 2; ++r) load16_to_lds(byte_off(ga[r] + kk, a_lo), st + la_off[r]);
#pragma unroll
    for (int r = 0; r < 2; ++r) load16_to_lds(byte_off(gb[r] + (long)kk * ldb, b_lo), st + lb_off[r]);
  };
  int aoff[4], boff[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = wm0 + 16 * i + lr;
    aoff[i] = m * 64 + ((lq ^ lds_swz(m)) << 4);
  }
#pragma unroll
  for (int s4 = 0; s4 < 4; ++s4) boff[s4] = BM * BK * 4 + (4 * lq + s4) * 512 + (wn0 + 4 * lr) * 4;

  // accumulator element (i, j, e) is C row 16 i + 4 lq + e, column 4 lr + j
  // of this wave's 64 x 64 block
  float *Cw = C + (long)(m0 + wm0) * ldc + n0 + wn0;
  unsigned c_lo = (unsigned)(4 * lq * (int)ldc + 4 * lr) * 4u;
  auto cquad = [&](int i, int e) {
    return byte_off(reinterpret_cast<float4_t *>(Cw + (long)(i * 16 + e) * ldc), c_lo);
  };
  static_assert(NST >= 2 && NST <= 4, "stages");
#pragma unroll
  for (int st = 0; st < NST - 1; ++st)
    if (st < nk) issue(st);
  float4_t acc[4][4];
  if (CINIT && beta != 0.f) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float4_t c4 = ld_nt4(cquad(i, e));
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j][e] = beta * c4[j];
      }
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = (float4_t){0.f, 0.f, 0.f, 0.f};
  }
  const bool neg = CINIT && alpha < 0.f;

  for (int kc = 0; kc < nk; ++kc) {
    const int ahead = nk - 1 - kc;
    if (NST >= 4 && ahead >= 2) wait_vm<4 * (NST >= 4 ? 2 : 0)>();
    else if (NST >= 3 && ahead >= 1) wait_vm<4>();
    else wait_vm<0>();
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
    if (kc + NST - 1 < nk) issue(kc + NST - 1);
    const char *st = smem + (kc % NST) * STAGE;
    float4_t a4[4], b4[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a4[i] = *reinterpret_cast<const float4_t *>(st + aoff[i]);
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) b4[s4] = *reinterpret_cast<const float4_t *>(st + boff[s4]);
    if (neg) {
#pragma unroll
      for (int i = 0; i < 4; ++i) a4[i] = -a4[i];
    }
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[i][s4], b4[s4][j], acc[i][j], 0, 0, 0);
  }
  wait_vm<0>();
  __asm__ volatile("" : "+v"(c_lo));
  if (CINIT) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        st_nt4(cquad(i, e), (float4_t){acc[i][0][e], acc[i][1][e], acc[i][2][e], acc[i][3][e]});
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float4_t old = beta != 0.f ? ld_nt4(cquad(i, e)) : (float4_t){0.f, 0.f, 0.f, 0.f};
        float4_t v;
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = alpha * acc[i][j][e] + beta * old[j];
        st_nt4(cquad(i, e), v);
      }
  }
}

template <int NST, bool CINIT, int OCC>
__global__ __launch_bounds__(256, OCC) void sgemm_nn_lds_kernel(
    int M, int N, int K, float alpha, const float *__restrict__ A, long lda,
    const float *__restrict__ B, long ldb, float beta, float *__restrict__ C, long ldc,
    int tiles_m, int tiles_n) {
  __shared__ __attribute__((aligned(1024))) char smem[NST * (128 * 16 + 16 * 128) * 4];
  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
  lds_tile_f32<NST, CINIT>(smem, tm * 128, tn * 128, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

template <typename T, int BM, int BN, int WM, int WN, bool TA, bool TB>
static int launch(int M, int N, int K, T alpha, const T *A, long lda, const T *B, long ldb,
                  T beta, T *C, long ldc, hipStream_t s, int ny = 1, int nz = 1,
                  BatchStrides bs = BatchStrides{0, 0, 0, 0, 0, 0}) {
  const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  const long nblk = (long)tiles_m * tiles_n;
  if (nblk > 0x7fffffffL || ny < 1 || ny > 65535 || nz < 1 || nz > 65535) return -1;
  hipLaunchKernelGGL((gemm_mfma_kernel<T, BM, BN, WM, WN, TA, TB>),
                     dim3((unsigned)nblk, (unsigned)ny, (unsigned)nz), dim3(256), 0, s, M, N, K,
                     alpha, A, lda, B, ldb, beta, C, ldc, tiles_m, tiles_n, bs);
  CLU_LAUNCHED("gemm_mfma_kernel");
  return CLU_OK;
}

static bool lds_aligned(int M, int N, int K, const double *A, long lda, const double *B,
                        long ldb, const double *C, long ldc) {
  return M % 128 == 0 && N % 128 == 0 && K % 8 == 0 && K > 0 && lda % 2 == 0 && ldb % 2 == 0 &&
         ldc % 2 == 0 && ((uintptr_t)A & 15) == 0 && ((uintptr_t)B & 15) == 0 &&
         ((uintptr_t)C & 15) == 0;
}
// the direct-to-LDS kernel for general callers: aligned and enough tiles to
// fill the chip (smaller grids do better with 64 x 64 tiles)
static bool lds_path_ok(int M, int N, int K, const double *A, long lda, const double *B,
                        long ldb, const double *C, long ldc) {
  return lds_aligned(M, N, K, A, lda, B, ldb, C, ldc) && (long)(M / 128) * (N / 128) >= 256;
}

static bool lds_path_ok_f32(int M, int N, int K, const float *A, long lda, const float *B,
                            long ldb, const float *C, long ldc) {
  return M % 128 == 0 && N % 128 == 0 && K % 16 == 0 && K > 0 && lda % 4 == 0 &&
         ldb % 4 == 0 && ldc % 4 == 0 && ((uintptr_t)A & 15) == 0 && ((uintptr_t)B & 15) == 0 &&
         ((uintptr_t)C & 15) == 0 && (long)(M / 128) * (N / 128) >= 256;
}

template <typename T, bool TA, bool TB>
static int dispatch(int M, int N, int K, T alpha, const T *A, long lda, const T *B, long ldb,
                    T beta, T *C, long ldc, hipStream_t s, int role) {
  if (sizeof(T) == 4 && !TA && !TB &&
      lds_path_ok_f32(M, N, K, (const float *)A, lda, (const float *)B, ldb, (const float *)C,
                      ldc)) {
    const int tiles_m = M / 128, tiles_n = N / 128;
    const bool cinit = alpha == (T)1 || alpha == (T)-1;
    // K >= 512: four stages, two workgroups per CU (1024 x 64512 x 1024:
    // 131.7 against 127.4 TFLOP/s as C = A B; 8192^2 x 1024: 130.5 against
    // 125.7); shorter K: three stages, three per CU (32768^2 x 256: 123.7
    // against 120.6)
#define CLU_SLDS_LAUNCH(NS, CI, OC)                                                            \
  hipLaunchKernelGGL((sgemm_nn_lds_kernel<NS, CI, OC>), dim3(tiles_m * tiles_n), dim3(256), 0, \
                     s, M, N, K, (float)alpha, (const float *)A, lda, (const float *)B, ldb,   \
                     (float)beta, (float *)C, ldc, tiles_m, tiles_n)
    if (K >= 512) {
      if (cinit) CLU_SLDS_LAUNCH(4, true, 2);
      else CLU_SLDS_LAUNCH(4, false, 2);
    } else {
      if (cinit) CLU_SLDS_LAUNCH(3, true, 3);
      else CLU_SLDS_LAUNCH(3, false, 3);
    }
#undef CLU_SLDS_LAUNCH
    CLU_LAUNCHED("sgemm_nn_lds_kernel");
    return CLU_OK;
  }
  const bool trailing = role == 1 && (alpha == (T)1 || alpha == (T)-1);
  if (sizeof(T) == 8 && !TA && !TB &&
      (trailing ? lds_aligned(M, N, K, (const double *)A, lda, (const double *)B, ldb,
                              (const double *)C, ldc)
                : lds_path_ok(M, N, K, (const double *)A, lda, (const double *)B, ldb,
                              (const double *)C, ldc))) {
    const int tiles_m = M / 128, tiles_n = N / 128;
    const bool cinit = alpha == (T)1 || alpha == (T)-1;
#define CLU_LDS_LAUNCH(R, CI)                                                                   \
  hipLaunchKernelGGL((dgemm_nn_lds_kernel<3, R, CI>), dim3(tiles_m * tiles_n), dim3(256), 0, s, \
                     M, N, K, (double)alpha, (const double *)A, lda, (const double *)B, ldb,    \
                     (double)beta, (double *)C, ldc, tiles_m, tiles_n)
    if (trailing) CLU_LDS_LAUNCH(1, true);
    else if (role == 2 && cinit) CLU_LDS_LAUNCH(2, true);
    else if (cinit) CLU_LDS_LAUNCH(0, true);
    else CLU_LDS_LAUNCH(0, false);
#undef CLU_LDS_LAUNCH
    CLU_LAUNCHED("dgemm_nn_lds_kernel");
    return CLU_OK;
  }
  // narrow right-hand sides (panel-internal updates, multi-RHS solves):
  // 64-row tiles, four times the workgroups of 256-row ones (these updates
  // are memory- and latency-bound, and the panel's run on 32 CUs)
  if (N <= 32)
    return launch<T, 64, 32, 4, 1, TA, TB>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  if (N <= 64)
    return launch<T, 64, 64, 2, 2, TA, TB>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  if (M <= 32)
    return launch<T, 32, 256, 1, 4, TA, TB>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  // small problems: more, smaller tiles to fill 256 CUs
  if ((long)((M + 127) / 128) * ((N + 127) / 128) < 512)
    return launch<T, 64, 64, 2, 2, TA, TB>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  return launch<T, 128, 128, 2, 2, TA, TB>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

template <typename T>
static int gemm_any(bool ta, bool tb, int M, int N, int K, T alpha, const T *A, long lda,
                    const T *B, long ldb, T beta, T *C, long ldc, hipStream_t s, int role = 0) {
  if (M <= 0 || N <= 0) return CLU_OK;
  if (K <= 0 || alpha == (T)0) K = 0;  // C = beta * C: the K loop runs zero chunks
  if (!ta && !tb)
    return dispatch<T, false, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s, role);
  if (ta && !tb)
    return dispatch<T, true, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s, role);
  if (!ta && tb)
    return dispatch<T, false, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s, role);
  return dispatch<T, true, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s, role);
}

int dgemm_lookahead(int M, int N, int K, double alpha, const double *A, long lda,
                    const double *B, long ldb, double beta, double *C, long ldc, hipStream_t s) {
  if (alpha != 1.0 && alpha != -1.0) return -1;
  return gemm_any<double>(false, false, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s, 2);
}

// The trailing update (alpha = +-1): the direct-to-LDS kernel whenever the
// operands are aligned, whatever the grid size, so both parts of a split
// update run the same kernel.
int dgemm_trailing(int M, int N, int K, double alpha, const double *A, long lda, const double *B,
                   long ldb, double beta, double *C, long ldc, hipStream_t s) {
  if (alpha != 1.0 && alpha != -1.0) return -1;
  return gemm_any<double>(false, false, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s, 1);
}
bool dgemm_trailing_lds(int M, int N, int K, const double *A, long lda, const double *B,
                        long ldb, const double *C, long ldc) {
  return lds_aligned(M, N, K, A, lda, B, ldb, C, ldc);
}

int dgemm(bool ta, bool tb, int M, int N, int K, double alpha, const double *A, long lda,
          const double *B, long ldb, double beta, double *C, long ldc, hipStream_t s) {
  return gemm_any<double>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

int sgemm(bool ta, bool tb, int M, int N, int K, float alpha, const float *A, long lda,
          const float *B, long ldb, float beta, float *C, long ldc, hipStream_t s) {
  return gemm_any<float>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

// NN products of a strided batch of small problems (the blocks of a
// triangular inverse): 32 x 256-row tiles for N <= 32, else 64 x 64.
template <typename T>
int gemm_batched(int M, int N, int K, T alpha, const T *A, long lda, const T *B, long ldb,
                 T beta, T *C, long ldc, int ny, int nz, long a_y, long a_z, long b_y, long b_z,
                 long c_y, long c_z, hipStream_t s) {
  if (M <= 0 || N <= 0 || ny <= 0 || nz <= 0) return CLU_OK;
  if (K <= 0 || alpha == (T)0) K = 0;
  const BatchStrides bs{a_y, a_z, b_y, b_z, c_y, c_z};
  if (N <= 32)
    return launch<T, 256, 32, 4, 1, false, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                                                  s, ny, nz, bs);
  return launch<T, 64, 64, 2, 2, false, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s,
                                              ny, nz, bs);
}
template int gemm_batched<double>(int, int, int, double, const double *, long, const double *,
                                  long, double, double *, long, int, int, long, long, long, long,
                                  long, long, hipStream_t);
template int gemm_batched<float>(int, int, int, float, const float *, long, const float *, long,
                                 float, float *, long, int, int, long, long, long, long, long,
                                 long, hipStream_t);

}  // namespace clu
