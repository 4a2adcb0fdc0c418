// panel.hip -- panel factorization and row interchanges.
//
// panel_base_kernel replaces the reference's active-panel loop
// (factorDiagBlock lu.ci:256-370 + factorBelowDiagBlock lu.ci:186-251) for a
// column strip of up to 16 columns: per column, an argmax over the rows
// (MaxElm / MaxElm_max, maxelm.h:7-13, lu.C:197-218), the pivot-row exchange
// (swapLocal / pivotRowSwap, lu.C:166-177, lu.ci:290-323), the multipliers by
// division plus the eager update of the next column and its candidate
// (computeMultipliersAndFindColMax, lu.C:259-288) and the rank-1 update of
// the rest of the strip (updateLsubBlock, lu.C:230-254).
//
// The strip lives in registers for the whole kernel: workgroup g, thread t
// of NT holds rows g*NT*R + r*NT + t (r < R), 16 values each.  With one
// workgroup a column costs one LDS barrier.  With G > 1 workgroups each
// column needs one grid-wide exchange: every workgroup publishes its best
// candidate (value, row, and the row's 16 values), and workgroup 0 the
// current top row, into a double-buffered slot array with relaxed
// agent-scope stores; each 8-byte word carries the column's sequence number
// beside its payload, so readers poll until every word they need carries
// the current one (one memory round trip, no counter barrier).  Every
// workgroup then reduces the G candidates itself (same data, same order, so
// all agree on the pivot) and swaps/eliminates its own rows.
//
// pivots_to_moves / apply_moves replace the reference's pivot agglomeration
// and cycle application (recordPivot / announceAgglomeratedPivots /
// sendPendingPivots, lu.C:46-163, lu.C:300-352): a run of sequential swaps
// becomes a list of (destination row, source row) moves, applied to whole
// column ranges (each workgroup gathers the moved row segments of its
// columns into registers, then writes them: one read and one write per
// moved row segment, coalesced along the row).
#include <type_traits>

#include "common.h"
#include "kernels.h"

namespace clu {

// Strip width SW: 8, 16 or 32 columns.  A 32-column strip halves the strip
// launches of a panel and removes the 32-column nodes of its recursion (a
// triangular solve and a GEMM each).  FP32 strips of at most 32 * 2048 rows
// are 32 wide (4 rows of 32 floats per thread; taller ones need 8 rows per
// thread).  FP64 strips are 32 wide up to 512 rows (one workgroup) and from
// 1025 to 9216 rows (2 rows of 32 doubles per thread, at most 18
// workgroups: their records fill the polling waves): per 32 columns 79 us
// against 2 x 36 us plus the node at 512 rows, 120-131 us against
// 2 x 62-66 us plus the node at 2048-8192 rows; from 513 to 1024 rows the
// one workgroup of 512 threads this would need is slower (119 us against
// 2 x 42).  8-column strips (8 rows of 8 doubles or 16 rows of 8 floats per
// thread, 512-thread workgroups) are the fallback for strips too tall for
// the wider ones on the workgroups that can be resident at once.
constexpr int PW8 = 8;      // the fallback strip
constexpr int PW = 16;      // the narrow strip
constexpr int PW32 = 32;    // the wide strip

// Cross-workgroup records.  Every 8-byte word carries a 32-bit payload in
// its low half and the column's sequence number in its high half, so a
// reader can check each word on its own: a column costs one memory round
// trip (stores are not waited for, nothing is ordered by fences).  Record
// of workgroup g for column parity par at recs + (par * GMAX + g) * RW,
// in 32-bit payload words (VW per value, FP64 low half first):
//   [0, SW*VW)          the workgroup's best row (its SW values; zeros if none)
//   [SW*VW, SW*VW+VW)   |value| of that row in the current column
//   SW*VW+VW            its row (local index, -1: none)
//   [TOPW, TOPW+SW*VW)  workgroup 0 only: row c (the top row) before the swap
constexpr int GMAX = 64;       // workgroups in one exchange
constexpr int RECW_MAX = 136;  // words per record, at most (RW: each kernel's own)
static_assert(2L * GMAX * RECW_MAX == PANEL_SLOT_WORDS, "slot array size");
constexpr int SPIN_MAX = 1 << 22;  // polling rounds before giving up (seconds)

// Payload word k of a value: FP64 low half first (little-endian), so the
// payloads restaged in LDS read back as the values themselves.
CLU_DEV unsigned value_word(double v, int k) {
  return (unsigned)(k == 0 ? __double2loint(v) : __double2hiint(v));
}
CLU_DEV unsigned value_word(float v, int) { return __float_as_uint(v); }
template <typename T>
CLU_DEV unsigned row_word(const T *row, int k) {  // word k of an LDS row
  return reinterpret_cast<const unsigned *>(row)[k];
}

template <typename T, int R, int NT, bool MULTI, bool PIV, int SW>
__global__ __launch_bounds__(NT) void panel_base_kernel(
    T *__restrict__ A, long lda, int j0, int m, int w, int pc0, int pc1,
    int *__restrict__ ipiv, unsigned long long *__restrict__ recs, unsigned seq0, int *info,
    unsigned *abortw) {
  static_assert(PIV || !MULTI, "the unpivoted strip is one workgroup");
  constexpr int VW = sizeof(T) / 4;      // payload words per value
  constexpr int CW = SW * VW;            // words of a row
  constexpr int L1 = CW + VW + 1;        // words of a record without the top row
  constexpr int TOPW = (L1 + 1) / 2 * 2; // the top row, 8-byte aligned
  constexpr int RPW = NT * R;            // rows per workgroup
  constexpr int NW = NT / 64;            // waves
  constexpr int CHW = NW >= 8 ? 4 : 5;   // record words per polling lane
  constexpr int RW = TOPW + CW;          // words per record
  static_assert(RW <= RECW_MAX, "record size");
  const int G = MULTI ? (int)gridDim.x : 1, g = MULTI ? (int)blockIdx.x : 0;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const T tiny = (T)100 * (sizeof(T) == 8 ? (T)2.220446049250313e-16 : (T)1.1920929e-07);

  // per-column LDS slots, double-buffered by column parity
  __shared__ T s_val[2][NW];
  __shared__ int s_idx[2][NW];
  __shared__ T s_wcand[2][NW][SW];  // each wave's candidate row
  __shared__ T s_top[2][SW];        // row c before the swap (in the workgroup holding it)
  __shared__ __attribute__((aligned(16))) unsigned s_rec[2][MULTI ? GMAX : 1][RW];
  __shared__ int s_piv[SW];
  __shared__ T s_pmag[SW];

  T a[R][SW];
#pragma clang loop unroll(full)
  for (int r = 0; r < R; ++r) {
    const int li = g * RPW + r * NT + t;
#pragma clang loop unroll(full)
    for (int q = 0; q < SW; ++q)
      a[r][q] = (li < m && q < w) ? A[(long)(j0 + li) * lda + j0 + q] : (T)0;
  }
  // Drain the strip loads here: otherwise the compiler waits for them at the
  // top of the column loop, where the wait then also covers the previous
  // column's ipiv store on every iteration.
  wait_all_mem();

  // MULTI: the record words this lane polls (word offsets inside one
  // parity's records; -1: none, or this workgroup's own record).  Wave v
  // takes words [v * 64 * CHW, (v + 1) * 64 * CHW) of the list; the host
  // keeps G * L1 + CW <= NW * 64 * CHW.
  int poff[MULTI ? CHW : 1];
  bool poller = false;
#pragma clang loop unroll(full)
  for (int i = 0; i < (MULTI ? CHW : 1); ++i) {
    const int e = wave * 64 * CHW + lane + 64 * i;
    const int rg = e < G * L1 ? e / L1 : 0;
    const int rw = e < G * L1 ? e % L1 : TOPW + (e - G * L1);
    poff[i] = (MULTI && e < G * L1 + CW && rg != g) ? rg * RW + rw : -1;
    poller = poller || poff[i] >= 0;
  }
  poller = __any(poller);  // wave-uniform

  // The column loop stays rolled (a fully unrolled strip kernel is tens of
  // thousands of instructions and runs from instruction-cache misses).  Four
  // columns are unrolled per iteration at static register offsets K = 0..3;
  // after each iteration every row is rotated left by four positions, so the
  // next four columns are again at offsets 0..3.
  auto rotate4 = [&]() {
#pragma clang loop unroll(full)
    for (int r = 0; r < R; ++r) {
      T keep[4];
#pragma clang loop unroll(full)
      for (int q = 0; q < 4; ++q) keep[q] = a[r][q];
#pragma clang loop unroll(full)
      for (int q = 0; q < SW - 4; ++q) a[r][q] = a[r][q + 4];
#pragma clang loop unroll(full)
      for (int q = 0; q < 4; ++q) a[r][SW - 4 + q] = keep[q];
    }
  };
  // this thread's candidate for column c over its rows >= c: strictly
  // larger magnitude, rows in increasing order (lowest row on ties)
  auto candidate = [&](int c, auto kc, T &bv, int &bi) {
    constexpr int K = decltype(kc)::value;
#pragma clang loop unroll(full)
    for (int r = 0; r < R; ++r) {
      const int li = g * RPW + r * NT + t;
      const T v = fabs(a[r][K]);
      if (li >= c && v > bv) { bv = v; bi = li; }
    }
  };
  // multipliers (division) and the rank-1 update of the later columns.  In
  // the rotated frame the columns still to come sit at offsets K+1 .. lim-1
  // (lim = w - base); the finished ones (rotated to the end) and the
  // padding get a zero pivot-row entry, so they are left as they are.
  // Branch-free: every row computes its multiplier, rows <= c keep their
  // values by selection and update with a zero multiplier.
  auto eliminate = [&](int c, int lim, const auto &u, auto kc) {
    constexpr int K = decltype(kc)::value;
    T uq[SW];
#pragma clang loop unroll(full)
    for (int q = 0; q < SW; ++q) uq[q] = u[q];
#pragma clang loop unroll(full)
    for (int q = K + 1; q < SW; ++q) uq[q] = q < lim ? uq[q] : (T)0;
    const T piv = uq[K];
    if (piv == (T)0) return;
#pragma clang loop unroll(full)
    for (int r = 0; r < R; ++r) {
      const bool below = g * RPW + r * NT + t > c;
      const T l = a[r][K] / piv;
      const T le = below ? l : (T)0;
      a[r][K] = below ? l : a[r][K];
#pragma clang loop unroll(full)
      for (int q = K + 1; q < SW; ++q) a[r][q] -= le * uq[q];
    }
  };
  using K0 = std::integral_constant<int, 0>;
  using K1 = std::integral_constant<int, 1>;
  using K2 = std::integral_constant<int, 2>;
  using K3 = std::integral_constant<int, 3>;

  auto column = [&](const int c, const int lim, auto kc) {
    constexpr int kk = decltype(kc)::value;
    {
      const int par = c & 1;
      T bv = (T)0;
      int bi = -1;
      if (PIV) {
        candidate(c, kc, bv, bi);
        wave_argmax(bv, bi);
        if (lane == 0) { s_val[par][wave] = bv; s_idx[par][wave] = bi; }
      }
#pragma clang loop unroll(full)
      for (int r = 0; r < R; ++r) {
        const int li = g * RPW + r * NT + t;
        if (PIV && bi >= 0 && li == bi) {
#pragma clang loop unroll(full)
          for (int q = 0; q < SW; ++q) s_wcand[par][wave][q] = a[r][q];
        }
        if (li == c) {
#pragma clang loop unroll(full)
          for (int q = 0; q < SW; ++q) s_top[par][q] = a[r][q];
        }
      }
      __syncthreads();
      // ---- this workgroup's best over its waves ----
      int bw = 0;
      if (PIV) {
        bv = s_val[par][0];
        bi = s_idx[par][0];
#pragma clang loop unroll(full)
        for (int k = 1; k < NW; ++k)
          if (better(s_val[par][k], s_idx[par][k], bv, bi)) {
            bv = s_val[par][k]; bi = s_idx[par][k]; bw = k;
          }
      }
      int p;
      const T *u;    // the pivot row (LDS)
      const T *top;  // row c before the swap (LDS)
      if (!MULTI) {
        p = (PIV && bi >= 0) ? bi : c;
        top = s_top[par];
        u = (p == c) ? top : s_wcand[par][bw];
      } else {
        const unsigned long long sq = (unsigned long long)(seq0 + (unsigned)c + 1u) << 32;
        if (wave == 0) {
          // publish: lane k writes word k of this workgroup's record (and
          // of the top row in workgroup 0), straight from LDS
          unsigned long long *mine = recs + ((long)par * GMAX + g) * RW;
          unsigned *own = &s_rec[par][g][0];
          for (int k = lane; k < (g == 0 ? L1 + CW : L1); k += 64) {
            unsigned pl;
            if (k < CW) pl = bi >= 0 ? row_word(s_wcand[par][bw], k) : 0u;
            else if (k < CW + VW) pl = value_word(bv, k - CW);
            else if (k == CW + VW) pl = (unsigned)bi;
            else pl = row_word(s_top[par], k - L1);
            const int kw = k < L1 ? k : TOPW + (k - L1);
            st_agent(mine + kw, sq | pl);
            own[kw] = pl;  // this workgroup's own record, without the round trip
          }
        }
        if (poller) {
          // poll the other records (and workgroup 0's top row) until every
          // word carries this column's sequence number.  Bounded: after
          // SPIN_MAX rounds the wave raises the factorization's abort word
          // and goes on with what it has; every poll of every strip of the
          // same factorization then stops at its first incomplete round, so
          // a lost record costs one bound per factorization, not one per
          // column (the factorization then reports info = -1).
          const unsigned long long *prec = recs + (long)par * GMAX * RW;
          unsigned long long wv[CHW];
          bool ok = false;
          for (int spin = 0; !ok; ++spin) {
#pragma clang loop unroll(full)
            for (int i = 0; i < CHW; ++i) wv[i] = poff[i] >= 0 ? ld_agent(prec + poff[i]) : sq;
            const bool aborted = __any(ld_agent(abortw) != 0u);
            ok = true;
#pragma clang loop unroll(full)
            for (int i = 0; i < CHW; ++i) ok = ok && (wv[i] & 0xffffffff00000000ull) == sq;
            ok = __all(ok);
            if (!ok && (aborted || spin >= SPIN_MAX)) {
              if (lane == 0 && !aborted) st_agent(abortw, 1u);
              ok = true;
            }
            if (!ok) __builtin_amdgcn_s_sleep(1);
          }
          unsigned *srec = &s_rec[par][0][0];
#pragma clang loop unroll(full)
          for (int i = 0; i < CHW; ++i)
            if (poff[i] >= 0) srec[poff[i]] = (unsigned)wv[i];
        }
        __syncthreads();
        // the pivot: the best of the G records (lane k holds record k; every
        // wave decides the same way from the same data)
        T cv = (T)0;
        int ci = -1;
        if (lane < G) {
          ci = (int)s_rec[par][lane][CW + VW];
          cv = ci >= 0 ? *reinterpret_cast<const T *>(&s_rec[par][lane][CW]) : (T)0;
        }
        const int my_row = lane < G ? ci : -1;
        wave_argmax(cv, ci);
        // rows are distinct across workgroups: one record holds row ci
        const unsigned long long hold = __ballot(ci >= 0 && my_row == ci);
        p = ci < 0 ? c : ci;
        top = reinterpret_cast<const T *>(&s_rec[par][0][TOPW]);
        const int holder = hold ? __ffsll((long long)hold) - 1 : 0;
        u = (p == c) ? top : reinterpret_cast<const T *>(&s_rec[par][holder][0]);
      }

      // pivots go to global memory after the loop: a store here would make
      // the next barrier wait for its acknowledgement from memory
      // ---- row exchange inside the strip, then the elimination ----
      if constexpr (NT <= 256 || (R <= 4 && R * SW * sizeof(T) <= 512)) {
        // the pivot row in registers, read from LDS once for the new row c
        // and the update (the wave holding row c then needs no LDS read of
        // its own: 1000 of the 3800 cycles of a column at m = 1024 were
        // that swap).  512-thread workgroups (two waves per SIMD, at most
        // 256 VGPRs) have the registers for it with 4 rows per thread of at
        // most 512 bytes: FP32 strips of 32 columns, 7.15 -> 6.53 us per
        // column at 65536 rows; FP64 strips of 16, 7.50 -> 7.26 (with 8 or
        // 16 rows per thread the row would spill).
        // Two loops for the swap: as one if/else the compiler turns the two
        // copies into one through a pointer that selects the source, which
        // puts the register row in scratch memory.  (A compile-time index
        // for |pivot|: a captured constexpr would index the row at run time.)
        T ur[SW];
#pragma clang loop unroll(full)
        for (int q = 0; q < SW; ++q) ur[q] = u[q];
        if (t == 0) { s_piv[c] = p; s_pmag[c] = fabs(ur[decltype(kc)::value]); }
        if (p != c) {
#pragma clang loop unroll(full)
          for (int r = 0; r < R; ++r) {
            if (g * RPW + r * NT + t == c) {
#pragma clang loop unroll(full)
              for (int q = 0; q < SW; ++q) a[r][q] = ur[q];
            }
          }
#pragma clang loop unroll(full)
          for (int r = 0; r < R; ++r) {
            if (g * RPW + r * NT + t == p) {
#pragma clang loop unroll(full)
              for (int q = 0; q < SW; ++q) a[r][q] = top[q];
            }
          }
        }
        eliminate(c, lim, ur, kc);
      } else {
        if (t == 0) { s_piv[c] = p; s_pmag[c] = fabs(u[kk]); }
        if (p != c) {
#pragma clang loop unroll(full)
          for (int r = 0; r < R; ++r) {
            const int li = g * RPW + r * NT + t;
            if (li == c) {
#pragma clang loop unroll(full)
              for (int q = 0; q < SW; ++q) a[r][q] = u[q];
            } else if (li == p) {
#pragma clang loop unroll(full)
              for (int q = 0; q < SW; ++q) a[r][q] = top[q];
            }
          }
        }
        eliminate(c, lim, u, kc);
      }
    }
  };
#pragma clang loop unroll(disable)
  for (int base = 0; base < w; base += 4) {
    const int lim = w - base;
    column(base + 0, lim, K0{});
    if (base + 1 < w) column(base + 1, lim, K1{});
    if (base + 2 < w) column(base + 2, lim, K2{});
    if (base + 3 < w) column(base + 3, lim, K3{});
    rotate4();
  }
#pragma clang loop unroll(disable)
  for (int base = (w + 3) / 4 * 4; base < SW; base += 4) rotate4();

#pragma clang loop unroll(full)
  for (int r = 0; r < R; ++r) {
    const int li = g * RPW + r * NT + t;
    if (li < m) {
#pragma clang loop unroll(full)
      for (int q = 0; q < SW; ++q)
        if (q < w) A[(long)(j0 + li) * lda + j0 + q] = a[r][q];
    }
  }
  __syncthreads();  // s_piv complete
  if (g == 0 && t < w) ipiv[j0 + t] = j0 + s_piv[t];
  if (g == 0 && t == 0) {
    for (int c = 0; c < w; ++c)
      if (s_pmag[c] <= tiny) {
        if (*info == 0) *info = j0 + c + 1;
        break;
      }
  }

  // ---- the strip's interchanges on the other columns of the panel
  //      (the reference exchanges whole b-wide block rows, lu.C:166-177) ----
  // Each candidate position (the strip's rows and their pivot rows) finds
  // the row that ends up there by walking the swaps backwards.
  __shared__ int s_pos[2 * SW], s_src[2 * SW];
  __shared__ int s_nmv;
  if (t < 2 * SW) {
    int x = -1;
    if (t < w) {
      x = t;
    } else if (t >= SW && t - SW < w) {
      x = s_piv[t - SW];
      bool dup = x < w;
      for (int j = 0; j < t - SW; ++j) dup = dup || s_piv[j] == x;
      if (dup) x = -1;
    }
    int y = x;
    for (int j = w - 1; j >= 0 && x >= 0; --j) {
      const int pj = s_piv[j];
      y = (y == j) ? pj : (y == pj ? j : y);
    }
    const bool mv = x >= 0 && y != x;
    const unsigned long long mask = __ballot(mv);  // the 32 candidates are in wave 0
    if (mv) {
      const int off = __popcll(mask & ((1ull << t) - 1ull));
      s_pos[off] = x;
      s_src[off] = y;
    }
    if (t == 0) s_nmv = __popcll(mask);
  }
  __syncthreads();
  const int nmv = s_nmv;
  const int left = j0 - pc0, right = pc1 - (j0 + w);
  // columns of the rest of the panel, spread over all threads of the grid
  for (int k = g * NT + t; nmv > 0 && k < left + right; k += G * NT) {
    const int col = k < left ? pc0 + k : j0 + w + (k - left);
    T v[2 * SW];
#pragma clang loop unroll(full)
    for (int i = 0; i < 2 * SW; ++i)
      if (i < nmv) v[i] = A[(long)(j0 + s_src[i]) * lda + col];
#pragma clang loop unroll(full)
    for (int i = 0; i < 2 * SW; ++i)
      if (i < nmv) A[(long)(j0 + s_pos[i]) * lda + col] = v[i];
  }
}

// ---- launch shapes ---------------------------------------------------------
// A strip of m rows and sw columns runs as G workgroups of NT threads with R
// rows per thread (R * NT rows each).  G > 1 needs every workgroup resident
// at once (they exchange a record per column) and the polling waves must
// cover the G records in one round.  Candidates in order of preference
// (measured fastest first); the first that fits is taken.
namespace {
struct Cand { int R, NT; };
constexpr int poll_words(int NT) { return (NT / 64) * 64 * (NT >= 512 ? 4 : 5); }

template <typename T>
bool shape_fits(int m, int sw, Cand c, int max_groups, PanelShape *ps) {
  const int rows = c.R * c.NT;
  const int G = (m + rows - 1) / rows;
  const int vw = (int)sizeof(T) / 4, l1 = sw * vw + vw + 1;
  if (G > max_groups || G > GMAX) return false;
  if (G > 1 && G * l1 + sw * vw > poll_words(c.NT)) return false;
  // FP64 32-column strips between 513 and 1024 rows would need one
  // 512-thread workgroup: not instantiated (slower than two 16-column strips)
  if (sizeof(T) == 8 && sw == PW32 && G > 1 && m <= 1024) return false;
  ps->sw = sw;
  ps->R = c.R;
  ps->NT = c.NT;
  ps->G = G;
  return true;
}

template <typename T>
bool plan_sw(int m, int sw, int max_groups, bool pivot, PanelShape *ps) {
  if (m <= 0) return false;
  if (!pivot) {  // one workgroup (the unpivoted strip): 256 threads, 512 above 1024 rows
    if (m > 2048 || sw == PW8 || (sizeof(T) == 8 && sw == PW32)) return false;
    if (m > 1024) return shape_fits<T>(m, sw, Cand{4, 512}, 1, ps);
    return shape_fits<T>(m, sw, Cand{m <= 256 ? 1 : m <= 512 ? 2 : 4, 256}, 1, ps);
  }
  if (sw == PW8) {  // tall strips only
    return shape_fits<T>(m, sw, Cand{sizeof(T) == 8 ? 8 : 16, 512}, max_groups, ps);
  }
  if (sizeof(T) == 8 && sw == PW32)
    return shape_fits<T>(m, sw, Cand{2, 256}, max_groups, ps);
  // one workgroup up to 2048 rows (256 threads up to 1024, then 512)
  if (m <= 256) return shape_fits<T>(m, sw, Cand{1, 256}, 1, ps);
  if (m <= 512) return shape_fits<T>(m, sw, Cand{2, 256}, 1, ps);
  if (m <= 1024) return shape_fits<T>(m, sw, Cand{4, 256}, 1, ps);
  if (m <= 2048) return shape_fits<T>(m, sw, Cand{4, 512}, 1, ps);
  // taller: 256-thread workgroups of 1024 rows while they fit, else
  // 512-thread ones of 2048 rows, else (FP32, 16 columns) 4096 rows
  if (shape_fits<T>(m, sw, Cand{4, 256}, max_groups, ps)) return true;
  if (shape_fits<T>(m, sw, Cand{4, 512}, max_groups, ps)) return true;
  return sizeof(T) == 4 && sw == PW && shape_fits<T>(m, sw, Cand{8, 512}, max_groups, ps);
}
}  // namespace

template <typename T>
int panel_strip_width(int m, int max_groups, bool pivot) {
  PanelShape ps;
  // the measured policy for 32 columns (FP32 up to 65536 rows; FP64 up to
  // 512 and from 1025 to 9216), then 16, then the 8-column fallback
  const bool want32 = sizeof(T) == 4 ? m <= 32 * 2048 : (m <= 512 || (m > 1024 && m <= 18 * 512));
  if (want32 && plan_sw<T>(m, PW32, max_groups, pivot, &ps)) return PW32;
  if (plan_sw<T>(m, PW, max_groups, pivot, &ps)) return PW;
  if (plan_sw<T>(m, PW8, max_groups, pivot, &ps)) return PW8;
  return 0;
}
template int panel_strip_width<double>(int, int, bool);
template int panel_strip_width<float>(int, int, bool);

template <typename T>
int panel_plan(int m, int w, int max_groups, bool pivot, PanelShape *ps) {
  if (m <= 0 || w <= 0 || w > PW32 || !ps) return -1;
  const int widest = panel_strip_width<T>(m, max_groups, pivot);
  if (widest == 0 || w > widest) return -1;
  const int sw = w <= PW8 && widest == PW8 ? PW8 : (w <= PW ? PW : PW32);
  if (!plan_sw<T>(m, sw, max_groups, pivot, ps)) return -1;
  return CLU_OK;
}
template int panel_plan<double>(int, int, int, bool, PanelShape *);
template int panel_plan<float>(int, int, int, bool, PanelShape *);

// Right-looking substitution for x U = a with the leaf's U in LDS (D): x_k
// /= u_kk, then x_j -= x_k u_kj, row k of U read as whole 16-byte pieces
// from its first one holding j > k.
#define CLU_LEAF_SOLVE(x)                                                       \
  _Pragma("unroll") for (int k = 0; k < W; ++k) {                               \
    constexpr int V = 16 / sizeof(T);                                           \
    T uk[W];                                                                    \
    _Pragma("unroll") for (int q = k / V * V; q < W; q += V) {                  \
      typedef T vec_t __attribute__((ext_vector_type(V)));                      \
      const vec_t v = *reinterpret_cast<const vec_t *>(&D[k][q]);               \
      _Pragma("unroll") for (int e = 0; e < V; ++e) uk[q + e] = v[e];           \
    }                                                                           \
    if (uk[k] != (T)0) x[k] = x[k] / uk[k];                                     \
    _Pragma("unroll") for (int j = k + 1; j < W; ++j) x[j] -= x[k] * uk[j];     \
  }

// Leaf of the recursive panel without interchanges: rows [j0, j0 + m) x
// columns [j0, j0 + w), w <= 32.  Without pivoting only the w x w diagonal
// block needs the column-by-column elimination; every row below is the
// solution x of x U = a with that block's U, independent of the others.
// One workgroup of 1024 threads: wave 0 eliminates the diagonal block in
// LDS (a row per lane, no workgroup barrier per column), then every thread
// solves rows below (one per thread, U read from LDS; the strip staged
// through LDS, 1024 rows per pass: 16-byte aligned rows only), and the
// diagonal block goes back with the pivots (the identity) and the first
// tiny pivot.
// (The exchanging strip, used for this before, spends about 2 us per
// column on a workgroup barrier and the pivot row through LDS for every
// row: 70 us for 32 columns of a 1024-row leaf.)  As in that strip, a zero
// pivot leaves its column as it is.
template <typename T>
__global__ __launch_bounds__(1024) void leaf_nopiv_kernel(T *__restrict__ A, long lda, int j0,
                                                          int m, int w, int *__restrict__ ipiv,
                                                          int *info) {
  constexpr int W = 32, SP = W + 1;  // SP: row stride of the staged strip
  static_assert(sizeof(T) == 4, "the staged strip holds 1024 rows of 32 floats");
  __shared__ __attribute__((aligned(16))) T D[W][W];
  __shared__ T S[1024 * SP];
  const int t = threadIdx.x;
  for (int e = t; e < W * W; e += 1024) {
    const int i = e / W, j = e % W;
    D[i][j] = (i < w && j < w) ? A[(long)(j0 + i) * lda + j0 + j] : (i == j ? (T)1 : (T)0);
  }
  __syncthreads();
  if (t < 64) {
    // wave 0: lane i < W holds row i; at column c lane c publishes its
    // (finished) row c, every lane below it reads it back and eliminates
    const int i = t;
    T r[W];
#pragma unroll
    for (int j = 0; j < W; ++j) r[j] = i < W ? D[i][j] : (T)0;
#pragma unroll
    for (int c = 0; c < W; ++c) {
      if (i == c) {
#pragma unroll
        for (int j = 0; j < W; ++j) D[c][j] = r[j];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      T u[W];
#pragma unroll
      for (int j = c; j < W; ++j) u[j] = D[c][j];
      if (i > c && i < W && u[c] != (T)0) {
        const T l = r[c] / u[c];
        r[c] = l;
#pragma unroll
        for (int j = c + 1; j < W; ++j) r[j] -= l * u[j];
      }
    }
    __builtin_amdgcn_wave_barrier();
    if (i < W) {
#pragma unroll
      for (int j = 0; j < W; ++j) D[i][j] = r[j];
    }
  }
  __syncthreads();
  typedef T vec4_t __attribute__((ext_vector_type(4)));
  {
    // rows below the block in passes of up to 1024 (the caller guarantees
    // 16-byte aligned rows): the strip moves between memory and LDS in
    // 16-byte pieces, 8 lanes per 128-byte row (a lane reading its own row
    // from memory touches 64 rows' cache lines per access: 47 us for a
    // 1024-row leaf against 28.5 staged), and each thread solves one row
    // from LDS (the odd row stride puts the lanes' rows in distinct banks)
    for (int r0 = w; r0 < m; r0 += 1024) {
      const int nr = min(1024, m - r0);
      for (int c = t; c < nr * 8; c += 1024) {
        const int rr = c >> 3, q = (c & 7) * 4;
        const T *src = A + (long)(j0 + r0 + rr) * lda + j0 + q;
        vec4_t v = {(T)0, (T)0, (T)0, (T)0};
        if (q + 4 <= w) {
          v = *reinterpret_cast<const vec4_t *>(src);
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (q + e < w) v[e] = src[e];
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) S[rr * SP + q + e] = v[e];
      }
      __syncthreads();
      if (t < nr) {
        __asm__ volatile("" ::: "memory");
        T x[W];
#pragma unroll
        for (int j = 0; j < W; ++j) x[j] = S[t * SP + j];
        CLU_LEAF_SOLVE(x)
#pragma unroll
        for (int j = 0; j < W; ++j) S[t * SP + j] = x[j];
      }
      __syncthreads();
      for (int c = t; c < nr * 8; c += 1024) {
        const int rr = c >> 3, q = (c & 7) * 4;
        T *dst = A + (long)(j0 + r0 + rr) * lda + j0 + q;
        const vec4_t v = {S[rr * SP + q], S[rr * SP + q + 1], S[rr * SP + q + 2],
                          S[rr * SP + q + 3]};
        if (q + 4 <= w) {
          *reinterpret_cast<vec4_t *>(dst) = v;
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (q + e < w) dst[e] = v[e];
        }
      }
      __syncthreads();  // before the next pass refills S
    }
  }
  for (int e = t; e < w * w; e += 1024) {
    const int i = e / w, j = e % w;
    A[(long)(j0 + i) * lda + j0 + j] = D[i][j];
  }
  if (t < w) ipiv[j0 + t] = j0 + t;
  if (t == 0) {
    const T tiny = (T)100 * (sizeof(T) == 8 ? (T)2.220446049250313e-16 : (T)1.1920929e-07);
    for (int c = 0; c < w; ++c)
      if (fabs(D[c][c]) <= tiny) {
        if (*info == 0) *info = j0 + c + 1;
        break;
      }
  }
}

#undef CLU_LEAF_SOLVE

template <typename T>
int panel_base(T *A, long lda, int j0, int m_end, int w, int pc0, int pc1, int *ipiv,
               PanelWork &pw, hipStream_t s, bool pivot) {
  const int m = m_end - j0;
  if (m <= 0 || w <= 0) return CLU_OK;
  if (m < w) return -4;  // row c of every column lies in the strip
  PanelShape ps;
  if (panel_plan<T>(m, w, pw.max_groups, pivot, &ps)) {
    set_error("no launch shape for a %d x %d panel strip with %d resident workgroups", m, w,
              pw.max_groups);
    return CLU_ERR_UNSUPPORTED;
  }
  const int G = ps.G;
  unsigned long long *recs = reinterpret_cast<unsigned long long *>(pw.slots);
  const unsigned seq0 = pw.epoch;
  bool launched = true;
#define CLU_PANEL(RR, NN, MU, PV, SWW)                                                         \
  hipLaunchKernelGGL((panel_base_kernel<T, RR, NN, MU, PV, SWW>), dim3(G), dim3(NN), 0, s,     \
                     A, lda, j0, m, w, pc0, pc1, ipiv, recs, seq0, pw.info, pw.abortw)
  // the instantiated shapes: (R, NT, multi-workgroup, pivoting, strip width)
  const bool multi = G > 1;
  if constexpr (sizeof(T) == 4) {  // (the FP64 factorization always exchanges)
    if (!pivot && w <= 32 && lda % 4 == 0 && j0 % 4 == 0 && ((uintptr_t)A & 15) == 0) {
      hipLaunchKernelGGL((leaf_nopiv_kernel<T>), dim3(1), dim3(1024), 0, s, A, lda, j0, m, w,
                         ipiv, pw.info);
      CLU_LAUNCHED("leaf_nopiv_kernel");
      pw.epoch += (unsigned)w;
      return CLU_OK;
    }
  }
  if (!pivot) {
    if (ps.sw == PW) {
      if (ps.NT == 512) CLU_PANEL(4, 512, false, false, PW);
      else if (ps.R == 1) CLU_PANEL(1, 256, false, false, PW);
      else if (ps.R == 2) CLU_PANEL(2, 256, false, false, PW);
      else CLU_PANEL(4, 256, false, false, PW);
    } else if constexpr (sizeof(T) == 4) {
      if (ps.NT == 512) CLU_PANEL(4, 512, false, false, PW32);
      else if (ps.R == 1) CLU_PANEL(1, 256, false, false, PW32);
      else if (ps.R == 2) CLU_PANEL(2, 256, false, false, PW32);
      else CLU_PANEL(4, 256, false, false, PW32);
    } else {
      launched = false;
    }
  } else if (ps.sw == PW8) {
    if (!multi) launched = false;
    else if constexpr (sizeof(T) == 8) CLU_PANEL(8, 512, true, true, PW8);
    else CLU_PANEL(16, 512, true, true, PW8);
  } else if (ps.sw == PW32) {
    if constexpr (sizeof(T) == 8) {
      if (multi) CLU_PANEL(2, 256, true, true, PW32);
      else CLU_PANEL(2, 256, false, true, PW32);
    } else {
      if (multi && ps.NT == 256) CLU_PANEL(4, 256, true, true, PW32);
      else if (multi) CLU_PANEL(4, 512, true, true, PW32);
      else if (ps.NT == 512) CLU_PANEL(4, 512, false, true, PW32);
      else if (ps.R == 1) CLU_PANEL(1, 256, false, true, PW32);
      else if (ps.R == 2) CLU_PANEL(2, 256, false, true, PW32);
      else CLU_PANEL(4, 256, false, true, PW32);
    }
  } else {  // 16 columns
    if (ps.R == 8) {
      if constexpr (sizeof(T) == 4) {
        if (multi) CLU_PANEL(8, 512, true, true, PW);
        else launched = false;
      } else {
        launched = false;
      }
    } else if (multi && ps.NT == 256) CLU_PANEL(4, 256, true, true, PW);
    else if (multi) CLU_PANEL(4, 512, true, true, PW);
    else if (ps.NT == 512) CLU_PANEL(4, 512, false, true, PW);
    else if (ps.R == 1) CLU_PANEL(1, 256, false, true, PW);
    else if (ps.R == 2) CLU_PANEL(2, 256, false, true, PW);
    else CLU_PANEL(4, 256, false, true, PW);
  }
#undef CLU_PANEL
  if (!launched) {
    set_error("panel strip shape (R %d, NT %d, G %d, width %d) not instantiated", ps.R, ps.NT, G,
              ps.sw);
    return CLU_ERR_UNSUPPORTED;
  }
  CLU_LAUNCHED("panel_base_kernel");
  pw.epoch += (unsigned)w;
  return CLU_OK;
}

template int panel_base<double>(double *, long, int, int, int, int, int, int *, PanelWork &, hipStream_t, bool);
template int panel_base<float>(float *, long, int, int, int, int, int, int *, PanelWork &, hipStream_t, bool);

// ---------------------------------------------------------------------------
// swaps -> moves
// ---------------------------------------------------------------------------
constexpr int MAXRUN = 256;

// The move of one thread of a run of swaps ipiv[k1, k2) (global rows).
// piv[0..MAXRUN) holds the run in LDS, padded past its end with k1 + j (a
// row swapped with itself).  Thread i < nk takes position k1 + i, thread
// MAXRUN + i the pivot row piv[i] if it lies below the run and appears there
// first; x is that position (-1: none) and y the old row that ends up at x,
// found by walking the swaps backwards.  Both scans are branch-free over
// four pivots per LDS read, so the reads pipeline instead of paying one LDS
// round trip per swap.
CLU_DEV void run_move(const int *piv, int k1, int k2, int t, int &x, int &y) {
  const int nk = k2 - k1;
  const int4 *p4 = reinterpret_cast<const int4 *>(piv);
  x = -1;
  if (t < nk) {
    x = k1 + t;
  } else if (t >= MAXRUN && t - MAXRUN < nk) {
    const int i = t - MAXRUN;
    x = piv[i];
    bool dup = x < k2;
    for (int j4 = 0; 4 * j4 < i; ++j4) {
      const int4 v = p4[j4];
      const int j = 4 * j4;
      dup = dup || (v.x == x) || (j + 1 < i && v.y == x) || (j + 2 < i && v.z == x) ||
            (j + 3 < i && v.w == x);
    }
    if (dup) x = -1;
  }
  int z = x;
#pragma unroll 4
  for (int j4 = MAXRUN / 4 - 1; j4 >= 0; --j4) {
    const int4 v = p4[j4];
    const int j = k1 + 4 * j4;
    z = (z == j + 3) ? v.w : (z == v.w ? j + 3 : z);
    z = (z == j + 2) ? v.z : (z == v.z ? j + 2 : z);
    z = (z == j + 1) ? v.y : (z == v.y ? j + 1 : z);
    z = (z == j) ? v.x : (z == v.x ? j : z);
  }
  y = z;
}

// 2*MAXRUN threads, one run: every moved position becomes a move {x, y}
// ("row x receives old row y"), compacted in thread order.
__global__ __launch_bounds__(2 * MAXRUN) void pivots_to_moves_kernel(
    const int *__restrict__ ipiv, int k1, int k2, int2 *__restrict__ moves,
    int *__restrict__ count) {
  __shared__ __attribute__((aligned(16))) int piv[MAXRUN];
  __shared__ int wave_cnt[2 * MAXRUN / 64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int nk = k2 - k1;
  if (t < MAXRUN) piv[t] = t < nk ? ipiv[k1 + t] : k1 + t;
  __syncthreads();
  int x, y;
  run_move(piv, k1, k2, t, x, y);
  const bool mv = x >= 0 && y != x;
  const unsigned long long mask = __ballot(mv);
  if (lane == 0) wave_cnt[wave] = __popcll(mask);
  __syncthreads();
  int base = 0;
  for (int k = 0; k < wave; ++k) base += wave_cnt[k];
  if (mv) moves[base + __popcll(mask & ((1ull << lane) - 1ull))] = make_int2(x, y);
  if (t == 0) {
    int tot = 0;
    for (int k = 0; k < 2 * MAXRUN / 64; ++k) tot += wave_cnt[k];
    *count = tot;
  }
}

int pivots_to_moves(const int *ipiv, int k1, int k2, int2 *moves, int *count,
                    hipStream_t s) {
  if (k2 - k1 > MAXRUN || k2 < k1) return -1;
  hipLaunchKernelGGL(pivots_to_moves_kernel, dim3(1), dim3(2 * MAXRUN), 0, s, ipiv, k1, k2,
                     moves, count);
  CLU_LAUNCHED("pivots_to_moves_kernel");
  return CLU_OK;
}

// One workgroup per chunk of CW columns (128 bytes of every row; 64 and 256
// measured slower: 1.7 and 2.1 TB/s over a step's interchanges at
// n = 32768 against 2.5): all moves'
// source values of the chunk are gathered into registers (up to 2*MAXRUN
// rows, NP per thread, every load issued before the first is used), then,
// once the whole workgroup holds them, written to the destination rows.
// Moves form cycles, so a row read by one thread is written by another:
// the barrier between the phases orders them; chunks are independent.
template <typename T, int CW>
__global__ __launch_bounds__(256) void apply_moves_kernel(
    const int2 *__restrict__ moves, const int *__restrict__ count, T *__restrict__ A,
    long lda, int c0, int c1, int skip0, int skip1, int inverse) {
  constexpr int RPP = 256 / CW;          // rows per pass
  constexpr int NP = 2 * MAXRUN / RPP;   // passes for the longest list
  __shared__ int2 s_mv[2 * MAXRUN];
  const int cnt = *count;
  const int t = threadIdx.x;
  for (int i = t; i < cnt; i += 256) s_mv[i] = moves[i];
  __syncthreads();
  const int left = skip0 - c0;  // columns before the skipped range
  const int logical = blockIdx.x * CW + t % CW;
  const int col = logical < left ? c0 + logical : logical - left + skip1;
  const bool colok = col < c1;
  const int r0 = t / CW;
  T v[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const int e = p * RPP + r0;
    if (colok && e < cnt) {
      const int2 m2 = s_mv[e];
      v[p] = A[(long)(inverse ? m2.x : m2.y) * lda + col];
    }
  }
  wait_all_mem();
  __syncthreads();
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const int e = p * RPP + r0;
    if (colok && e < cnt) {
      const int2 m2 = s_mv[e];
      A[(long)(inverse ? m2.y : m2.x) * lda + col] = v[p];
    }
  }
}

// Several runs' move lists in one launch: each workgroup applies them to
// its own columns in order (reverse order for the inverse permutation), a
// barrier between runs (its columns belong to it alone).  Kept apart from
// the one-run kernel, which needs fewer registers.
template <typename T, int CW>
__global__ __launch_bounds__(256) void apply_runs_kernel(
    const int2 *__restrict__ moves, const int *__restrict__ counts, int nruns, long mstride,
    T *__restrict__ A, long lda, int c0, int c1, int skip0, int skip1, int inverse) {
  constexpr int RPP = 256 / CW;          // rows per pass
  constexpr int NP = 2 * MAXRUN / RPP;   // passes for the longest list
  __shared__ int2 s_mv[2 * MAXRUN];
  const int t = threadIdx.x;
  const int left = skip0 - c0;  // columns before the skipped range
  const int logical = blockIdx.x * CW + t % CW;
  const int col = logical < left ? c0 + logical : logical - left + skip1;
  const bool colok = col < c1;
  const int r0 = t / CW;
  for (int i = 0; i < nruns; ++i) {
    const int run = inverse ? nruns - 1 - i : i;
    const int cnt = counts[run];
    const int2 *mv = moves + (long)run * mstride;
    if (i > 0) __syncthreads();  // the previous run's writes and list reads are done
    for (int k = t; k < cnt; k += 256) s_mv[k] = mv[k];
    __syncthreads();
    T v[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int e = p * RPP + r0;
      if (colok && e < cnt) {
        const int2 m2 = s_mv[e];
        v[p] = A[(long)(inverse ? m2.x : m2.y) * lda + col];
      }
    }
    wait_all_mem();
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int e = p * RPP + r0;
      if (colok && e < cnt) {
        const int2 m2 = s_mv[e];
        A[(long)(inverse ? m2.y : m2.x) * lda + col] = v[p];
      }
    }
    wait_all_mem();
  }
}

template <typename T>
int apply_moves(const int2 *moves, const int *count, int max_moves, T *A, long lda,
                int c0, int c1, int skip0, int skip1, bool inverse, hipStream_t s,
                int nruns, long run_stride) {
  // normalise the skipped range to lie inside [c0, c1]
  skip0 = max(c0, min(skip0, c1));
  skip1 = max(skip0, min(skip1, c1));
  const int ncols = (c1 - c0) - (skip1 - skip0);
  if (ncols <= 0 || max_moves <= 0 || nruns <= 0) return CLU_OK;
  constexpr int CW = 128 / (int)sizeof(T);
  if (max_moves > 2 * MAXRUN) return -1;
  if (nruns > 1 && run_stride < max_moves) return -1;
  const int nchunks = (ncols + CW - 1) / CW;
  if (nruns == 1)
    hipLaunchKernelGGL((apply_moves_kernel<T, CW>), dim3(nchunks), dim3(256), 0, s, moves, count,
                       A, lda, c0, c1, skip0, skip1, inverse ? 1 : 0);
  else
    hipLaunchKernelGGL((apply_runs_kernel<T, CW>), dim3(nchunks), dim3(256), 0, s, moves, count,
                       nruns, run_stride, A, lda, c0, c1, skip0, skip1, inverse ? 1 : 0);
  CLU_LAUNCHED("apply_moves_kernel");
  return CLU_OK;
}

// One workgroup per run of MAXRUN pivots of ipiv[k0, kend): run r covers
// [k0 + r * MAXRUN, ...), its move list (as pivots_to_moves_kernel computes
// it) at moves + r * 2 * MAXRUN.
__global__ __launch_bounds__(2 * MAXRUN) void pivots_to_moves_all_kernel(
    const int *__restrict__ ipiv, int k0, int kend, int2 *__restrict__ moves,
    int *__restrict__ counts) {
  __shared__ __attribute__((aligned(16))) int piv[MAXRUN];
  __shared__ int wave_cnt[2 * MAXRUN / 64];
  const int r = blockIdx.x, k1 = k0 + r * MAXRUN, k2 = min(kend, k1 + MAXRUN), nk = k2 - k1;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  if (t < MAXRUN) piv[t] = t < nk ? ipiv[k1 + t] : k1 + t;
  __syncthreads();
  int x, y;
  run_move(piv, k1, k2, t, x, y);
  const bool mv = x >= 0 && y != x;
  const unsigned long long mask = __ballot(mv);
  if (lane == 0) wave_cnt[wave] = __popcll(mask);
  __syncthreads();
  int base = 0;
  for (int k = 0; k < wave; ++k) base += wave_cnt[k];
  int2 *out = moves + (long)r * 2 * MAXRUN;
  if (mv) out[base + __popcll(mask & ((1ull << lane) - 1ull))] = make_int2(x, y);
  if (t == 0) {
    int tot = 0;
    for (int k = 0; k < 2 * MAXRUN / 64; ++k) tot += wave_cnt[k];
    counts[r] = tot;
  }
}

int pivots_to_moves_all(const int *ipiv, int k1, int k2, int2 *moves, int *counts,
                        hipStream_t s) {
  if (k2 <= k1) return CLU_OK;
  const int nruns = (k2 - k1 + MAXRUN - 1) / MAXRUN;
  hipLaunchKernelGGL(pivots_to_moves_all_kernel, dim3(nruns), dim3(2 * MAXRUN), 0, s, ipiv, k1,
                     k2, moves, counts);
  CLU_LAUNCHED("pivots_to_moves_all_kernel");
  return CLU_OK;
}

// The runs' move lists applied to one vector, in order (reverse order and
// inverted moves for the inverse permutation), by one workgroup: all of a
// run's sources are read before any destination is written; the next run's
// moves are fetched while the current one is applied.
template <typename T>
__global__ __launch_bounds__(2 * MAXRUN) void permute_vec_kernel(const int2 *__restrict__ moves,
                                                                 const int *__restrict__ counts,
                                                                 int nruns, T *__restrict__ v,
                                                                 int inverse) {
  const int t = threadIdx.x;
  auto fetch = [&](int i, int2 &m, int &c) {
    const int r = inverse ? nruns - 1 - i : i;
    c = counts[r];
    m = t < c ? moves[(long)r * 2 * MAXRUN + t] : make_int2(0, 0);
  };
  int2 m;
  int c;
  fetch(0, m, c);
  for (int i = 0; i < nruns; ++i) {
    int2 mn = make_int2(0, 0);
    int cn = 0;
    if (i + 1 < nruns) fetch(i + 1, mn, cn);
    // forward: new[x] = old[y]; inverse: new[y] = old[x]
    T val = (T)0;
    if (t < c) val = v[inverse ? m.x : m.y];
    wait_all_mem();
    __syncthreads();  // every source of this run read
    if (t < c) v[inverse ? m.y : m.x] = val;
    wait_all_mem();
    __syncthreads();  // written before the next run reads
    m = mn;
    c = cn;
  }
}

template <typename T>
int permute_vec(const int *ipiv, int n, T *v, bool inverse, int2 *moves, int *counts,
                hipStream_t s) {
  if (n <= 0) return CLU_OK;
  int rc = pivots_to_moves_all(ipiv, 0, n, moves, counts, s);
  if (rc) return rc;
  const int nruns = (n + MAXRUN - 1) / MAXRUN;
  hipLaunchKernelGGL((permute_vec_kernel<T>), dim3(1), dim3(2 * MAXRUN), 0, s, moves, counts,
                     nruns, v, inverse ? 1 : 0);
  CLU_LAUNCHED("permute_vec_kernel");
  return CLU_OK;
}
template int permute_vec<double>(const int *, int, double *, bool, int2 *, int *, hipStream_t);
template int permute_vec<float>(const int *, int, float *, bool, int2 *, int *, hipStream_t);

template int apply_moves<double>(const int2 *, const int *, int, double *, long, int, int,
                                 int, int, bool, hipStream_t, int, long);
template int apply_moves<float>(const int2 *, const int *, int, float *, long, int, int,
                                int, int, bool, hipStream_t, int, long);

}  // namespace clu
