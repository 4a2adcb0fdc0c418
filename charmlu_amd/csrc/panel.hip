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
// holds rows g*256*R + r*256 + t (r < R), 16 values each.  With one
// workgroup a column costs two LDS barriers.  With G > 1 workgroups each
// column needs one grid-wide exchange: every workgroup publishes its best
// candidate (value, row, and the row's 16 values) and the owner of the
// current top row publishes that row, into a double-buffered slot array
// written with relaxed agent-scope atomics; a counter barrier follows; every
// workgroup then reduces the G candidates itself (same data, same order, so
// all agree on the pivot) and swaps/eliminates its own rows.
//
// pivots_to_moves / apply_moves replace the reference's pivot agglomeration
// and cycle application (recordPivot / announceAgglomeratedPivots /
// sendPendingPivots, lu.C:46-163, lu.C:300-352): a run of sequential swaps
// becomes a list of (destination row, source row) moves, applied to whole
// column ranges as a gather through LDS (one read and one write per moved
// row segment, coalesced along the row).
#include "common.h"
#include "kernels.h"

namespace clu {

constexpr int PW = 16;  // max strip width

template <typename T>
struct SlotView {
  T *val;   // [2][G]
  int *idx; // [2][G]
  T *row;   // [2][G][PW]
  T *top;   // [2][PW]
};

template <typename T, int R>
__global__ __launch_bounds__(256) void panel_base_kernel(
    T *__restrict__ A, long lda, int j0, int m, int w, int *__restrict__ ipiv,
    T *slot_val, int *slot_idx, T *slot_row, T *top_row, unsigned *counter,
    unsigned epoch, int *info) {
  const int G = gridDim.x, g = blockIdx.x, t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  constexpr int RPW = 256 * R;
  const T tiny = (T)100 * (sizeof(T) == 8 ? (T)2.220446049250313e-16 : (T)1.1920929e-07);

  __shared__ T s_val[4];
  __shared__ int s_idx[4];
  __shared__ T s_u[PW];
  __shared__ T s_top[PW];
  __shared__ int s_p;

  T a[R][PW];
#pragma clang loop unroll(full)
  for (int r = 0; r < R; ++r) {
    const int li = g * RPW + r * 256 + t;
#pragma clang loop unroll(full)
    for (int q = 0; q < PW; ++q)
      a[r][q] = (li < m && q < w) ? A[(long)(j0 + li) * lda + j0 + q] : (T)0;
  }

  unsigned target = epoch;
#pragma clang loop unroll(full)
  for (int c = 0; c < PW; ++c) {
    // a constant trip count keeps the loop unrollable, so every a[r][c]
    // below is a register; columns past w do nothing (w is uniform)
    if (c < w) {
    const int par = c & 1;
    // ---- candidate for column c over local rows >= c (strict '>', rows
    //      visited in increasing order, so the lowest row wins ties) ----
    T bv = (T)0;
    int bi = -1;
#pragma clang loop unroll(full)
    for (int r = 0; r < R; ++r) {
      const int li = g * RPW + r * 256 + t;
      const T v = fabs(a[r][c]);
      if (li < m && li >= c && v > bv) { bv = v; bi = li; }
    }
    wave_argmax(bv, bi);
    if (lane == 0) { s_val[wave] = bv; s_idx[wave] = bi; }
    __syncthreads();
    bv = s_val[0]; bi = s_idx[0];
#pragma clang loop unroll(full)
    for (int k = 1; k < 4; ++k)
      if (better(s_val[k], s_idx[k], bv, bi)) { bv = s_val[k]; bi = s_idx[k]; }

    const int owner_c = c / RPW;  // workgroup holding the top row
    if (G == 1) {
      const int p = bi < 0 ? c : bi;
#pragma clang loop unroll(full)
      for (int r = 0; r < R; ++r) {
        const int li = r * 256 + t;
        if (li == p) {
#pragma clang loop unroll(full)
          for (int q = 0; q < PW; ++q) s_u[q] = a[r][q];
        }
        if (li == c) {
#pragma clang loop unroll(full)
          for (int q = 0; q < PW; ++q) s_top[q] = a[r][q];
        }
      }
      if (t == 0) s_p = p;
      __syncthreads();
    } else {
      // publish this workgroup's candidate and, if held here, the top row
      if (t == 0) {
        st_agent(&slot_val[par * G + g], bv);
        st_agent(&slot_idx[par * G + g], bi);
      }
#pragma clang loop unroll(full)
      for (int r = 0; r < R; ++r) {
        const int li = g * RPW + r * 256 + t;
        if (bi >= 0 && li == bi) {
#pragma clang loop unroll(full)
          for (int q = 0; q < PW; ++q) st_agent(&slot_row[(par * G + g) * PW + q], a[r][q]);
        }
        if (li == c) {
#pragma clang loop unroll(full)
          for (int q = 0; q < PW; ++q) st_agent(&top_row[par * PW + q], a[r][q]);
        }
      }
      wait_all_mem();
      __syncthreads();
      target += G;
      if (t == 0) {
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (ld_agent(counter) - target > 0x80000000u) __builtin_amdgcn_s_sleep(1);
      }
      __syncthreads();
      // every workgroup reduces the same G candidates the same way
      T cv = (T)0;
      int ci = -1;
      if (t < G) { cv = ld_agent(&slot_val[par * G + t]); ci = ld_agent(&slot_idx[par * G + t]); }
      wave_argmax(cv, ci);
      if (lane == 0) { s_val[wave] = cv; s_idx[wave] = ci; }
      __syncthreads();
      cv = s_val[0]; ci = s_idx[0];
#pragma clang loop unroll(full)
      for (int k = 1; k < 4; ++k)
        if (better(s_val[k], s_idx[k], cv, ci)) { cv = s_val[k]; ci = s_idx[k]; }
      const int p = ci < 0 ? c : ci;
      if (t < w) {
        const int gw = p / RPW;
        s_u[t] = (p == c) ? ld_agent(&top_row[par * PW + t])
                          : ld_agent(&slot_row[(par * G + gw) * PW + t]);
        s_top[t] = ld_agent(&top_row[par * PW + t]);
      }
      if (t == 0) s_p = p;
      (void)owner_c;
      __syncthreads();
    }

    const int p = s_p;
    if (g == 0 && t == 0) {
      ipiv[j0 + c] = j0 + p;
      if (fabs(s_u[c]) <= tiny && *info == 0) *info = j0 + c + 1;
    }
    // ---- row exchange inside the strip ----
    if (p != c) {
#pragma clang loop unroll(full)
      for (int r = 0; r < R; ++r) {
        const int li = g * RPW + r * 256 + t;
        if (li == c) {
#pragma clang loop unroll(full)
          for (int q = 0; q < PW; ++q) a[r][q] = s_u[q];
        } else if (li == p) {
#pragma clang loop unroll(full)
          for (int q = 0; q < PW; ++q) a[r][q] = s_top[q];
        }
      }
    }
    // ---- multipliers (division) and rank-1 update of columns > c ----
    const T piv = s_u[c];
    if (piv != (T)0) {
#pragma clang loop unroll(full)
      for (int r = 0; r < R; ++r) {
        const int li = g * RPW + r * 256 + t;
        if (li < m && li > c) {
          const T l = a[r][c] / piv;
          a[r][c] = l;
#pragma clang loop unroll(full)
          for (int q = c + 1; q < PW; ++q) a[r][q] -= l * s_u[q];
        }
      }
    }
    __syncthreads();  // s_u / s_top / s_val reused by the next column
    }
  }

#pragma clang loop unroll(full)
  for (int r = 0; r < R; ++r) {
    const int li = g * RPW + r * 256 + t;
    if (li < m) {
#pragma clang loop unroll(full)
      for (int q = 0; q < PW; ++q)
        if (q < w) A[(long)(j0 + li) * lda + j0 + q] = a[r][q];
    }
  }
}

template <typename T>
int panel_base(T *A, long lda, int j0, int m_end, int w, int *ipiv, PanelWork &pw,
               hipStream_t s) {
  const int m = m_end - j0;
  if (m <= 0 || w <= 0) return CLU_OK;
  if (w > PW) return -1;
  int R, G;
  if (m <= 256) { R = 1; G = 1; }
  else if (m <= 512) { R = 2; G = 1; }
  else if (m <= 1024) { R = 4; G = 1; }
  else { R = 8; G = (m + 2047) / 2048; }
  if (G > pw.max_groups) return -2;
  T *sv = reinterpret_cast<T *>(pw.slots);
  int *si = reinterpret_cast<int *>(sv + 2 * pw.max_groups);
  T *sr = reinterpret_cast<T *>(si + 2 * pw.max_groups);
  T *tr = reinterpret_cast<T *>(pw.top);
  // the slot arrays are laid out for max_groups; index them with G
  switch (R) {
    case 1: hipLaunchKernelGGL((panel_base_kernel<T, 1>), dim3(G), dim3(256), 0, s, A, lda, j0, m, w, ipiv, sv, si, sr, tr, pw.counter, pw.epoch, pw.info); break;
    case 2: hipLaunchKernelGGL((panel_base_kernel<T, 2>), dim3(G), dim3(256), 0, s, A, lda, j0, m, w, ipiv, sv, si, sr, tr, pw.counter, pw.epoch, pw.info); break;
    case 4: hipLaunchKernelGGL((panel_base_kernel<T, 4>), dim3(G), dim3(256), 0, s, A, lda, j0, m, w, ipiv, sv, si, sr, tr, pw.counter, pw.epoch, pw.info); break;
    default: hipLaunchKernelGGL((panel_base_kernel<T, 8>), dim3(G), dim3(256), 0, s, A, lda, j0, m, w, ipiv, sv, si, sr, tr, pw.counter, pw.epoch, pw.info); break;
  }
  CLU_LAUNCHED("panel_base_kernel");
  if (G > 1) pw.epoch += (unsigned)G * (unsigned)w;
  return CLU_OK;
}

template int panel_base<double>(double *, long, int, int, int, int *, PanelWork &, hipStream_t);
template int panel_base<float>(float *, long, int, int, int, int *, PanelWork &, hipStream_t);

// ---------------------------------------------------------------------------
// swaps -> moves
// ---------------------------------------------------------------------------
constexpr int MAXRUN = 256;

__global__ __launch_bounds__(64) void pivots_to_moves_kernel(
    const int *__restrict__ ipiv, int k1, int k2, int2 *__restrict__ moves,
    int *__restrict__ count) {
  __shared__ int cur[MAXRUN];
  __shared__ int ext_pos[MAXRUN];
  __shared__ int ext_src[MAXRUN];
  const int lane = threadIdx.x;
  const int nk = k2 - k1;
  for (int i = lane; i < nk; i += 64) cur[i] = k1 + i;
  int n_ext = 0;
  __syncthreads();
  for (int j = k1; j < k2; ++j) {
    const int p = ipiv[j];
    if (p == j) continue;
    if (p < k2) {
      if (lane == 0) {
        const int tmp = cur[j - k1];
        cur[j - k1] = cur[p - k1];
        cur[p - k1] = tmp;
      }
    } else {
      int found = -1;
      for (int base = 0; base < n_ext; base += 64) {
        const int e = base + lane;
        const bool hit = e < n_ext && ext_pos[e] == p;
        const unsigned long long mask = __ballot(hit);
        if (mask) { found = base + __ffsll((long long)mask) - 1; break; }
      }
      if (found < 0) {
        found = n_ext;
        if (lane == 0) { ext_pos[found] = p; ext_src[found] = p; }
        ++n_ext;
      }
      __syncthreads();
      if (lane == 0) {
        const int tmp = cur[j - k1];
        cur[j - k1] = ext_src[found];
        ext_src[found] = tmp;
      }
    }
    __syncthreads();
  }
  // emit moves: positions whose content changed
  int base = 0;
  for (int i0 = 0; i0 < nk + n_ext; i0 += 64) {
    const int i = i0 + lane;
    int dst = -1, src = -1;
    if (i < nk) { dst = k1 + i; src = cur[i]; }
    else if (i < nk + n_ext) { dst = ext_pos[i - nk]; src = ext_src[i - nk]; }
    const bool mv = dst >= 0 && dst != src;
    const unsigned long long mask = __ballot(mv);
    if (mv) {
      const int off = __popcll(mask & ((1ull << lane) - 1ull));
      moves[base + off] = make_int2(dst, src);
    }
    base += __popcll(mask);
  }
  if (lane == 0) *count = base;
}

int pivots_to_moves(const int *ipiv, int k1, int k2, int2 *moves, int *count,
                    hipStream_t s) {
  if (k2 - k1 > MAXRUN || k2 < k1) return -1;
  hipLaunchKernelGGL(pivots_to_moves_kernel, dim3(1), dim3(64), 0, s, ipiv, k1, k2,
                     moves, count);
  CLU_LAUNCHED("pivots_to_moves_kernel");
  return CLU_OK;
}

template <typename T, int CW>
__global__ __launch_bounds__(256) void apply_moves_kernel(
    const int2 *__restrict__ moves, const int *__restrict__ count, T *__restrict__ A,
    long lda, int c0, int c1, int skip0, int skip1, int inverse) {
  extern __shared__ unsigned char smem_raw[];
  T *buf = reinterpret_cast<T *>(smem_raw);
  const int cnt = *count;
  const int t = threadIdx.x;
  const int left = skip0 - c0;  // columns before the skipped range
  auto phys = [&](int logical) { return logical < left ? c0 + logical : logical - left + skip1; };
  const int chunk = blockIdx.x * CW;
  for (int e = t; e < cnt * CW; e += 256) {
    const int mv = e / CW, q = e % CW;
    const int col = phys(chunk + q);
    if (col < c1) {
      const int2 m2 = moves[mv];
      const int src = inverse ? m2.x : m2.y;
      buf[e] = A[(long)src * lda + col];
    }
  }
  __syncthreads();
  for (int e = t; e < cnt * CW; e += 256) {
    const int mv = e / CW, q = e % CW;
    const int col = phys(chunk + q);
    if (col < c1) {
      const int2 m2 = moves[mv];
      const int dst = inverse ? m2.y : m2.x;
      A[(long)dst * lda + col] = buf[e];
    }
  }
}

template <typename T>
int apply_moves(const int2 *moves, const int *count, int max_moves, T *A, long lda,
                int c0, int c1, int skip0, int skip1, bool inverse, hipStream_t s) {
  // normalise the skipped range to lie inside [c0, c1]
  skip0 = max(c0, min(skip0, c1));
  skip1 = max(skip0, min(skip1, c1));
  const int ncols = (c1 - c0) - (skip1 - skip0);
  if (ncols <= 0 || max_moves <= 0) return CLU_OK;
  constexpr int CW = 16;
  const int nchunks = (ncols + CW - 1) / CW;
  const size_t shm = (size_t)max_moves * CW * sizeof(T);
  hipLaunchKernelGGL((apply_moves_kernel<T, CW>), dim3(nchunks), dim3(256), shm, s,
                     moves, count, A, lda, c0, c1, skip0, skip1, inverse ? 1 : 0);
  CLU_LAUNCHED("apply_moves_kernel");
  return CLU_OK;
}

template int apply_moves<double>(const int2 *, const int *, int, double *, long, int, int,
                                 int, int, bool, hipStream_t);
template int apply_moves<float>(const int2 *, const int *, int, float *, long, int, int,
                                int, int, bool, hipStream_t);

}  // namespace clu
