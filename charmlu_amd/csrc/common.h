// common.h -- shared helpers for the charmlu_amd HIP kernels (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define CLU_DEV __device__ __forceinline__

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));

// Status codes of the C ABI (also in include/charmlu_amd.h).
#define CLU_OK 0
#define CLU_ERR_HIP (-100)
#define CLU_ERR_NOMEM (-101)
#define CLU_ERR_CONFIG (-102)
#define CLU_ERR_UNSUPPORTED (-103)

namespace clu {

// Record a HIP error as the library's last error; returns CLU_ERR_HIP.
int hip_fail(hipError_t e, const char *what, const char *file, int line);
void set_error(const char *fmt, ...);

}  // namespace clu

#define CLU_HIP(call)                                                       \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) return clu::hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)

#define CLU_LAUNCHED(what)                                                  \
  do {                                                                      \
    hipError_t e_ = hipGetLastError();                                      \
    if (e_ != hipSuccess) return clu::hip_fail(e_, what, __FILE__, __LINE__); \
  } while (0)

// Relaxed agent-scope atomics: coherent across the XCDs' L2 caches without
// the whole-L2 write-back / invalidate an agent-scope fence would cost.
template <typename T>
CLU_DEV void st_agent(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
CLU_DEV T ld_agent(const T *p) {
  return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// XOR swizzle of the 4 16-byte chunks of a 64-byte LDS operand row: chunk k
// of row r sits at chunk k ^ lds_swz(r).  A fragment read takes one chunk
// of 16 consecutive rows (16 bytes per lane), 8 lanes per LDS cycle:
// lds_swz gives every 8 consecutive rows of one parity, and every 4 rows of
// one residue modulo 4, four different chunks, so neither 8- nor 16-lane
// groups of a read share banks ((r >> 2) & 3 alone left rows r and r + 2 on
// the same banks: half the LDS cycles of the BF16 update, 40 % of the FP64
// one's, were bank conflicts).
CLU_DEV int lds_swz(int r) { return ((r >> 2) & 1) | ((((r >> 1) ^ (r >> 3)) & 1) << 1); }

// Wait until every outstanding memory operation of this thread completed.
CLU_DEV void wait_all_mem() { __builtin_amdgcn_s_waitcnt(0); }

// (|value|, row) argmax with the reference's semantics: strictly larger
// magnitude wins; among equal magnitudes the lowest row wins (the reference
// keeps the first one seen in row order inside a block, lu.C:197-205).
// A candidate is only ever formed from a nonzero value, so idx == -1 means
// "nothing found" and compares below every real candidate.
template <typename T>
CLU_DEV bool better(T av, int ai, T bv, int bi) {
  // is (av, ai) strictly better than (bv, bi)?  (branch-free)
  const bool tie_wins = ai >= 0 && (bi < 0 || ai < bi);
  return (av > bv) | ((av == bv) & tie_wins);
}

// One DPP lane move of a 64-bit value (two 32-bit halves); lanes whose
// source is out of range or whose row is masked off read 0.
template <int CTRL, int ROW_MASK>
CLU_DEV double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xF, true);
  return __hiloint2double(hi, lo);
}
template <int CTRL, int ROW_MASK>
CLU_DEV float dpp_move(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, true));
}

// Wave-wide maximum of non-negative values with DPP lane moves (row_shr 1,
// 2, 4, 8 leave each row's maximum in its lane 15; row_bcast 15 and 31 carry
// it across rows to lane 63), then read from lane 63.
template <typename T>
CLU_DEV T wave_max_nonneg(T v) {
  v = fmax(v, dpp_move<0x111, 0xF>(v));  // row_shr:1
  v = fmax(v, dpp_move<0x112, 0xF>(v));  // row_shr:2
  v = fmax(v, dpp_move<0x114, 0xF>(v));  // row_shr:4
  v = fmax(v, dpp_move<0x118, 0xF>(v));  // row_shr:8
  v = fmax(v, dpp_move<0x142, 0xA>(v));  // row_bcast:15 into rows 1, 3
  v = fmax(v, dpp_move<0x143, 0xC>(v));  // row_bcast:31 into rows 2, 3
  if (sizeof(T) == 8) {
    const int lo = __builtin_amdgcn_readlane(__double2loint((double)v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint((double)v), 63);
    return (T)__hiloint2double(hi, lo);
  }
  return (T)__int_as_float(__builtin_amdgcn_readlane(__float_as_int((float)v), 63));
}

// Wave-wide argmax of non-negative magnitudes: the maximum by DPP, the
// lanes holding it by ballot, and only on an exact tie a reduction for the
// lowest row.  (v, i) = (0, -1) means no candidate.
template <typename T>
CLU_DEV void wave_argmax(T &v, int &i) {
  const T m = wave_max_nonneg(v);
  const bool mine = i >= 0 && v == m;
  const unsigned long long b = __ballot(mine);
  if (b == 0ull) { v = (T)0; i = -1; return; }
  if (__popcll(b) == 1) {
    i = __builtin_amdgcn_readlane(i, __ffsll((long long)b) - 1);
    v = m;
    return;
  }
  int ci = mine ? i : 0x7fffffff;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) ci = min(ci, __shfl_xor(ci, off, 64));
  i = ci;
  v = m;
}

CLU_DEV double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
CLU_DEV double wave_max(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// atomic max for non-negative doubles via their ordered bit patterns.
CLU_DEV void atomic_max_nonneg(double *addr, double v) {
  unsigned long long *a = reinterpret_cast<unsigned long long *>(addr);
  atomicMax(a, (unsigned long long)__double_as_longlong(v));
}
