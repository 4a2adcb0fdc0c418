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

// Wait until every outstanding memory operation of this thread completed.
CLU_DEV void wait_all_mem() { __builtin_amdgcn_s_waitcnt(0); }

// (|value|, row) argmax with the reference's semantics: strictly larger
// magnitude wins; among equal magnitudes the lowest row wins (the reference
// keeps the first one seen in row order inside a block, lu.C:197-205).
// A candidate is only ever formed from a nonzero value, so idx == -1 means
// "nothing found" and compares below every real candidate.
template <typename T>
CLU_DEV bool better(T av, int ai, T bv, int bi) {
  // is (av, ai) strictly better than (bv, bi)?
  if (av > bv) return true;
  if (av < bv) return false;
  return ai >= 0 && (bi < 0 || ai < bi);
}

template <typename T>
CLU_DEV void wave_argmax(T &v, int &i) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    T ov = __shfl_xor(v, off, 64);
    int oi = __shfl_xor(i, off, 64);
    if (better(ov, oi, v, i)) { v = ov; i = oi; }
  }
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
