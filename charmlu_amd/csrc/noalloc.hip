// noalloc.hip -- the FP8 block op of charmlu_gemm_fp8_dev on a fixed,
// statically sized device scratch array, for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// The quantized operands (M x Kp and N x Kp bytes of e4m3), their per-tile
// power-of-two exponents and the quantizer's partial maxima live in
// `scratch`, a file-scope device
// array of NOALLOC_SCRATCH_BYTES that every device holding the library
// keeps until the process ends.  In host code `scratch` names a host-side
// stand-in, not device memory: its device address is looked up once per
// device with hipGetSymbolAddress and only that address is passed on.
// Every call on one device shares that device's array and keeps the last
// call's contents; calls that overlap in time on one device race.  Nothing
// here needs to start at zero: the quantizers write every byte the GEMM
// reads (the K padding included) and every exponent.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/charmlu_amd.h"
#include "common.h"
#include "kernels.h"

namespace clu {

constexpr size_t NOALLOC_SCRATCH_BYTES = CHARMLU_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[64];

// The device address of this device's scratch array (looked up once per
// device, with that device current).
static int noalloc_scratch(void **out) {
  int dev = 0;
  CLU_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    set_error("device %d out of range", dev);
    return CLU_ERR_HIP;
  }
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void *p = nullptr;
    CLU_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error("hipGetSymbolAddress gave no address for the scratch array");
      return CLU_ERR_HIP;
    }
    g_noalloc_ptr[dev] = p;
  }
  *out = g_noalloc_ptr[dev];
  return CLU_OK;
}

}  // namespace clu

using namespace clu;

extern "C" int charmlu_gemm_fp8_dev_noalloc(int M, int N, int K, float alpha, const float *dA,
                                            long lda, const float *dB, long ldb, float beta,
                                            float *dC, long ldc, void *stream) {
  if (M < 0) return -1;
  if (N < 0) return -2;
  if (K < 0) return -3;
  if (lda < (K > 1 ? K : 1)) return -6;
  if (ldb < (N > 1 ? N : 1)) return -8;
  if (ldc < (N > 1 ? N : 1)) return -11;
  if (M == 0 || N == 0) return CLU_OK;
  hipStream_t s = (hipStream_t)stream;
  if (K == 0 || alpha == 0.f) {
    // C = beta C (no operands to quantize)
    return sgemm(false, false, M, N, 0, 0.0f, dA, lda, dB, ldb, beta, dC, ldc, s);
  }
  const long Kp = (K + 63) / 64 * 64;
  const long tm = (M + 127) / 128, tn = (N + 127) / 128;
  // per 128-row tile of either operand: one exponent and the partial maxima
  // of the quantizer
  const size_t per_tile = sizeof(int) + sizeof(float) * FP8_PART_PER_TILE;
  const size_t need = (size_t)(M + (long)N) * Kp + per_tile * (size_t)(tm + tn);
  if (need > NOALLOC_SCRATCH_BYTES) {
    set_error("charmlu_gemm_fp8_dev_noalloc: (M + N) * ceil64(K) + %zu * (ceil(M/128) + "
              "ceil(N/128)) = %zu bytes exceeds the static scratch of %zu bytes",
              per_tile, need, (size_t)NOALLOC_SCRATCH_BYTES);
    return CLU_ERR_UNSUPPORTED;
  }
  void *base = nullptr;
  int rc = noalloc_scratch(&base);
  if (rc) return rc;
  uint8_t *qa = (uint8_t *)base, *qb = qa + (size_t)M * Kp;
  int *ea = (int *)(qb + (size_t)N * Kp), *eb = ea + tm;
  float *part = (float *)(eb + tn);
  if ((rc = quant_rows_fp8(M, K, (int)Kp, 1.0f, dA, lda, qa, ea, part, s))) return rc;
  if ((rc = quant_cols_fp8(K, N, (int)Kp, dB, ldb, qb, eb, part + FP8_PART_PER_TILE * tm, s)))
    return rc;
  return fgemm_nt(M, N, (int)Kp, alpha, qa, Kp, ea, qb, Kp, eb, beta, dC, ldc, s);
}
