// misc.hip -- seeded generation, the scaled-residual check, conversions.
//
// gen_rhs restates MatGen (benchmark.C:27-74) as a jump-ahead: draw k of a
// Park-Miller stream seeded with s is s * 48271^k mod (2^31 - 1), so every
// entry of b is computed independently (genVec, benchmark.C:454-460: global
// row i holds draw i+1).
//
// gen_uniform is the library's own counter generator for A (the reference
// uses Charm++'s CrnDouble, benchmark.C:436-452): entry (i, j) is the
// splitmix64 output function of seed*phi ^ (i<<32 | j), top 53 bits scaled
// to [0,1), minus 0.5.  The CPU oracle computes the same bits.
//
// residual computes the maxima that the reference's validation gathers
// (recvXvec / sumBvec / calcResiduals, benchmark.C:485-573) in one pass over
// A: one wave per row, FP64 throughout.
#include "common.h"
#include "kernels.h"

namespace clu {

CLU_DEV uint64_t splitmix64_mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

CLU_DEV double gen_value(uint64_t seed, long i, long j) {
  const uint64_t key = (seed * 0x9E3779B97F4A7C15ULL) ^ (((uint64_t)i << 32) | (uint64_t)(uint32_t)j);
  const uint64_t u = splitmix64_mix(key) >> 11;
  return (double)u * 0x1.0p-53 - 0.5;
}

__global__ void gen_uniform_kernel(uint64_t seed, long row0, long col0, long m, long n,
                                   double *__restrict__ A, long lda) {
  const long total = m * n;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const long i = e / n, j = e % n;
    A[i * lda + j] = gen_value(seed, row0 + i, col0 + j);
  }
}

int gen_uniform(uint64_t seed, long row0, long col0, long m, long n, double *A, long lda,
                hipStream_t s) {
  if (m <= 0 || n <= 0) return CLU_OK;
  const long total = m * n;
  long blocks = (total + 255) / 256;
  if (blocks > 65536L * 16) blocks = 65536L * 16;
  hipLaunchKernelGGL(gen_uniform_kernel, dim3((unsigned)blocks), dim3(256), 0, s, seed, row0,
                     col0, m, n, A, lda);
  CLU_LAUNCHED("gen_uniform_kernel");
  return CLU_OK;
}

// one workgroup per row: write the row, then a_ii = sum_j |a_ij|
__global__ __launch_bounds__(256) void gen_diagdom_kernel(uint64_t seed, long n,
                                                          double *__restrict__ A, long lda) {
  __shared__ double part[4];
  const long i = blockIdx.x;
  double s = 0.0;
  for (long j = threadIdx.x; j < n; j += 256) {
    const double v = gen_value(seed, i, j);
    A[i * lda + j] = v;
    s += fabs(v);
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) A[i * lda + i] = (part[0] + part[1]) + (part[2] + part[3]);
}

int gen_diagdom(uint64_t seed, long n, double *A, long lda, hipStream_t s) {
  if (n <= 0) return CLU_OK;
  hipLaunchKernelGGL(gen_diagdom_kernel, dim3((unsigned)n), dim3(256), 0, s, seed, n, A, lda);
  CLU_LAUNCHED("gen_diagdom_kernel");
  return CLU_OK;
}

// MatGen draw k (k >= 1): seed * 48271^k mod (2^31 - 1)
CLU_DEV int matgen_draw(int seed, long k) {
  const uint64_t M = 2147483647ULL;
  uint64_t base = 48271ULL, acc = 1ULL;
  uint64_t e = (uint64_t)k;
  while (e) {
    if (e & 1ULL) acc = (acc * base) % M;
    base = (base * base) % M;
    e >>= 1;
  }
  return (int)(((uint64_t)(uint32_t)seed * acc) % M);
}

__global__ void gen_rhs_kernel(int seed, long skip, long n, double *__restrict__ b) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)matgen_draw(seed, skip + i + 1) / 2147483647 - 0.5;
}

int gen_rhs(int seed, long skip, long n, double *b, hipStream_t s) {
  if (n <= 0) return CLU_OK;
  hipLaunchKernelGGL(gen_rhs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, seed,
                     skip, n, b);
  CLU_LAUNCHED("gen_rhs_kernel");
  return CLU_OK;
}

__global__ void residual_init_kernel(double *out) {
  if (threadIdx.x < 5) out[threadIdx.x] = 0.0;
}

// 16 lanes per row (16 rows per workgroup); unit-stride 16-byte aligned
// rows take 16-byte loads with four partial sums, so each lane keeps four
// loads in flight
__global__ __launch_bounds__(256) void residual_kernel(int n, const double *__restrict__ A,
                                                       long lda, const double *__restrict__ x,
                                                       const double *__restrict__ b,
                                                       double *__restrict__ r,
                                                       double *__restrict__ out) {
  const long i = (long)blockIdx.x * 16 + (threadIdx.x >> 4);
  const int sub = threadIdx.x & 15;
  double s[4] = {0.0, 0.0, 0.0, 0.0}, as[4] = {0.0, 0.0, 0.0, 0.0}, am = 0.0;
  const bool vec = ((lda | n) & 1) == 0 && (((unsigned long)A | (unsigned long)x) & 15) == 0;
  if (i < n) {
    const double *Ai = A + i * lda;
    if (vec) {
      const double2 *a2 = reinterpret_cast<const double2 *>(Ai);
      const double2 *x2 = reinterpret_cast<const double2 *>(x);
      const int n2 = n / 2;
      int k = sub;
      for (; k + 48 < n2; k += 64) {
        double2 p[4], q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          p[u] = a2[k + 16 * u];
          q[u] = x2[k + 16 * u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          s[u] += p[u].x * q[u].x + p[u].y * q[u].y;
          as[u] += fabs(p[u].x) + fabs(p[u].y);
          am = fmax(am, fmax(fabs(p[u].x), fabs(p[u].y)));
        }
      }
      for (; k < n2; k += 16) {
        const double2 p0 = a2[k], q0 = x2[k];
        s[0] += p0.x * q0.x + p0.y * q0.y;
        as[0] += fabs(p0.x) + fabs(p0.y);
        am = fmax(am, fmax(fabs(p0.x), fabs(p0.y)));
      }
    } else {
      for (int j = sub; j < n; j += 16) {
        const double a = Ai[j];
        s[0] += a * x[j];
        as[0] += fabs(a);
        am = fmax(am, fabs(a));
      }
    }
  }
  double st = (s[0] + s[1]) + (s[2] + s[3]), ast = (as[0] + as[1]) + (as[2] + as[3]);
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) {
    st += __shfl_xor(st, off, 16);
    ast += __shfl_xor(ast, off, 16);
    am = fmax(am, __shfl_xor(am, off, 16));
  }
  // the five maxima over the workgroup's 16 rows (wave, then LDS), then one
  // atomic each per workgroup: one per row on the same five words
  // serialised in L2 and held the pass at 3.9 TB/s
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (i < n) {
    const double ri = b[i] - st;
    if (r && sub == 0) r[i] = ri;
    v[0] = am;
    v[1] = fabs(b[i]);
    v[2] = fabs(x[i]);
    v[3] = fabs(ri);
    v[4] = ast;
  }
  __shared__ double wmax[4][5];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    v[q] = wave_max_nonneg(v[q]);
    if (lane == 0) wmax[wave][q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    const int q = threadIdx.x;
    atomic_max_nonneg(&out[q], fmax(fmax(wmax[0][q], wmax[1][q]), fmax(wmax[2][q], wmax[3][q])));
  }
}

int residual(int n, const double *A, long lda, const double *x, const double *b, double *r,
             double *out, hipStream_t s) {
  hipLaunchKernelGGL(residual_init_kernel, dim3(1), dim3(64), 0, s, out);
  CLU_LAUNCHED("residual_init_kernel");
  if (n <= 0) return CLU_OK;
  hipLaunchKernelGGL(residual_kernel, dim3((unsigned)((n + 15) / 16)), dim3(256), 0, s, n, A,
                     lda, x, b, r, out);
  CLU_LAUNCHED("residual_kernel");
  return CLU_OK;
}

// out[0] = max_ij |a_ij|, out[1] = max_j sum_i |a_ij| (the infinity norm
// of A^T): one thread per column, rows in order (coalesced across lanes).
// 128 columns per workgroup as in the transposed product: two columns per
// lane, the 8 waves over every 8th row, four rows in flight per lane, the
// column sums added over the waves in a fixed order
__global__ __launch_bounds__(512) void abs_colsum_kernel(int n, const double *__restrict__ A,
                                                         long lda, double *__restrict__ out) {
  __shared__ double part[8][2][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c0 = blockIdx.x * 128 + 2 * lane;
  const bool vec = ((lda | n) & 1) == 0 && ((unsigned long)A & 15) == 0;
  double sx[4] = {0.0, 0.0, 0.0, 0.0}, sy[4] = {0.0, 0.0, 0.0, 0.0}, cm = 0.0;
  if (c0 < n) {
    if (vec) {
      int k = wave;
      for (; k + 24 < n; k += 32) {
        double2 a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
          a[u] = *reinterpret_cast<const double2 *>(A + (long)(k + 8 * u) * lda + c0);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          sx[u] += fabs(a[u].x);
          sy[u] += fabs(a[u].y);
          cm = fmax(cm, fmax(fabs(a[u].x), fabs(a[u].y)));
        }
      }
      for (; k < n; k += 8) {
        const double2 a0 = *reinterpret_cast<const double2 *>(A + (long)k * lda + c0);
        sx[0] += fabs(a0.x);
        sy[0] += fabs(a0.y);
        cm = fmax(cm, fmax(fabs(a0.x), fabs(a0.y)));
      }
    } else {
      for (int k = wave; k < n; k += 8) {
        const double a0 = fabs(A[(long)k * lda + c0]);
        sx[0] += a0;
        cm = fmax(cm, a0);
        if (c0 + 1 < n) {
          const double a1 = fabs(A[(long)k * lda + c0 + 1]);
          sy[0] += a1;
          cm = fmax(cm, a1);
        }
      }
    }
  }
  part[wave][0][lane] = (sx[0] + sx[1]) + (sx[2] + sx[3]);
  part[wave][1][lane] = (sy[0] + sy[1]) + (sy[2] + sy[3]);
  cm = wave_max(cm);
  if (lane == 0) atomic_max_nonneg(&out[0], cm);
  __syncthreads();
  if (wave < 2) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < 8; ++w) t += part[w][wave][lane];
    t = wave_max(c0 + wave < n ? t : 0.0);
    if (lane == 0) atomic_max_nonneg(&out[1], t);
  }
}

int abs_colsum(int n, const double *A, long lda, double *out, hipStream_t s) {
  hipLaunchKernelGGL(residual_init_kernel, dim3(1), dim3(64), 0, s, out);
  CLU_LAUNCHED("residual_init_kernel");
  if (n <= 0) return CLU_OK;
  hipLaunchKernelGGL(abs_colsum_kernel, dim3((unsigned)((n + 127) / 128)), dim3(512), 0, s, n, A,
                     lda, out);
  CLU_LAUNCHED("abs_colsum_kernel");
  return CLU_OK;
}

template <typename TD, typename TS>
__global__ void convert_kernel(long m, long n, const TS *__restrict__ src, long lds,
                               TD *__restrict__ dst, long ldd) {
  const long total = m * n;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const long i = e / n, j = e % n;
    dst[i * ldd + j] = (TD)src[i * lds + j];
  }
}

// The same for rows of 16-byte-aligned groups of four: workgroup b converts
// up to 4096 elements of one row (one division per workgroup, not per
// element), four per thread per pass, the four passes' loads issued first.
template <typename TD, typename TS>
__global__ __launch_bounds__(256) void convert_rows_kernel(long n, long chunks,
                                                           const TS *__restrict__ src, long lds,
                                                           TD *__restrict__ dst, long ldd) {
  typedef TS sv __attribute__((ext_vector_type(4)));
  typedef TD dv __attribute__((ext_vector_type(4)));
  const long b = blockIdx.x, i = b / chunks, c0 = (b - i * chunks) * 4096;
  const TS *sr = src + i * lds;
  TD *dr = dst + i * ldd;
  sv v[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const long j = c0 + (long)(u * 256 + threadIdx.x) * 4;
    if (j < n) v[u] = *reinterpret_cast<const sv *>(sr + j);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const long j = c0 + (long)(u * 256 + threadIdx.x) * 4;
    if (j < n) {
      dv o;
      o.x = (TD)v[u].x;
      o.y = (TD)v[u].y;
      o.z = (TD)v[u].z;
      o.w = (TD)v[u].w;
      *reinterpret_cast<dv *>(dr + j) = o;
    }
  }
}

template <typename TD, typename TS>
int convert(long m, long n, const TS *src, long lds, TD *dst, long ldd, hipStream_t s) {
  if (m <= 0 || n <= 0) return CLU_OK;
  if (n % 4 == 0 && lds % 4 == 0 && ldd % 4 == 0 && (uintptr_t)src % (4 * sizeof(TS)) == 0 &&
      (uintptr_t)dst % (4 * sizeof(TD)) == 0) {
    const long chunks = (n + 4095) / 4096;
    if (m * chunks < (1L << 31)) {
      hipLaunchKernelGGL((convert_rows_kernel<TD, TS>), dim3((unsigned)(m * chunks)), dim3(256), 0,
                         s, n, chunks, src, lds, dst, ldd);
      CLU_LAUNCHED("convert_rows_kernel");
      return CLU_OK;
    }
  }
  long blocks = (m * n + 255) / 256;
  if (blocks > 65536L * 16) blocks = 65536L * 16;
  hipLaunchKernelGGL((convert_kernel<TD, TS>), dim3((unsigned)blocks), dim3(256), 0, s, m, n,
                     src, lds, dst, ldd);
  CLU_LAUNCHED("convert_kernel");
  return CLU_OK;
}
template int convert<float, double>(long, long, const double *, long, float *, long, hipStream_t);
template int convert<double, float>(long, long, const float *, long, double *, long, hipStream_t);
template int convert<double, double>(long, long, const double *, long, double *, long, hipStream_t);
template int convert<float, float>(long, long, const float *, long, float *, long, hipStream_t);

__global__ void fill_int_kernel(int *p, int v, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Check of a panel eliminated without interchanges: *flag = 1 unless every
// multiplier (the strictly lower part of the nb x nb diagonal block, in the
// matrix, and the M x nb block L21 below it) is below thr in magnitude and
// every diagonal entry of U11 above tiny; NaN fails.  Workgroup i takes row
// i of the diagonal block, then row i - nb of L21.
__global__ __launch_bounds__(256) void spec_check_kernel(int M, int nb,
                                                         const float *__restrict__ L21, long ld21,
                                                         const float *__restrict__ A11, long lda,
                                                         float thr, float tiny,
                                                         int *__restrict__ flag) {
  const long i = blockIdx.x;
  bool bad = false;
  if (i < nb) {
    const float *r = A11 + i * lda;
    for (int j = threadIdx.x; j < i; j += 256) bad = bad || !(fabsf(r[j]) < thr);
    if (threadIdx.x == 0) bad = bad || !(fabsf(r[i]) > tiny);
  } else if (i - nb < M) {
    const float *r = L21 + (i - nb) * ld21;
    for (int j = threadIdx.x; j < nb; j += 256) bad = bad || !(fabsf(r[j]) < thr);
  }
  if (bad) *flag = 1;
}

int spec_check(int M, int nb, const float *L21, long ld21, const float *A11, long lda,
               float thr, float tiny, int *flag, hipStream_t s) {
  int rc = fill_int(flag, 0, 1, s);
  if (rc) return rc;
  if (nb <= 0) return CLU_OK;
  hipLaunchKernelGGL(spec_check_kernel, dim3((unsigned)(nb + (M > 0 ? M : 0))), dim3(256), 0, s,
                     M, nb, L21, ld21, A11, lda, thr, tiny, flag);
  CLU_LAUNCHED("spec_check_kernel");
  return CLU_OK;
}

int fill_int(int *p, int v, long n, hipStream_t s) {
  if (n <= 0) return CLU_OK;
  hipLaunchKernelGGL(fill_int_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, v, n);
  CLU_LAUNCHED("fill_int_kernel");
  return CLU_OK;
}

// Co-residency probe of a stream's CU mask: one workgroup per CU (96 KB of
// LDS each), every workgroup stamps the constant 100 MHz clock, sleeps a
// fixed time, stamps again and writes {XCC id, hardware id, start, end}.
// Workgroups that were not resident together start after another one ended.
// Every workgroup ends after its fixed sleep: nothing waits on another.
__global__ __launch_bounds__(256) void residency_probe_kernel(unsigned long long *out, int sleeps) {
  __shared__ int hold[96 * 1024 / 4];
  // every thread touches its part of the 96 KB (the array must stay
  // allocated: it is what holds one workgroup per CU)
  hold[threadIdx.x * 96] = (int)threadIdx.x + sleeps;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int pick = (int)((unsigned)sleeps % 256u);
    const int zero = hold[pick * 96] - pick - sleeps;  // 0
    const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    for (int i = 0; i < sleeps; ++i) __builtin_amdgcn_s_sleep(127);
    const unsigned long long t1 = __builtin_amdgcn_s_memrealtime();
    const unsigned hw = __builtin_amdgcn_s_getreg((31 << 11) | 4);    // HW_REG_HW_ID
    const unsigned xcc = __builtin_amdgcn_s_getreg((31 << 11) | 20);  // HW_REG_XCC_ID
    out[4 * blockIdx.x + 0] = xcc + (unsigned)zero;
    out[4 * blockIdx.x + 1] = hw;
    out[4 * blockIdx.x + 2] = t0;
    out[4 * blockIdx.x + 3] = t1;
  }
}

int residency_probe(int groups, int sleeps, unsigned long long *out, hipStream_t s) {
  if (groups <= 0) return CLU_OK;
  hipLaunchKernelGGL(residency_probe_kernel, dim3(groups), dim3(256), 0, s, out, sleeps);
  CLU_LAUNCHED("residency_probe_kernel");
  return CLU_OK;
}

__global__ void finish_info_kernel(int *info, const unsigned *abortw) {
  if (threadIdx.x == 0 && abortw[0] != 0u) info[0] = -1;
}

int finish_info(int *info, const unsigned *abortw, hipStream_t s) {
  hipLaunchKernelGGL(finish_info_kernel, dim3(1), dim3(64), 0, s, info, abortw);
  CLU_LAUNCHED("finish_info_kernel");
  return CLU_OK;
}

}  // namespace clu
