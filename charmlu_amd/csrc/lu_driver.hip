// lu_driver.hip -- host side of libcharmlu_amd.so: the C ABI of
// include/charmlu_amd.h, per-device workspaces, and the step loop of the
// blocked LU with a one-step look-ahead on two HIP streams.
//
// The reference runs one SDAG control script per b x b block (lu.ci:31-370)
// and a per-PE BlockScheduler that decides when ready trailing updates run
// (scheduler.C:235-484).  On one GPU the whole matrix is one device-resident
// row-major array and a step is a handful of launches:
//
//   panel stream:  panel(t+1) .................. [waits for the look-ahead
//                                                  block column of step t]
//   main stream:   swaps(t) -> U row(t) -> update(block col t+1) -> update(rest, part 1)
//   full stream:                          [after panel(t+1)] update(rest, part 2)
//
// so the latency-bound panel of step t+1 (the critical path, prio_notes:1-27
// in the reference) overlaps the bulk of step t's trailing update, which is
// what the reference's priorities (manager.h:54-59) aim at.  The panel and
// main streams have disjoint CU masks; the full stream has every CU and
// takes the part of the FP64 update left when the panel is done.  The reference's
// write-after-read hazard -- an L block may not be row-permuted by later
// pivots before every trailing update of its step has read it
// (lu.ci:240-242) -- is kept by stream order: the left-side interchanges of
// step t+1 run on the main stream after step t's update.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/charmlu_amd.h"
#include "common.h"
#include "kernels.h"

namespace clu {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
  set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  return CLU_ERR_HIP;
}

// ---------------------------------------------------------------------------
// per-device workspace
// ---------------------------------------------------------------------------
// Workgroups of one multi-workgroup panel strip, at most (the records of
// one exchange), and panel CUs, at most.
constexpr int MAX_GROUPS = 64;
constexpr int MAX_PANEL_CUS = 64;
constexpr int RUN = 256;         // pivots per move list

// Workgroups of one panel strip that are certain to be resident at once on
// the panel stream's CU mask (the top panel_cus of ncu CUs).  The strip
// kernels that exchange records hold one workgroup per CU (their register
// use allows one wave per SIMD, or two of a 512-thread workgroup).  CU mask
// bit c lies on XCD c % X (X = ncu / 32 XCDs of 32 CUs) and there on shader
// engine (c / X) % 4.  The hardware deals a launch's workgroups round-robin
// to the XCDs (starting where the previous launch left off) and, inside an
// XCD, round-robin to the shader engines that hold masked CUs, whatever
// their free CUs: so every XCD must take ceil(G / X) workgroups with each of
// its engines taking its share.  Measured with a bounded-sleep probe (one
// workgroup per CU) on MI355X: 32 of 32 CUs (one per engine) hold 32
// workgroups and not 36; 48 CUs (engines with 1 or 2) hold 32, not 36;
// 24 CUs hold 24, not 28; 16 hold 16, not 24.  With no CU masks
// (panel_cus = 0) the strip may run on any CU: 32.
static int resident_groups(int panel_cus, int ncu) {
  if (panel_cus <= 0) return 32;
  const int X = ncu >= 64 ? ncu / 32 : 1;
  int cap = 1 << 30;
  for (int x = 0; x < X; ++x) {
    int per_se[4] = {0, 0, 0, 0};
    for (int c = ncu - panel_cus; c < ncu; ++c)
      if (c % X == x) ++per_se[(c / X) % 4];
    int engines = 0, least = 1 << 30;
    for (int e = 0; e < 4; ++e)
      if (per_se[e] > 0) {
        ++engines;
        least = per_se[e] < least ? per_se[e] : least;
      }
    const int cap_x = engines > 0 ? engines * least : 0;
    cap = cap_x < cap ? cap_x : cap;
  }
  cap *= X;
  return cap < MAX_GROUPS ? cap : MAX_GROUPS;
}

struct PanelHistory {
  int n = 0, nb = 0, kind = -1;
  std::vector<hipEvent_t> ev;  // 2 per step
  std::vector<float> ms;       // per step, < 0 when unknown
};

struct Workspace {
  int device = -1;
  hipStream_t s_panel = nullptr;  // panels: the reserved CUs only
  hipStream_t s_main = nullptr;   // everything else: the other CUs
  hipStream_t s_full = nullptr;   // the end of each trailing update: every CU
  hipStream_t s_pgemm = nullptr;  // the large GEMMs inside a panel: every CU
  hipEvent_t ev_pg_in = nullptr, ev_pg_out = nullptr;
  int panel_cus = 0, ncu = 0;
  hipEvent_t ev_panel = nullptr, ev_la = nullptr, ev_start = nullptr, ev_end = nullptr;
  hipEvent_t ev_urow = nullptr;  // the step's interchanges and U row done (update stream)
  hipEvent_t ev_ula = nullptr;   // the same for the look-ahead block column only
  hipEvent_t ev_full = nullptr;
  // Durations of the last factorization's panels (begin/end events on the
  // panel stream, per step), read back at the next call of the same kind
  // with the same n and block size to size the part of each trailing update
  // that runs beside the panel: one history for FP64, one for the mixed
  // form.
  PanelHistory hist64, histmx;
  // panel exchange slots (sized for doubles, reused for floats)
  void *slots = nullptr;
  unsigned *abortw = nullptr;  // a strip's exchange gave up (cleared per factorization)
  // fence of the library's streams before an error returns
  hipEvent_t ev_fence[3] = {nullptr, nullptr, nullptr};
  // move lists: in-panel (one run) and per-step (nruns runs, two parities)
  int2 *pmoves = nullptr;
  int *pcount = nullptr;
  int2 *smoves = nullptr;
  int *scount = nullptr;
  int srun_cap = 0;  // runs per parity
  // host-pointer shim staging
  void *stage = nullptr;
  size_t stage_bytes = 0;
  // U-row solve by inverse: inv(L11) (nb x nb) and the U row block (nb x n)
  double *winv = nullptr;
  double *ubuf = nullptr;
  size_t winv_elems = 0, ubuf_elems = 0;
  // mixed precision: inv(L11) and the U row block in FP32, the 16-bit packed
  // operands of the trailing update, and the refinement's FP64 vectors
  void *winv32 = nullptr, *ubuf32 = nullptr, *pa16 = nullptr, *pu16 = nullptr;
  size_t winv32_b = 0, ubuf32_b = 0, pa16_b = 0, pu16_b = 0;
  // unpivoted mixed form: inv(U11) and L21 = A21 inv(U11) before it is copied back
  void *uinv32 = nullptr, *l21buf = nullptr;
  size_t uinv32_b = 0, l21buf_b = 0;
  // speculative identity pivots: the diagonal block before its unpivoted
  // elimination, the check's flag on the device and in pinned host memory
  void *specbak = nullptr;
  size_t specbak_b = 0;
  int *spec_flag = nullptr, *spec_hflag = nullptr;
  hipEvent_t ev_chk = nullptr;
  // FP8 update: power-of-two exponents per 128-row tile of L21 and of U12^T
  void *esc = nullptr;
  size_t esc_b = 0;
  void *irbuf = nullptr;
  size_t irbuf_b = 0;
  // solve: inverses of the diagonal blocks of L and U, the intermediate y
  void *sinv = nullptr, *sy = nullptr, *smv = nullptr;  // solve: inverses, Y, move lists
  size_t sinv_b = 0, sy_b = 0, smv_b = 0;
  // one-launch vector solves (trsv_tiles): published values with their
  // sequence numbers (zeroed when allocated), the last number used
  void *stag = nullptr;
  size_t stag_b = 0;
  unsigned trsv_seq = 0;
  // scratch of the triangular inverses (products of the doubling steps)
  void *tinv = nullptr;
  size_t tinv_b = 0;
  // recursive panel: inverse of the left half's L11 and the product
  // inv(L11) A12 before it replaces A12 (panel stream only)
  void *pinv = nullptr, *ptmp = nullptr;
  size_t pinv_b = 0, ptmp_b = 0;
  PanelWork pw{};
};

static std::mutex g_mu;
static Workspace g_ws[64];

// ---------------------------------------------------------------------------
// optional per-launch timing of the factorization's kernel classes
// ---------------------------------------------------------------------------
enum ProfKind { PK_GEMM_BIG = 0, PK_GEMM_OTHER = 1, PK_PANEL = 2, PK_SWAP = 3, PK_TRSM = 4,
                PK_NKINDS = 5 };
struct Prof {
  bool on = false;
  struct Rec { hipEvent_t a, b; int kind; double work; };
  std::vector<Rec> recs;
  size_t used = 0;
};
static Prof g_prof;

static int prof_begin(int kind, double work, hipStream_t s) {
  if (!g_prof.on) return -1;
  if (g_prof.used == g_prof.recs.size()) {
    Prof::Rec r{};
    if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return -1;
    g_prof.recs.push_back(r);
  }
  Prof::Rec &r = g_prof.recs[g_prof.used];
  r.kind = kind;
  r.work = work;
  if (hipEventRecord(r.a, s) != hipSuccess) return -1;
  return (int)g_prof.used++;
}
static void prof_end(int id, hipStream_t s) {
  if (id >= 0) (void)hipEventRecord(g_prof.recs[id].b, s);
}

static int ensure_runs(Workspace &w, int runs) {
  if (runs <= w.srun_cap) return CLU_OK;
  if (w.smoves) { CLU_HIP(hipFree(w.smoves)); w.smoves = nullptr; }
  if (w.scount) { CLU_HIP(hipFree(w.scount)); w.scount = nullptr; }
  CLU_HIP(hipMalloc(&w.smoves, sizeof(int2) * 2 * RUN * 2 * (size_t)runs));
  CLU_HIP(hipMalloc(&w.scount, sizeof(int) * 2 * (size_t)runs));
  w.srun_cap = runs;
  return CLU_OK;
}

static int get_ws(Workspace **out) {
  int dev = 0;
  CLU_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) { set_error("device %d out of range", dev); return CLU_ERR_HIP; }
  std::lock_guard<std::mutex> lk(g_mu);
  Workspace &w = g_ws[dev];
  if (w.device < 0) {
    // The update stream and the panel stream get disjoint CU masks, so the
    // latency-bound panel of the next step runs on compute units of its own
    // instead of queueing behind (and sharing SIMDs with) tens of thousands
    // of GEMM workgroups.
    hipDeviceProp_t prop;
    CLU_HIP(hipGetDeviceProperties(&prop, dev));
    const int ncu = prop.multiProcessorCount;
    int reserve = 32;
    if (const char *e = getenv("CHARMLU_PANEL_CUS")) {
      // 0 (no masks) or a multiple of the XCD count: the launches of a
      // masked stream are dealt to every XCD, so each XCD needs the same
      // number of its CUs (a mask without CUs on some XCD never runs the
      // workgroups dealt there)
      char *end = nullptr;
      const long v = strtol(e, &end, 10);
      const int X = ncu >= 64 ? ncu / 32 : 1;
      if (end == e || *end != '\0' || v < 0 || v > MAX_PANEL_CUS || v % X != 0 ||
          (v > 0 && v >= ncu)) {
        set_error("CHARMLU_PANEL_CUS=%s: expected 0 or a multiple of %d from %d to %d", e, X, X,
                  MAX_PANEL_CUS);
        return CLU_ERR_CONFIG;
      }
      reserve = (int)v;
    }
    if (reserve >= ncu) reserve = 0;
    w.s_main = w.s_panel = nullptr;
    if (reserve > 0) {
      std::vector<uint32_t> mm((ncu + 31) / 32, 0u), pm((ncu + 31) / 32, 0u);
      for (int c = 0; c < ncu; ++c) {
        if (c < ncu - reserve) mm[c / 32] |= 1u << (c % 32);
        else pm[c / 32] |= 1u << (c % 32);
      }
      if (hipExtStreamCreateWithCUMask(&w.s_main, (uint32_t)mm.size(), mm.data()) != hipSuccess ||
          hipExtStreamCreateWithCUMask(&w.s_panel, (uint32_t)pm.size(), pm.data()) != hipSuccess) {
        (void)hipGetLastError();
        if (w.s_main) (void)hipStreamDestroy(w.s_main);
        if (w.s_panel) (void)hipStreamDestroy(w.s_panel);
        w.s_main = w.s_panel = nullptr;
        reserve = 0;
      }
    }
    if (!w.s_main) CLU_HIP(hipStreamCreateWithFlags(&w.s_main, hipStreamNonBlocking));
    if (!w.s_panel) CLU_HIP(hipStreamCreateWithFlags(&w.s_panel, hipStreamNonBlocking));
    CLU_HIP(hipStreamCreateWithFlags(&w.s_full, hipStreamNonBlocking));
    CLU_HIP(hipStreamCreateWithFlags(&w.s_pgemm, hipStreamNonBlocking));
    CLU_HIP(hipEventCreateWithFlags(&w.ev_pg_in, hipEventDisableTiming));
    CLU_HIP(hipEventCreateWithFlags(&w.ev_pg_out, hipEventDisableTiming));
    w.panel_cus = reserve;
    w.ncu = ncu;
    CLU_HIP(hipEventCreateWithFlags(&w.ev_full, hipEventDisableTiming));
    CLU_HIP(hipEventCreateWithFlags(&w.ev_end, hipEventDisableTiming));
    CLU_HIP(hipEventCreateWithFlags(&w.ev_panel, hipEventDisableTiming));
    CLU_HIP(hipEventCreateWithFlags(&w.ev_la, hipEventDisableTiming));
    CLU_HIP(hipEventCreateWithFlags(&w.ev_urow, hipEventDisableTiming));
    CLU_HIP(hipEventCreateWithFlags(&w.ev_ula, hipEventDisableTiming));
    CLU_HIP(hipEventCreateWithFlags(&w.ev_start, hipEventDisableTiming));
    CLU_HIP(hipEventCreateWithFlags(&w.ev_chk, hipEventDisableTiming));
    CLU_HIP(hipMalloc(&w.spec_flag, 64));
    CLU_HIP(hipHostMalloc((void **)&w.spec_hflag, 64, hipHostMallocDefault));
    // panel candidate records: 2 parities x 64 workgroups x 136 words of 8
    // bytes, zeroed once; their sequence numbers only ever grow (pw.epoch),
    // so a record left by an earlier call can never look current
    const size_t slot_bytes = (size_t)PANEL_SLOT_WORDS * 8;
    CLU_HIP(hipMalloc(&w.slots, slot_bytes));
    CLU_HIP(hipMemset(w.slots, 0, slot_bytes));
    CLU_HIP(hipMalloc(&w.abortw, 64));
    CLU_HIP(hipMemset(w.abortw, 0, 64));
    for (int i = 0; i < 3; ++i)
      CLU_HIP(hipEventCreateWithFlags(&w.ev_fence[i], hipEventDisableTiming));
    CLU_HIP(hipMalloc(&w.pmoves, sizeof(int2) * 2 * RUN));
    CLU_HIP(hipMalloc(&w.pcount, 64));
    int rc = ensure_runs(w, 1);
    if (rc) return rc;
    w.pw.slots = w.slots;
    w.pw.abortw = w.abortw;
    w.pw.epoch = 1;
    w.pw.max_groups = resident_groups(reserve, ncu);
    w.device = dev;
  }
  *out = &w;
  return CLU_OK;
}

static int ensure_stage(Workspace &w, size_t bytes) {
  if (bytes <= w.stage_bytes) return CLU_OK;
  if (w.stage) CLU_HIP(hipFree(w.stage));
  w.stage = nullptr;
  w.stage_bytes = 0;
  size_t b = bytes < (1u << 20) ? (1u << 20) : bytes;
  if (hipMalloc(&w.stage, b) != hipSuccess) {
    set_error("cannot allocate %zu bytes of staging memory", b);
    return CLU_ERR_NOMEM;
  }
  w.stage_bytes = b;
  return CLU_OK;
}

static int ensure_dbuf(double **p, size_t *have, size_t need) {
  if (need <= *have) return CLU_OK;
  if (*p) CLU_HIP(hipFree(*p));
  *p = nullptr;
  *have = 0;
  if (hipMalloc(p, need * sizeof(double)) != hipSuccess) {
    set_error("cannot allocate %zu bytes of workspace", need * sizeof(double));
    return CLU_ERR_NOMEM;
  }
  *have = need;
  return CLU_OK;
}

static int ensure_bytes(void **p, size_t *have, size_t need) {
  if (need <= *have) return CLU_OK;
  if (*p) CLU_HIP(hipFree(*p));
  *p = nullptr;
  *have = 0;
  if (hipMalloc(p, need) != hipSuccess) {
    set_error("cannot allocate %zu bytes of workspace", need);
    return CLU_ERR_NOMEM;
  }
  *have = need;
  return CLU_OK;
}

// ---------------------------------------------------------------------------
// building blocks shared by the factorizations and solves
// ---------------------------------------------------------------------------

template <typename T>
static int gemm_t(bool ta, bool tb, int M, int N, int K, T alpha, const T *A, long lda,
                  const T *B, long ldb, T beta, T *C, long ldc, hipStream_t s);
template <>
int gemm_t<double>(bool ta, bool tb, int M, int N, int K, double alpha, const double *A,
                   long lda, const double *B, long ldb, double beta, double *C, long ldc,
                   hipStream_t s) {
  return dgemm(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}
template <>
int gemm_t<float>(bool ta, bool tb, int M, int N, int K, float alpha, const float *A, long lda,
                  const float *B, long ldb, float beta, float *C, long ldc, hipStream_t s) {
  return sgemm(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

// Products with the explicit inverse of a triangular factor, whose entries
// on the other side of the diagonal are zeros: in q blocks of the
// triangle's order (4 from order 1024, 2 from 512), each block's K ending
// at the diagonal, (q + 1) / (2 q) of the full product's flops.  Only zero
// terms at the end of each sum are dropped.
static int tri_blocks(int nb) { return nb >= 1024 && nb % 4 == 0 ? 4 : nb >= 512 && nb % 2 == 0 ? 2 : 1; }

// C (nb x N) = X B for X (nb x nb) lower triangular
template <typename T>
static int gemm_lower_left(int nb, int N, const T *X, long ldx, const T *B, long ldb, T *C,
                           long ldc, hipStream_t s) {
  const int q = tri_blocks(nb), bs = nb / q;
  for (int i = 0; i < q; ++i) {
    int rc = gemm_t<T>(false, false, bs, N, (i + 1) * bs, (T)1, X + (long)i * bs * ldx, ldx, B,
                       ldb, (T)0, C + (long)i * bs * ldc, ldc, s);
    if (rc) return rc;
  }
  return CLU_OK;
}

// C (M x nb) = A Y for Y (nb x nb) upper triangular
template <typename T>
static int gemm_upper_right(int M, int nb, const T *A, long lda, const T *Y, long ldy, T *C,
                            long ldc, hipStream_t s) {
  const int q = tri_blocks(nb), bs = nb / q;
  for (int j = 0; j < q; ++j) {
    int rc = gemm_t<T>(false, false, M, bs, (j + 1) * bs, (T)1, A, lda, Y + (long)j * bs, ldy,
                       (T)0, C + (long)j * bs, ldc, s);
    if (rc) return rc;
  }
  return CLU_OK;
}

// B = L^{-1} B, L unit lower M x M (any M): 256-row diagonal solves plus
// GEMM updates of the rows below.
template <typename T>
static int trsm_llnu_big(int M, int N, const T *L, long ldl, T *B, long ldb, hipStream_t s) {
  for (int i0 = 0; i0 < M; i0 += 256) {
    const int mi = M - i0 < 256 ? M - i0 : 256;
    int rc = trsm_llnu<T>(mi, N, L + (long)i0 * ldl + i0, ldl, B + (long)i0 * ldb, ldb, s);
    if (rc) return rc;
    if (i0 + mi < M) {
      rc = gemm_t<T>(false, false, M - i0 - mi, N, mi, (T)-1, L + (long)(i0 + mi) * ldl + i0, ldl,
                     B + (long)i0 * ldb, ldb, (T)1, B + (long)(i0 + mi) * ldb, ldb, s);
      if (rc) return rc;
    }
  }
  return CLU_OK;
}

// Factor rows [c0, n) x columns [c0, c0+w) of the panel [pc0, pc0+pw)
// (recursive halving down to strips of panel_strip_width columns -- 16, or
// 32 for FP32 -- factored by panel_base, which
// also applies each strip's interchanges to the rest of the panel, so both
// halves of every split are already permuted when they are used).
template <typename T>
static int factor_panel(T *A, long lda, int n, int c0, int w, int pc0, int pcw, int *ipiv,
                        Workspace &ws, hipStream_t s, bool pivot = true) {
  const int sw = panel_strip_width<T>(n - c0, ws.pw.max_groups, pivot);
  if (sw == 0) {
    set_error("a panel of %d rows needs more than %d resident workgroups", n - c0,
              ws.pw.max_groups);
    return CLU_ERR_UNSUPPORTED;
  }
  if (w <= sw) return panel_base<T>(A, lda, c0, n, w, pc0, pc0 + pcw, ipiv, ws.pw, s, pivot);
  int w1 = (w / 2 + 15) / 16 * 16;
  if (w1 >= w) w1 = w / 2;
  const int w2 = w - w1;
  int rc = factor_panel<T>(A, lda, n, c0, w1, pc0, pcw, ipiv, ws, s, pivot);
  if (rc) return rc;
  T *A11 = A + (long)c0 * lda + c0;
  // A12 <- inv(L11) A12: for w1 >= 128 through the inverse (about ten short
  // launches, most of them matrix-core GEMMs) rather than the 16-row blocked
  // substitution, whose serial steps cost more at that size
  if (w1 >= 128 && ws.pinv_b >= sizeof(T) * w1 * w1 && ws.ptmp_b >= sizeof(T) * w1 * w2 &&
      ws.tinv_b >= sizeof(T) * w1 * w1) {
    T *pinv = (T *)ws.pinv, *ptmp = (T *)ws.ptmp;
    if ((rc = tri_inv<T>(false, true, w1, 1, A11, lda, 0, pinv, w1, 0, (T *)ws.tinv, s)))
      return rc;
    if ((rc = gemm_t<T>(false, false, w1, w2, w1, (T)1, pinv, w1, A11 + w1, lda, (T)0, ptmp, w2,
                        s)))
      return rc;
    if ((rc = convert<T, T>(w1, w2, ptmp, w2, A11 + w1, lda, s))) return rc;
  } else if ((rc = trsm_llnu_big<T>(w1, w2, A11, lda, A11 + w1, lda, s))) {
    return rc;
  }
  // A22 -= A21 A12.  With the panel on CUs of its own, a GEMM of 1 GFLOP or
  // more runs on a stream of every CU instead (ordered by events): it still
  // has the panel's CUs, which are idle meanwhile, and takes update CUs as
  // they free up.  The strips stay on the panel stream.
  const int Mr = n - c0 - w1;
  if (Mr > 0) {
    const bool wide = ws.panel_cus > 0 && 2.0 * Mr * w2 * w1 >= 1e9;
    hipStream_t sg = wide ? ws.s_pgemm : s;
    if (wide) {
      CLU_HIP(hipEventRecord(ws.ev_pg_in, s));
      CLU_HIP(hipStreamWaitEvent(sg, ws.ev_pg_in, 0));
    }
    if ((rc = gemm_t<T>(false, false, Mr, w2, w1, (T)-1, A11 + (long)w1 * lda, lda, A11 + w1, lda,
                        (T)1, A11 + (long)w1 * lda + w1, lda, sg)))
      return rc;
    if (wide) {
      CLU_HIP(hipEventRecord(ws.ev_pg_out, sg));
      CLU_HIP(hipStreamWaitEvent(s, ws.ev_pg_out, 0));
    }
  }
  return factor_panel<T>(A, lda, n, c0 + w1, w2, pc0, pcw, ipiv, ws, s, pivot);
}

// CHARMLU_FORCE_SPLIT=1 (read at every factorization): every trailing
// update splits its rows in two halves (part 1 rounded up to 128 rows), so
// that small test problems run both launches and their events; results do
// not depend on the split.
static bool split_forced() {
  const char *e = getenv("CHARMLU_FORCE_SPLIT");
  return e && e[0] == '1';
}
// speculative identity pivots for the pivoted low-precision form
// (CHARMLU_SPECULATE=0: every panel exchanges candidate rows)
static bool speculate_enabled() {
  const char *e = getenv("CHARMLU_SPECULATE");
  return !(e && e[0] == '0');
}
// steps per group of the unpivoted low-precision form (CHARMLU_GROUP: 1 for
// one trailing update per step, as the other forms, up to 16; default
// group_default)
static int group_steps(int group_default) {
  const char *e = getenv("CHARMLU_GROUP");
  if (!e) return group_default;
  char *end = nullptr;
  const long g = strtol(e, &end, 10);
  if (end == e || *end != 0 || g < 1 || g > 16) return group_default;
  return (int)g;
}
// one-launch triangles for the solve of one right-hand side
// (CHARMLU_TRSV_TILES=0: a pair of launches per diagonal block instead)
static bool trsv_tiles_enabled() {
  const char *e = getenv("CHARMLU_TRSV_TILES");
  return !(e && e[0] == '0');
}
static int forced_split(int M) {
  const int Ma = (M / 2 + 127) / 128 * 128;
  return Ma > M ? M : Ma;
}

// Panel durations of the previous call of the same kind and shape (events
// that are not complete, or a different kind or shape, leave -1), then
// events for this call's NB steps.
static int history_prepare(PanelHistory &h, int n, int nb, int kind, int NB) {
  const bool same = h.n == n && h.nb == nb && h.kind == kind && (int)h.ev.size() >= 2 * NB;
  if (!same) h.ms.assign(NB, -1.0f);
  if (same) {
    for (int t = 0; t < NB; ++t) {
      float ms = -1.0f;
      if (hipEventQuery(h.ev[2 * t + 1]) == hipSuccess &&
          hipEventElapsedTime(&ms, h.ev[2 * t], h.ev[2 * t + 1]) == hipSuccess)
        h.ms[t] = ms;
      (void)hipGetLastError();
    }
  }
  while ((int)h.ev.size() < 2 * NB) {
    hipEvent_t e;
    CLU_HIP(hipEventCreate(&e));
    h.ev.push_back(e);
  }
  h.n = n;
  h.nb = nb;
  h.kind = kind;
  return CLU_OK;
}

// On an error in the middle of a factorization: order the caller's stream
// after everything already enqueued on the library's streams, so that no
// kernel the call launched still writes the caller's buffers once the
// error has been returned and the caller synchronizes its stream.
static void fence_streams(Workspace &ws, hipStream_t su) {
  hipStream_t ss[3] = {ws.s_panel, ws.s_full, ws.s_pgemm};
  for (int i = 0; i < 3; ++i)
    if (hipEventRecord(ws.ev_fence[i], ss[i]) == hipSuccess)
      (void)hipStreamWaitEvent(ws.s_main, ws.ev_fence[i], 0);
  if (hipEventRecord(ws.ev_end, ws.s_main) == hipSuccess)
    (void)hipStreamWaitEvent(su, ws.ev_end, 0);
  (void)hipGetLastError();
}

// The tallest strip of a factorization is its first (n rows): if a launch
// shape exists for it, every later strip has one too.
template <typename T>
static int check_panel_shape(const Workspace &ws, int n, bool pivot) {
  if (n <= 0 || panel_strip_width<T>(n, ws.pw.max_groups, pivot) > 0) return CLU_OK;
  set_error("n = %d: the first panel needs more than the %d workgroups that can be resident at "
            "once on %d panel CUs (largest FP64 n with 32: 131072, FP32: 262144)",
            n, ws.pw.max_groups, ws.panel_cus);
  return CLU_ERR_UNSUPPORTED;
}

// move lists of ipiv[k1, k2) in runs of RUN pivots: one launch of one
// workgroup per run (one run: the single-run kernel)
static int build_moves(const int *ipiv, int k1, int k2, int2 *moves, int *counts, hipStream_t s) {
  if (k2 - k1 <= RUN) return pivots_to_moves(ipiv, k1, k2, moves, counts, s);
  return pivots_to_moves_all(ipiv, k1, k2, moves, counts, s);
}

// the runs' move lists applied in one launch
template <typename T>
static int apply_runs(int nruns, const int2 *moves, const int *counts, T *A, long lda, int c0,
                      int c1, int skip0, int skip1, bool inverse, hipStream_t s) {
  return apply_moves<T>(moves, counts, 2 * RUN, A, lda, c0, c1, skip0, skip1, inverse, s, nruns,
                        2L * RUN);
}


// Solve op(A) X = B from P A = L U (forwardSolve / backwardSolve /
// offDiagSolve, lu.ci:387-458): the row interchanges, then block forward and
// backward substitution over S-row blocks.  The diagonal blocks of L and U
// are inverted up front, all in one batched launch each, so the sequential
// part of each block step is one small matrix-vector (or matrix) product
// with the block's inverse instead of an S-step substitution; the block's
// contribution is then subtracted from every row still to be solved (a GEMV
// for one right-hand side, a GEMM for several).  S = 256 (explicit inverses
// of 512-row triangles doubled the scaled residual of the FP64 solve at
// n = 32768).  The refinement's FP32 preconditioner solves of systems of
// 8192 and more pass S = 1024: a vector solve is two launches per block, at
// n = 65536 its 1024 launches of 256-row blocks spent a third of the solve
// between kernels, and the refinement corrects the extra rounding.
// The transposed solve keeps the inverses transposed (one in-place pass
// after inverting), so both directions apply them by rows.
// inv_ready: the inverses of the diagonal blocks are still in the workspace
// from the previous call with the same factors, S and trans (the
// refinement's repeated solves)
template <typename T>
static int getrs_impl(int trans, int n, int nrhs, const T *dLU, long lda, const int *dipiv, T *dB,
                      long ldb, hipStream_t s, bool inv_ready = false, int S = 256) {
  if (trans != CHARMLU_NO_TRANS && trans != CHARMLU_TRANS) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (!dLU && n > 0) return -4;
  if (lda < (n > 1 ? n : 1)) return -5;
  if (!dipiv && n > 0) return -6;
  if (!dB && n > 0 && nrhs > 0) return -7;
  if (ldb < (nrhs > 1 ? nrhs : 1)) return -8;
  if (n == 0 || nrhs == 0) return CLU_OK;
  Workspace *wsp;
  int rc = get_ws(&wsp);
  if (rc) return rc;
  Workspace &ws = *wsp;
  if (S != 256 && S != 1024) return -1;
  const int nblk = (n + S - 1) / S;
  const int nruns = (n + RUN - 1) / RUN;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if ((rc = ensure_bytes(&ws.sinv, &ws.sinv_b, sizeof(T) * 3 * (size_t)nblk * S * S))) return rc;
    if ((rc = ensure_bytes(&ws.sy, &ws.sy_b, sizeof(T) * (size_t)n * nrhs))) return rc;
    if ((rc = ensure_bytes(&ws.smv, &ws.smv_b,
                           (sizeof(int2) * 2 * RUN + sizeof(int)) * (size_t)nruns)))
      return rc;
  }
  T *Linv = (T *)ws.sinv, *Uinv = Linv + (size_t)nblk * S * S, *tbuf = Uinv + (size_t)nblk * S * S;
  T *Y = (T *)ws.sy;
  const long ldy = nrhs;
  int2 *vmoves = (int2 *)ws.smv;
  int *vcounts = (int *)(vmoves + (size_t)nruns * 2 * RUN);
  // inverses of the diagonal blocks, by recursive doubling, all blocks of a
  // triangle in one batch (the last, shorter block on its own)
  const int nfull = n / S, last = n - nfull * S;
  for (int up = 0; up < 2 && !inv_ready; ++up) {
    T *Inv = up ? Uinv : Linv;
    if (nfull > 0 && (rc = tri_inv<T>(up, !up, S, nfull, dLU, lda, (long)S * (lda + 1), Inv, S,
                                      (long)S * S, tbuf, s)))
      return rc;
    if (last > 0 && (rc = tri_inv<T>(up, !up, last, 1, dLU + (long)nfull * S * (lda + 1), lda, 0,
                                     Inv + (size_t)nfull * S * S, S, 0, tbuf, s)))
      return rc;
    // (a shorter last block's inverse fills the top left of its S x S slot,
    // which the transpose maps onto itself)
    if (trans == CHARMLU_TRANS && (rc = blk_transpose<T>(S, nblk, Inv, s))) return rc;
  }
  const bool vec = nrhs == 1 && ldb == 1;
  auto interchanges = [&](bool inverse) -> int {
    if (vec) return permute_vec<T>(dipiv, n, dB, inverse, vmoves, vcounts, s);
    // every run's move list in one launch, then all runs in order in another
    int e = pivots_to_moves_all(dipiv, 0, n, vmoves, vcounts, s);
    if (!e)
      e = apply_moves<T>(vmoves, vcounts, 2 * RUN, dB, ldb, 0, nrhs, nrhs, nrhs, inverse, s, nruns,
                         2L * RUN);
    return e;
  };
  // dst = op(G) src for the inverse G of one diagonal block (G^T stored for
  // the transposed solve: pass tr = false there)
  auto diag = [&](bool tr, int m, const T *G, const T *src, long lds, T *dst, long ldd) -> int {
    if (vec) return blk_gemv<T>(tr, m, G, S, src, dst, s);
    return gemm_t<T>(tr, false, m, nrhs, m, (T)1, G, S, src, lds, (T)0, dst, ldd, s);
  };
  // off-diagonal contribution of a solved block (offDiagSolve, lu.ci:439-456)
  auto update = [&](bool ta, int M, int K, const T *Ablk, const T *X, long ldx, T *D,
                    long ldd) -> int {
    if (M <= 0) return CLU_OK;
    if (vec) return gemv_sub<T>(ta, M, K, Ablk, lda, X, D, s);
    return gemm_t<T>(ta, false, M, nrhs, K, (T)-1, Ablk, lda, X, ldx, (T)1, D, ldd, s);
  };
  const T *F = dLU;
  constexpr int VW = sizeof(T) / 4;
  if (vec && n % S == 0 && lda % (16 / sizeof(T)) == 0 && ((uintptr_t)dLU & 15) == 0 &&
      trsv_tiles_enabled()) {
    // both triangles in one launch each (trsv_tiles): L y = P b, then U x = y;
    // transposed: U^T z = b, then L^T w = z, x = P^T w
    unsigned long long *tag;
    unsigned seq;
    {
      std::lock_guard<std::mutex> lk(g_mu);
      const size_t had = ws.stag_b, need = 2 * sizeof(unsigned long long) * VW * (size_t)n;
      if ((rc = ensure_bytes(&ws.stag, &ws.stag_b, need))) return rc;
      if (ws.stag_b != had || ws.trsv_seq >= 0xfffffff0u) {
        if ((rc = fill_int((int *)ws.stag, 0, (long)(ws.stag_b / sizeof(int)), s))) return rc;
        ws.trsv_seq = 0;
      }
      seq = ws.trsv_seq + 1;
      ws.trsv_seq += 2;
      tag = (unsigned long long *)ws.stag;
    }
    unsigned *abortw = (unsigned *)ws.abortw + 8;
    unsigned long long *tagx = tag, *tagt = tag + (size_t)VW * n;
    const bool tr = trans == CHARMLU_TRANS;
    if (!tr && (rc = interchanges(false))) return rc;
    if ((rc = trsv_tiles<T>(false, tr, n, S, F, lda, tr ? Uinv : Linv, dB, Y, tagx, tagt, seq,
                            abortw, s)))
      return rc;
    if ((rc = trsv_tiles<T>(true, tr, n, S, F, lda, tr ? Linv : Uinv, Y, dB, tagx, tagt, seq + 1,
                            abortw, s)))
      return rc;
    return tr ? interchanges(true) : CLU_OK;
  }
  if (trans == CHARMLU_NO_TRANS) {
    if ((rc = interchanges(false))) return rc;
    // L y = P b: y into Y, B is the running right-hand side
    for (int k = 0; k < nblk; ++k) {
      const int x0 = k * S, m = n - x0 < S ? n - x0 : S;
      if ((rc = diag(false, m, Linv + (size_t)k * S * S, dB + (long)x0 * ldb, ldb,
                     Y + (long)x0 * ldy, ldy)))
        return rc;
      if ((rc = update(false, n - x0 - m, m, F + (long)(x0 + m) * lda + x0, Y + (long)x0 * ldy,
                       ldy, dB + (long)(x0 + m) * ldb, ldb)))
        return rc;
    }
    // U x = y: x into B, Y is the running right-hand side
    for (int k = nblk - 1; k >= 0; --k) {
      const int x0 = k * S, m = n - x0 < S ? n - x0 : S;
      if ((rc = diag(false, m, Uinv + (size_t)k * S * S, Y + (long)x0 * ldy, ldy,
                     dB + (long)x0 * ldb, ldb)))
        return rc;
      if ((rc = update(false, x0, m, F + x0, dB + (long)x0 * ldb, ldb, Y, ldy))) return rc;
    }
    return CLU_OK;
  }
  // A^T X = B with P A = L U:  U^T Z = B (Z into Y), L^T W = Z (W into B),
  // X = P^T W
  for (int k = 0; k < nblk; ++k) {
    const int x0 = k * S, m = n - x0 < S ? n - x0 : S;
    if ((rc = diag(false, m, Uinv + (size_t)k * S * S, dB + (long)x0 * ldb, ldb,
                   Y + (long)x0 * ldy, ldy)))
      return rc;
    if ((rc = update(true, n - x0 - m, m, F + (long)x0 * lda + x0 + m, Y + (long)x0 * ldy, ldy,
                     dB + (long)(x0 + m) * ldb, ldb)))
      return rc;
  }
  for (int k = nblk - 1; k >= 0; --k) {
    const int x0 = k * S, m = n - x0 < S ? n - x0 : S;
    if ((rc = diag(false, m, Linv + (size_t)k * S * S, Y + (long)x0 * ldy, ldy,
                   dB + (long)x0 * ldb, ldb)))
      return rc;
    if ((rc = update(true, x0, m, F + (long)x0 * lda, dB + (long)x0 * ldb, ldb, Y, ldy)))
      return rc;
  }
  return interchanges(true);
}

}  // namespace clu

using namespace clu;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int charmlu_version(void) { return 10000; }

int charmlu_profile(int on) {
  g_prof.on = on != 0;
  g_prof.used = 0;
  return CLU_OK;
}

int charmlu_profile_read(double *out, int nout) {
  if (!out || nout < 3 * PK_NKINDS) return -1;
  const int nfill = nout >= 4 * PK_NKINDS ? 4 * PK_NKINDS : 3 * PK_NKINDS;
  for (int i = 0; i < nfill; ++i) out[i] = 0.0;
  // intervals of every record on one time axis (ms after the first start)
  std::vector<std::vector<std::pair<double, double>>> iv(PK_NKINDS);
  for (size_t i = 0; i < g_prof.used; ++i) {
    Prof::Rec &r = g_prof.recs[i];
    CLU_HIP(hipEventSynchronize(r.b));
    float ms = 0.f, t0 = 0.f, t1 = 0.f;
    CLU_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    CLU_HIP(hipEventElapsedTime(&t0, g_prof.recs[0].a, r.a));
    CLU_HIP(hipEventElapsedTime(&t1, g_prof.recs[0].a, r.b));
    out[3 * r.kind + 0] += 1.0;
    out[3 * r.kind + 1] += r.work;
    out[3 * r.kind + 2] += ms;
    iv[r.kind].push_back({(double)t0, (double)t1});
  }
  if (nfill == 4 * PK_NKINDS) {
    // time during which at least one record of the class was running
    for (int k = 0; k < PK_NKINDS; ++k) {
      auto &v = iv[k];
      std::sort(v.begin(), v.end());
      double busy = 0.0, cs = 0.0, ce = -1.0;
      for (auto &x : v) {
        if (x.first > ce) {
          if (ce > cs) busy += ce - cs;
          cs = x.first;
          ce = x.second;
        } else if (x.second > ce) {
          ce = x.second;
        }
      }
      if (ce > cs) busy += ce - cs;
      out[3 * PK_NKINDS + k] = busy;
    }
  }
  g_prof.used = 0;
  return CLU_OK;
}

const char *charmlu_last_error(void) { return g_err; }

int charmlu_init(int device) {
  CLU_HIP(hipSetDevice(device));
  Workspace *ws = nullptr;
  return get_ws(&ws);
}

// Everything a workspace holds, released on its device; the workspace is
// then as new (the next call on that device makes it again).
static int release_ws(Workspace &w) {
  if (w.device < 0) return CLU_OK;
  CLU_HIP(hipSetDevice(w.device));
  CLU_HIP(hipDeviceSynchronize());
  hipStream_t ss[4] = {w.s_panel, w.s_main, w.s_full, w.s_pgemm};
  for (hipStream_t st : ss)
    if (st) CLU_HIP(hipStreamDestroy(st));
  hipEvent_t ev[] = {w.ev_pg_in, w.ev_pg_out, w.ev_panel,    w.ev_la,       w.ev_start,
                     w.ev_end,   w.ev_urow,   w.ev_ula,      w.ev_full,     w.ev_fence[0],
                     w.ev_fence[1], w.ev_fence[2], w.ev_chk};
  for (hipEvent_t e : ev)
    if (e) CLU_HIP(hipEventDestroy(e));
  for (PanelHistory *h : {&w.hist64, &w.histmx})
    for (hipEvent_t e : h->ev) CLU_HIP(hipEventDestroy(e));
  void *bufs[] = {w.slots,  w.abortw, w.pmoves, w.pcount, w.smoves, w.scount, w.stage,
                  w.winv,   w.ubuf,   w.winv32, w.ubuf32, w.pa16,   w.pu16,   w.uinv32,
                  w.l21buf, w.esc,    w.irbuf,  w.sinv,   w.sy,     w.smv,    w.tinv,
                  w.pinv,   w.ptmp,   w.specbak, w.spec_flag, w.stag};
  for (void *b : bufs)
    if (b) CLU_HIP(hipFree(b));
  if (w.spec_hflag) CLU_HIP(hipHostFree(w.spec_hflag));
  w = Workspace();
  return CLU_OK;
}

int charmlu_finalize(void) {
  int prev = -1;
  (void)hipGetDevice(&prev);
  std::lock_guard<std::mutex> lk(g_mu);
  int rc = CLU_OK;
  for (Workspace &w : g_ws) {
    const int r = release_ws(w);
    if (r && !rc) rc = r;
  }
  for (Prof::Rec &r : g_prof.recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  g_prof.recs.clear();
  g_prof.used = 0;
  g_prof.on = false;
  if (prev >= 0) (void)hipSetDevice(prev);
  return rc;
}

int charmlu_panel_probe(int groups, int sleeps, unsigned long long *dout, void *stream) {
  if (groups < 0 || groups > 4096) return -1;
  if (sleeps < 0 || sleeps > 4096) return -2;
  if (!dout && groups > 0) return -3;
  if (groups == 0) return CLU_OK;
  Workspace *wsp;
  int rc;
  if ((rc = get_ws(&wsp))) return rc;
  Workspace &ws = *wsp;
  hipStream_t su = (hipStream_t)stream;
  // on the panel stream, ordered after the caller's stream and back
  CLU_HIP(hipEventRecord(ws.ev_start, su));
  CLU_HIP(hipStreamWaitEvent(ws.s_panel, ws.ev_start, 0));
  if ((rc = residency_probe(groups, sleeps, dout, ws.s_panel))) return rc;
  CLU_HIP(hipEventRecord(ws.ev_end, ws.s_panel));
  CLU_HIP(hipStreamWaitEvent(su, ws.ev_end, 0));
  return CLU_OK;
}

int charmlu_panel_plan(int prec, int m, int w, int panel_cus, int ncu, int pivot, int *out) {
  if (prec != CHARMLU_PREC_FP64 && prec != CHARMLU_PREC_FP32) return -1;
  if (m < 1) return -2;
  if (w < 1) return -3;
  if (panel_cus < 0 || panel_cus > MAX_PANEL_CUS) return -4;
  if (ncu < 1) return -5;
  if (!out) return -7;
  const int groups = resident_groups(panel_cus, ncu);
  PanelShape ps{0, 0, 0, 0};
  const int r = prec == CHARMLU_PREC_FP64 ? panel_plan<double>(m, w, groups, pivot != 0, &ps)
                                          : panel_plan<float>(m, w, groups, pivot != 0, &ps);
  out[0] = ps.sw;
  out[1] = ps.R;
  out[2] = ps.NT;
  out[3] = ps.G;
  out[4] = groups;
  out[5] = prec == CHARMLU_PREC_FP64 ? panel_strip_width<double>(m, groups, pivot != 0)
                                     : panel_strip_width<float>(m, groups, pivot != 0);
  if (r) {
    set_error("no launch shape for a %d x %d strip with %d resident workgroups", m, w, groups);
    return CLU_ERR_UNSUPPORTED;
  }
  return CLU_OK;
}

// ---- device-pointer block ops ----------------------------------------------
int charmlu_dgemm_dev(int transa, int transb, int M, int N, int K, double alpha, const double *dA,
                      long lda, const double *dB, long ldb, double beta, double *dC, long ldc,
                      void *stream) {
  if (transa != CHARMLU_NO_TRANS && transa != CHARMLU_TRANS) return -1;
  if (transb != CHARMLU_NO_TRANS && transb != CHARMLU_TRANS) return -2;
  if (M < 0) return -3;
  if (N < 0) return -4;
  if (K < 0) return -5;
  const bool ta = transa == CHARMLU_TRANS, tb = transb == CHARMLU_TRANS;
  if (lda < ((ta ? M : K) > 1 ? (ta ? M : K) : 1)) return -8;
  if (ldb < ((tb ? K : N) > 1 ? (tb ? K : N) : 1)) return -10;
  if (ldc < (N > 1 ? N : 1)) return -13;
  return dgemm(ta, tb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, (hipStream_t)stream);
}

int charmlu_sgemm_dev(int transa, int transb, int M, int N, int K, float alpha, const float *dA,
                      long lda, const float *dB, long ldb, float beta, float *dC, long ldc,
                      void *stream) {
  if (transa != CHARMLU_NO_TRANS && transa != CHARMLU_TRANS) return -1;
  if (transb != CHARMLU_NO_TRANS && transb != CHARMLU_TRANS) return -2;
  if (M < 0) return -3;
  if (N < 0) return -4;
  if (K < 0) return -5;
  const bool ta = transa == CHARMLU_TRANS, tb = transb == CHARMLU_TRANS;
  if (lda < ((ta ? M : K) > 1 ? (ta ? M : K) : 1)) return -8;
  if (ldb < ((tb ? K : N) > 1 ? (tb ? K : N) : 1)) return -10;
  if (ldc < (N > 1 ? N : 1)) return -13;
  return sgemm(ta, tb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, (hipStream_t)stream);
}

int charmlu_gemm_lp_dev(int prec, int M, int N, int K, float alpha, const void *dA, long lda,
                        const void *dB, long ldb, float beta, float *dC, long ldc,
                        void *stream) {
  if (prec != CHARMLU_PREC_BF16 && prec != CHARMLU_PREC_FP16) return -1;
  if (M < 0) return -2;
  if (N < 0) return -3;
  if (K < 0 || K % 32 != 0) return -4;
  if (lda < (K > 1 ? K : 1) || lda % 8 != 0) return -7;
  if (ldb < (K > 1 ? K : 1) || ldb % 8 != 0) return -9;
  if (ldc < (N > 1 ? N : 1)) return -12;
  if (M == 0 || N == 0 || K == 0) return CLU_OK;
  hipStream_t s = (hipStream_t)stream;
  if (prec == CHARMLU_PREC_BF16)
    return hgemm_nt<__bf16>(M, N, K, alpha, (const __bf16 *)dA, lda, (const __bf16 *)dB, ldb,
                            beta, dC, ldc, s);
  return hgemm_nt<_Float16>(M, N, K, alpha, (const _Float16 *)dA, lda, (const _Float16 *)dB, ldb,
                            beta, dC, ldc, s);
}

int charmlu_gemm_fp8_dev(int M, int N, int K, float alpha, const float *dA, long lda,
                         const float *dB, long ldb, float beta, float *dC, long ldc,
                         void *stream) {
  if (M < 0) return -1;
  if (N < 0) return -2;
  if (K < 0) return -3;
  if (lda < (K > 1 ? K : 1)) return -6;
  if (ldb < (N > 1 ? N : 1)) return -8;
  if (ldc < (N > 1 ? N : 1)) return -11;
  if (M == 0 || N == 0) return CLU_OK;
  hipStream_t s = (hipStream_t)stream;
  if (K == 0 || alpha == 0.f) {
    // C = beta C
    return sgemm(false, false, M, N, 0, 0.0f, dA, lda, dB, ldb, beta, dC, ldc, s);
  }
  Workspace *wsp;
  int rc;
  if ((rc = get_ws(&wsp))) return rc;
  Workspace &ws = *wsp;
  const int Kp = (K + 63) / 64 * 64;
  const int tm = (M + 127) / 128, tn = (N + 127) / 128;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if ((rc = ensure_bytes(&ws.pa16, &ws.pa16_b, (size_t)M * Kp))) return rc;
    if ((rc = ensure_bytes(&ws.pu16, &ws.pu16_b, (size_t)N * Kp))) return rc;
    if ((rc = ensure_bytes(&ws.esc, &ws.esc_b,
                           (sizeof(int) + sizeof(float) * FP8_PART_PER_TILE) * (size_t)(tm + tn))))
      return rc;
  }
  uint8_t *qa = (uint8_t *)ws.pa16, *qb = (uint8_t *)ws.pu16;
  int *ea = (int *)ws.esc, *eb = ea + tm;
  float *part = (float *)(eb + tn);
  if ((rc = quant_rows_fp8(M, K, Kp, 1.0f, dA, lda, qa, ea, part, s))) return rc;
  if ((rc = quant_cols_fp8(K, N, Kp, dB, ldb, qb, eb, part + (size_t)FP8_PART_PER_TILE * tm, s)))
    return rc;
  return fgemm_nt(M, N, Kp, alpha, qa, Kp, ea, qb, Kp, eb, beta, dC, ldc, s);
}

// op(T) is M x M; for M > 256 the solve proceeds in 256-row blocks with
// GEMM updates between them.
static int dtrsm_dev_any(bool upper, bool trans, bool unit, int M, int N, double alpha,
                         const double *T, long ldt, double *B, long ldb, hipStream_t s) {
  if (M <= 256) return trsm_small<double>(upper, trans, unit, M, N, alpha, T, ldt, B, ldb, s);
  const bool oplow = (!upper) != trans;
  const int nblk = (M + 255) / 256;
  for (int bi = 0; bi < nblk; ++bi) {
    const int blk = oplow ? bi : nblk - 1 - bi;
    const int i0 = blk * 256, mi = M - i0 < 256 ? M - i0 : 256;
    // alpha scales the first block here and every other row through the
    // beta of the first update, so later blocks solve with alpha = 1
    const double a = bi == 0 ? alpha : 1.0;
    int rc = trsm_small<double>(upper, trans, unit, mi, N, a, T + (long)i0 * ldt + i0, ldt,
                                B + (long)i0 * ldb, ldb, s);
    if (rc) return rc;
    // update the rows still to be solved: B_rest -= op(T)[rest, blk] * X_blk
    if (oplow && i0 + mi < M) {
      const int r0 = i0 + mi, mr = M - r0;
      // op(T)[r0.., i0..i0+mi): trans ? T[i0.., r0..]^T : T[r0.., i0..]
      rc = trans ? dgemm(true, false, mr, N, mi, -1.0, T + (long)i0 * ldt + r0, ldt,
                         B + (long)i0 * ldb, ldb, alpha, B + (long)r0 * ldb, ldb, s)
                 : dgemm(false, false, mr, N, mi, -1.0, T + (long)r0 * ldt + i0, ldt,
                         B + (long)i0 * ldb, ldb, alpha, B + (long)r0 * ldb, ldb, s);
      if (rc) return rc;
    } else if (!oplow && i0 > 0) {
      rc = trans ? dgemm(true, false, i0, N, mi, -1.0, T + (long)i0 * ldt, ldt,
                         B + (long)i0 * ldb, ldb, alpha, B, ldb, s)
                 : dgemm(false, false, i0, N, mi, -1.0, T + i0, ldt, B + (long)i0 * ldb, ldb,
                         alpha, B, ldb, s);
      if (rc) return rc;
    }
    alpha = 1.0;  // the remaining rows were scaled by the update above
  }
  return CLU_OK;
}

int charmlu_dtrsm_dev(int uplo, int transa, int diag, int M, int N, const double *dT, long ldt,
                      double *dB, long ldb, void *stream) {
  if (uplo != CHARMLU_UPPER && uplo != CHARMLU_LOWER) return -1;
  if (transa != CHARMLU_NO_TRANS && transa != CHARMLU_TRANS) return -2;
  if (diag != CHARMLU_UNIT && diag != CHARMLU_NON_UNIT) return -3;
  if (M < 0) return -4;
  if (N < 0) return -5;
  if (ldt < (M > 1 ? M : 1)) return -7;
  if (ldb < (N > 1 ? N : 1)) return -9;
  return dtrsm_dev_any(uplo == CHARMLU_UPPER, transa == CHARMLU_TRANS, diag == CHARMLU_UNIT, M,
                       N, 1.0, dT, ldt, dB, ldb, (hipStream_t)stream);
}

int charmlu_dger_dev(int M, int N, double alpha, const double *dx, long incx, const double *dy,
                     long incy, double *dA, long lda, void *stream) {
  if (M < 0) return -1;
  if (N < 0) return -2;
  if (incx == 0) return -5;
  if (incy == 0) return -7;
  if (lda < (N > 1 ? N : 1)) return -9;
  return dger(M, N, alpha, dx, incx, dy, incy, dA, lda, (hipStream_t)stream);
}

// ---- host-pointer CBLAS shims -----------------------------------------------
// Stage operands through device memory, run the device op on the null stream,
// copy the result back; synchronous like the CBLAS calls they replace.
static int stage_in(double *dst, const double *src, long rows, long cols, long ld) {
  if (rows <= 0 || cols <= 0) return CLU_OK;
  CLU_HIP(hipMemcpy2D(dst, cols * sizeof(double), src, ld * sizeof(double), cols * sizeof(double),
                      rows, hipMemcpyHostToDevice));
  return CLU_OK;
}
static int stage_out(double *dst, long ld, const double *src, long rows, long cols) {
  if (rows <= 0 || cols <= 0) return CLU_OK;
  CLU_HIP(hipMemcpy2D(dst, ld * sizeof(double), src, cols * sizeof(double), cols * sizeof(double),
                      rows, hipMemcpyDeviceToHost));
  return CLU_OK;
}

int charmlu_cblas_dgemm(int order, int transa, int transb, int M, int N, int K, double alpha,
                        const double *A, int lda, const double *B, int ldb, double beta,
                        double *C, int ldc) {
  if (order != CHARMLU_ROW_MAJOR && order != CHARMLU_COL_MAJOR) return -1;
  if (transa != CHARMLU_NO_TRANS && transa != CHARMLU_TRANS) return -2;
  if (transb != CHARMLU_NO_TRANS && transb != CHARMLU_TRANS) return -3;
  if (M < 0) return -4;
  if (N < 0) return -5;
  if (K < 0) return -6;
  if (order == CHARMLU_COL_MAJOR) {
    // column-major C = op(A) op(B)  <=>  row-major C^T = op(B)^T op(A)^T
    std::swap(transa, transb);
    std::swap(M, N);
    std::swap(A, B);
    std::swap(lda, ldb);
  }
  const bool ta = transa == CHARMLU_TRANS, tb = transb == CHARMLU_TRANS;
  const long ar = ta ? K : M, ac = ta ? M : K, br = tb ? N : K, bc = tb ? K : N;
  if (lda < (ac > 1 ? ac : 1)) return order == CHARMLU_COL_MAJOR ? -11 : -9;
  if (ldb < (bc > 1 ? bc : 1)) return order == CHARMLU_COL_MAJOR ? -9 : -11;
  if (ldc < (N > 1 ? N : 1)) return -14;
  if (M == 0 || N == 0) return CLU_OK;
  Workspace *ws;
  int rc = get_ws(&ws);
  if (rc) return rc;
  const size_t na = (size_t)ar * ac, nbb = (size_t)br * bc, nc = (size_t)M * N;
  if ((rc = ensure_stage(*ws, (na + nbb + nc) * sizeof(double)))) return rc;
  double *dA = (double *)ws->stage, *dB = dA + na, *dC = dB + nbb;
  if ((rc = stage_in(dA, A, ar, ac, lda)) || (rc = stage_in(dB, B, br, bc, ldb))) return rc;
  if (beta != 0.0 && (rc = stage_in(dC, C, M, N, ldc))) return rc;
  if ((rc = dgemm(ta, tb, M, N, K, alpha, dA, ac, dB, bc, beta, dC, N, nullptr))) return rc;
  if ((rc = stage_out(C, ldc, dC, M, N))) return rc;
  return CLU_OK;
}

int charmlu_cblas_dtrsm(int order, int side, int uplo, int transa, int diag, int M, int N,
                        double alpha, const double *A, int lda, double *B, int ldb) {
  if (order != CHARMLU_ROW_MAJOR && order != CHARMLU_COL_MAJOR) return -1;
  if (side != CHARMLU_LEFT && side != CHARMLU_RIGHT) return -2;
  if (uplo != CHARMLU_UPPER && uplo != CHARMLU_LOWER) return -3;
  if (transa != CHARMLU_NO_TRANS && transa != CHARMLU_TRANS) return -4;
  if (diag != CHARMLU_UNIT && diag != CHARMLU_NON_UNIT) return -5;
  if (M < 0) return -6;
  if (N < 0) return -7;
  bool upper = uplo == CHARMLU_UPPER;
  // column-major with the triangle on the right (the ESSL/ACML call of
  // lu.C:185) is the row-major left solve on the transposed views
  int rM = M, rN = N;
  if (order == CHARMLU_COL_MAJOR) {
    if (side != CHARMLU_RIGHT) { set_error("dtrsm: column-major needs side=Right"); return -2; }
    upper = !upper;
    rM = N;
    rN = M;
  } else if (side != CHARMLU_LEFT) {
    set_error("dtrsm: row-major needs side=Left");
    return -2;
  }
  if (lda < (rM > 1 ? rM : 1)) return -10;
  if (ldb < (rN > 1 ? rN : 1)) return -12;
  if (rM == 0 || rN == 0) return CLU_OK;
  Workspace *ws;
  int rc = get_ws(&ws);
  if (rc) return rc;
  const size_t nt = (size_t)rM * rM, nbb = (size_t)rM * rN;
  if ((rc = ensure_stage(*ws, (nt + nbb) * sizeof(double)))) return rc;
  double *dT = (double *)ws->stage, *dB = dT + nt;
  if ((rc = stage_in(dT, A, rM, rM, lda)) || (rc = stage_in(dB, B, rM, rN, ldb))) return rc;
  if ((rc = dtrsm_dev_any(upper, transa == CHARMLU_TRANS, diag == CHARMLU_UNIT, rM, rN, alpha, dT,
                          rM, dB, rN, nullptr)))
    return rc;
  return stage_out(B, ldb, dB, rM, rN);
}

int charmlu_cblas_dger(int order, int M, int N, double alpha, const double *x, int incx,
                       const double *y, int incy, double *A, int lda) {
  if (order != CHARMLU_ROW_MAJOR && order != CHARMLU_COL_MAJOR) return -1;
  if (M < 0) return -2;
  if (N < 0) return -3;
  if (incx == 0) return -6;
  if (incy == 0) return -8;
  if (order == CHARMLU_COL_MAJOR) {
    // column-major A += x y^T  <=>  row-major A^T += y x^T
    std::swap(M, N);
    std::swap(x, y);
    std::swap(incx, incy);
  }
  if (lda < (N > 1 ? N : 1)) return -10;
  if (M == 0 || N == 0) return CLU_OK;
  Workspace *ws;
  int rc = get_ws(&ws);
  if (rc) return rc;
  const long ax = labs((long)incx), ay = labs((long)incy);
  const size_t nx = (size_t)(M - 1) * ax + 1, ny = (size_t)(N - 1) * ay + 1, na = (size_t)M * N;
  if ((rc = ensure_stage(*ws, (nx + ny + na) * sizeof(double)))) return rc;
  double *dx = (double *)ws->stage, *dy = dx + nx, *dA = dy + ny;
  CLU_HIP(hipMemcpy(dx, x, nx * sizeof(double), hipMemcpyHostToDevice));
  CLU_HIP(hipMemcpy(dy, y, ny * sizeof(double), hipMemcpyHostToDevice));
  if ((rc = stage_in(dA, A, M, N, lda))) return rc;
  if ((rc = dger(M, N, alpha, dx, incx, dy, incy, dA, N, nullptr))) return rc;
  return stage_out(A, lda, dA, M, N);
}

int charmlu_cblas_dgemv(int order, int trans, int M, int N, double alpha, const double *A,
                        int lda, const double *x, int incx, double beta, double *y, int incy) {
  if (order != CHARMLU_ROW_MAJOR && order != CHARMLU_COL_MAJOR) return -1;
  if (trans != CHARMLU_NO_TRANS && trans != CHARMLU_TRANS) return -2;
  if (M < 0) return -3;
  if (N < 0) return -4;
  if (incx == 0) return -9;
  if (incy == 0) return -12;
  bool tr = trans == CHARMLU_TRANS;
  int rM = M, rN = N;
  if (order == CHARMLU_COL_MAJOR) {
    // column-major M x N is the row-major N x M transpose
    tr = !tr;
    rM = N;
    rN = M;
  }
  if (lda < (rN > 1 ? rN : 1)) return -7;
  const long lx = tr ? rM : rN, ly = tr ? rN : rM;  // op(A) is ly x lx
  if (ly == 0) return CLU_OK;
  Workspace *ws;
  int rc = get_ws(&ws);
  if (rc) return rc;
  const long ax = labs((long)incx), ay = labs((long)incy);
  const size_t nx = lx > 0 ? (size_t)(lx - 1) * ax + 1 : 1, ny = (size_t)(ly - 1) * ay + 1;
  const size_t na = (size_t)rM * rN;
  if ((rc = ensure_stage(*ws, (nx + ny + na + 1) * sizeof(double)))) return rc;
  double *dx = (double *)ws->stage, *dy = dx + nx, *dA = dy + ny;
  if (lx > 0) CLU_HIP(hipMemcpy(dx, x, nx * sizeof(double), hipMemcpyHostToDevice));
  CLU_HIP(hipMemcpy(dy, y, ny * sizeof(double), hipMemcpyHostToDevice));
  if ((rc = stage_in(dA, A, rM, rN, lda))) return rc;
  if (lx == 0) {
    // y = beta * y
    if ((rc = dgemv(false, (int)ly, 1, 0.0, dy, 1, dy, 1, beta, dy, incy, nullptr))) return rc;
  } else if ((rc = dgemv(tr, rM, rN, alpha, dA, rN, dx, incx, beta, dy, incy, nullptr))) {
    return rc;
  }
  CLU_HIP(hipMemcpy(y, dy, ny * sizeof(double), hipMemcpyDeviceToHost));
  return CLU_OK;
}

// ---- factorization ------------------------------------------------------------
int charmlu_dgetrf(int n, int nb, double *dA, long lda, int *dipiv, int *dinfo, void *stream) {
  if (n < 0) return -1;
  if (nb < 1 || (n > 0 && n % nb != 0)) {
    set_error("block size %d must divide n = %d (benchmark.C:143-147)", nb, n);
    return -2;
  }
  if (!dA && n > 0) return -3;
  if (lda < (n > 1 ? n : 1)) return -4;
  if (!dipiv && n > 0) return -5;
  if (!dinfo) return -6;
  hipStream_t su = (hipStream_t)stream;
  int rc;
  if ((rc = fill_int(dinfo, 0, 1, su))) return rc;
  if (n == 0) return CLU_OK;
  Workspace *wsp;
  if ((rc = get_ws(&wsp))) return rc;
  Workspace &ws = *wsp;
  if ((rc = check_panel_shape<double>(ws, n, true))) return rc;
  // all work runs on the library's streams, fenced to the caller's
  hipStream_t sm = ws.s_main;
  const int nruns = (nb + RUN - 1) / RUN;
  // inv(L11) * A12 on the matrix cores replaces the substitution
  const bool use_inv = nb <= 1024;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if ((rc = ensure_runs(ws, nruns))) return rc;
    if (use_inv) {
      if ((rc = ensure_dbuf(&ws.winv, &ws.winv_elems, (size_t)2 * nb * nb))) return rc;
      if ((rc = ensure_dbuf(&ws.ubuf, &ws.ubuf_elems, (size_t)nb * n))) return rc;
      if ((rc = ensure_bytes(&ws.tinv, &ws.tinv_b, sizeof(double) * nb * nb))) return rc;
    }
    if ((rc = ensure_bytes(&ws.pinv, &ws.pinv_b, sizeof(double) * nb * nb / 4))) return rc;
    if ((rc = ensure_bytes(&ws.ptmp, &ws.ptmp_b, sizeof(double) * nb * nb / 4))) return rc;
  }
  // inv(L11) of step t, by step parity: the panel of step t+1 writes its
  // own while the U row of step t still reads this one
  auto winv_of = [&](int t) { return ws.winv + (size_t)(t & 1) * nb * nb; };
  hipStream_t sp = ws.s_panel;
  ws.pw.info = dinfo;
  if ((rc = fill_int((int *)ws.abortw, 0, 1, su))) return rc;
  CLU_HIP(hipEventRecord(ws.ev_start, su));
  CLU_HIP(hipStreamWaitEvent(sp, ws.ev_start, 0));
  CLU_HIP(hipStreamWaitEvent(sm, ws.ev_start, 0));

  const int NB = n / nb;
  auto moves_of = [&](int t) { return ws.smoves + (size_t)(t & 1) * nruns * 2 * RUN; };
  auto counts_of = [&](int t) { return ws.scount + (size_t)(t & 1) * nruns; };
  auto panel = [&](int t) -> int {
    const int r0 = t * nb;
    const int pid = prof_begin(PK_PANEL, (double)(n - r0) * nb, sp);
    int r = factor_panel<double>(dA, lda, n, r0, nb, r0, nb, dipiv, ws, sp);
    if (r) return r;
    r = build_moves(dipiv, r0, r0 + nb, moves_of(t), counts_of(t), sp);
    if (r) return r;
    if (use_inv && r0 + nb < n)  // inv(L11), off the update stream's path
      r = tri_inv<double>(false, true, nb, 1, dA + (long)r0 * lda + r0, lda, 0, winv_of(t), nb, 0,
                          (double *)ws.tinv, sp);
    prof_end(pid, sp);
    return r;
  };
  // the look-ahead block column, timed when profiling is on
  auto update_la = [&](int M, int N, const double *L, const double *U, long ldu, double *C,
                       hipStream_t st) -> int {
    const int pid = prof_begin(PK_GEMM_OTHER, 2.0 * M * N * nb, st);
    int r = dgemm_lookahead(M, N, nb, -1.0, L, lda, U, ldu, 1.0, C, lda, st);
    prof_end(pid, st);
    return r;
  };
  // The rest of the trailing update (updateMatrix, lu.C:26-43).  With the
  // panel on CUs of its own, the update is split by rows: the first Ma rows
  // run on the update stream beside the panel of step t+1, the rest on every
  // CU (s_full) once that panel is done.  Ma is sized so that the first part
  // lasts about as long as the panel (its duration in the previous call with
  // this n and block size, else a model), so neither set of CUs waits for
  // long.  Each tile is computed the same way whatever the split.
  const bool split = ws.panel_cus > 0;
  const bool force_split = split_forced();
  bool full_pending = false;
  auto split_rows = [&](int t, int M, int N, const double *L, const double *U, long ldu,
                        const double *C) -> int {
    if (!split || !dgemm_trailing_lds(M, N, nb, L, lda, U, ldu, C, lda)) return M;
    if (force_split) return forced_split(M);
    const int m = n - (t + 1) * nb;  // rows of panel t+1
    double p_ms = (t + 1 < (int)ws.hist64.ms.size()) ? ws.hist64.ms[t + 1] : -1.0;
    if (p_ms <= 0.0) p_ms = nb / 256.0 * (1.3 + 2.0 * (m < 32768 ? m / 32768.0 : 1.0));
    // FP64 matrix-core rate of the update stream's CUs at 80 % of peak,
    // plus 5 % margin (with 10 % the first part outlasted the panel more
    // often and overlapped the second for longer: per launch 0.73 against
    // 0.75 of peak at n = 32768, the LU within 0.2 %; with 0 %, 0.74 and
    // 0.9 % slower; with 20 %, 0.70 and 0.1 % faster)
    const double rate = 0.8 * 78.6e9 * (ws.ncu - ws.panel_cus) / ws.ncu;  // flop per ms
    const double rows = 1.05 * p_ms * rate / (2.0 * N * nb);
    if (rows >= M) return M;
    const int Ma = ((int)rows + 127) / 128 * 128;
    return Ma > M ? M : Ma;
  };
  // part 1: rows [0, Ma) on the update stream, enqueued before the panel's
  // many small launches so the host does not delay it
  auto update_part1 = [&](int Ma, int N, const double *L, const double *U, long ldu,
                          double *C) -> int {
    if (Ma <= 0) return CLU_OK;
    const int pid = prof_begin(PK_GEMM_BIG, 2.0 * Ma * N * nb, sm);
    int r = dgemm_trailing(Ma, N, nb, -1.0, L, lda, U, ldu, 1.0, C, lda, sm);
    prof_end(pid, sm);
    return r;
  };
  // part 2: rows [Ma, M) on every CU once the panel (ev_panel) is done
  auto update_part2 = [&](int Ma, int M, int N, const double *L, const double *U, long ldu,
                          double *C) -> int {
    if (Ma >= M) return CLU_OK;
    hipStream_t sf = ws.s_full;
    CLU_HIP(hipStreamWaitEvent(sf, ws.ev_urow, 0));
    CLU_HIP(hipStreamWaitEvent(sf, ws.ev_panel, 0));
    const int pid = prof_begin(PK_GEMM_BIG, 2.0 * (M - Ma) * N * nb, sf);
    int r = dgemm_trailing(M - Ma, N, nb, -1.0, L + (long)Ma * lda, lda, U, ldu, 1.0,
                           C + (long)Ma * lda, lda, sf);
    prof_end(pid, sf);
    if (r) return r;
    CLU_HIP(hipEventRecord(ws.ev_full, sf));
    full_pending = true;
    return CLU_OK;
  };
  if ((rc = history_prepare(ws.hist64, n, nb, 0, NB))) {
    fence_streams(ws, su);
    return rc;
  }
  rc = [&]() -> int {  // the steps; an error returns from here to the fence below
  int rc;
  if ((rc = panel(0))) return rc;
  CLU_HIP(hipEventRecord(ws.ev_panel, sp));
  for (int t = 0; t < NB; ++t) {
    const int r0 = t * nb, r1 = r0 + nb, Nt = n - r1;
    CLU_HIP(hipStreamWaitEvent(sm, ws.ev_panel, 0));
    if (full_pending) {  // the previous update's second launch
      CLU_HIP(hipStreamWaitEvent(sm, ws.ev_full, 0));
      full_pending = false;
    }
    // Interchanges of step t on every column outside the panel and the U
    // row, A12 <- L11^{-1} A12 (computeU, lu.C:180-189): first for the
    // look-ahead block column alone, whose update the next panel waits for,
    // then for the other columns while that panel runs.
    const int la = Nt < nb ? Nt : nb;
    double *A11 = dA + (long)r0 * lda + r0;
    const double *U = A11 + nb;  // the U row block the trailing update reads
    long ldu = lda;
    if (use_inv) {
      U = ws.ubuf;
      ldu = Nt;
    }
    auto swap_cols = [&](int c0, int c1, int skip0, int skip1) -> int {
      const int pid = prof_begin(PK_SWAP, (double)(c1 - c0 - (skip1 - skip0)), sm);
      int e = apply_runs<double>(nruns, moves_of(t), counts_of(t), dA, lda, c0, c1, skip0, skip1,
                                 false, sm);
      prof_end(pid, sm);
      return e;
    };
    auto u_row = [&](int j0, int j1) -> int {  // U row columns [j0, j1) of the trailing part
      if (j1 <= j0) return CLU_OK;
      const int pid = prof_begin(PK_TRSM, (double)(j1 - j0) * nb * (nb - 1), sm);
      int e = use_inv ? dgemm(false, false, nb, j1 - j0, nb, 1.0, winv_of(t), nb, A11 + nb + j0,
                              lda, 0.0, ws.ubuf + j0, Nt, sm)
                      : trsm_llnu_big<double>(nb, j1 - j0, A11, lda, A11 + nb + j0, lda, sm);
      prof_end(pid, sm);
      return e;
    };
    if (Nt <= 0) {
      if ((rc = swap_cols(0, n, r0, r1))) return rc;
      break;
    }
    if ((rc = swap_cols(r1, r1 + la, r1 + la, r1 + la))) return rc;
    if ((rc = u_row(0, la))) return rc;
    // look-ahead block column, then the panel of step t+1 on the panel
    // stream.  While the trailing matrix is wide, the look-ahead GEMM runs
    // on the stream of every CU, beside the rest of the interchanges and of
    // the U row (other columns), which part 1 of the update waits for; in
    // the narrow steps the panel waits on it, and an extra stream hop there
    // costs more than the overlap saves.
    const double *L21 = A11 + (long)nb * lda;
    const bool la_beside = Nt >= 16384;
    if (la_beside) {
      CLU_HIP(hipEventRecord(ws.ev_ula, sm));
      CLU_HIP(hipStreamWaitEvent(ws.s_full, ws.ev_ula, 0));
    }
    hipStream_t sla = la_beside ? ws.s_full : sm;
    if ((rc = update_la(Nt, la, L21, U, ldu, A11 + (long)nb * lda + nb, sla))) return rc;
    CLU_HIP(hipEventRecord(ws.ev_la, sla));
    if ((rc = swap_cols(0, n, r0, r1 + la))) return rc;
    if ((rc = u_row(la, Nt))) return rc;
    CLU_HIP(hipEventRecord(ws.ev_urow, sm));
    double *C2 = A11 + (long)nb * lda + nb + la;
    const int Ma = Nt > la ? split_rows(t, Nt, Nt - la, L21, U + la, ldu, C2) : 0;
    if (Nt > la && (rc = update_part1(Ma, Nt - la, L21, U + la, ldu, C2))) return rc;
    CLU_HIP(hipStreamWaitEvent(sp, ws.ev_la, 0));
    CLU_HIP(hipEventRecord(ws.hist64.ev[2 * (t + 1)], sp));
    if ((rc = panel(t + 1))) return rc;
    CLU_HIP(hipEventRecord(ws.hist64.ev[2 * (t + 1) + 1], sp));
    CLU_HIP(hipEventRecord(ws.ev_panel, sp));
    if (Nt > la && (rc = update_part2(Ma, Nt, Nt - la, L21, U + la, ldu, C2))) return rc;
    // the U row block into the factor storage (nothing reads A12 before)
    if (use_inv && (rc = convert<double, double>(nb, Nt, ws.ubuf, Nt, A11 + nb, lda, sm)))
      return rc;
  }
  if (full_pending) CLU_HIP(hipStreamWaitEvent(sm, ws.ev_full, 0));
  // info = -1 if a strip's exchange gave up
  if ((rc = finish_info(dinfo, ws.abortw, sm))) return rc;
  CLU_HIP(hipEventRecord(ws.ev_end, sm));
  CLU_HIP(hipStreamWaitEvent(su, ws.ev_end, 0));
  return CLU_OK;
  }();
  if (rc) fence_streams(ws, su);
  return rc;
}

// ---- solve ---------------------------------------------------------------------
int charmlu_dgetrs(int trans, int n, int nrhs, const double *dLU, long lda, const int *dipiv,
                   double *dB, long ldb, void *stream) {
  return getrs_impl<double>(trans, n, nrhs, dLU, lda, dipiv, dB, ldb, (hipStream_t)stream);
}

int charmlu_sgetrs(int trans, int n, int nrhs, const float *dLU, long lda, const int *dipiv,
                   float *dB, long ldb, void *stream) {
  return getrs_impl<float>(trans, n, nrhs, dLU, lda, dipiv, dB, ldb, (hipStream_t)stream);
}

// ---- mixed-precision factorization -------------------------------------------
// The right-looking blocked LU of charmlu_dgetrf on an FP32 working matrix:
// FP32 panels (the same strip kernel and recursive halving), the U row by
// inv(L11) in FP32, and the trailing update C -= L21 U12 with L21 and U12
// rounded to BF16 or FP16 and products accumulated in FP32 on the matrix
// cores (or a plain FP32 update).  The factors precondition the FP64
// refinement of charmlu_dgesv_ir, which restores FP64 accuracy.
int charmlu_sgetrf_mixed(int n, int nb, int prec, float *dA, long lda, int *dipiv, int *dinfo,
                         void *stream) {
  if (n < 0) return -1;
  if (nb < 1 || (n > 0 && n % nb != 0)) {
    set_error("block size %d must divide n = %d (benchmark.C:143-147)", nb, n);
    return -2;
  }
  const bool nopiv = (prec & CHARMLU_MIXED_NOPIV) != 0;
  prec &= ~CHARMLU_MIXED_NOPIV;
  if (prec != CHARMLU_PREC_FP32 && prec != CHARMLU_PREC_BF16 && prec != CHARMLU_PREC_FP16 &&
      prec != CHARMLU_PREC_FP8) {
    set_error("trailing-update precision %d is not supported", prec);
    return -3;
  }
  if (prec == CHARMLU_PREC_FP8 && nb % 128 != 0) {
    set_error("the FP8 update needs a block size that is a multiple of 128 (got %d)", nb);
    return -2;
  }
  if (nopiv && nb > 2048) {
    set_error("the unpivoted form needs block size <= 2048 (got %d)", nb);
    return -2;
  }
  if (!dA && n > 0) return -4;
  if (lda < (n > 1 ? n : 1)) return -5;
  if (!dipiv && n > 0) return -6;
  if (!dinfo) return -7;
  hipStream_t su = (hipStream_t)stream;
  int rc;
  if ((rc = fill_int(dinfo, 0, 1, su))) return rc;
  if (n == 0) return CLU_OK;
  Workspace *wsp;
  if ((rc = get_ws(&wsp))) return rc;
  Workspace &ws = *wsp;
  if ((rc = check_panel_shape<float>(ws, n, true))) return rc;
  hipStream_t sm = ws.s_main, sp = ws.s_panel;
  const int nruns = (nb + RUN - 1) / RUN;
  const bool lp = prec != CHARMLU_PREC_FP32;
  const bool fp8 = prec == CHARMLU_PREC_FP8;
  // Unpivoted low-precision form: the next step's L21, U row and packed
  // operands are computed beside the second part of this step's update, so
  // the packed operands (and FP8 scales) alternate between two buffers by
  // step parity (the second part of step t still reads step t's)
  const bool early = lp && nopiv;
  // ... and its steps run in groups of G that update the rest of the matrix
  // once (below).  Measured (block 1024, pivoted with speculative identity
  // pivots), G = 4, 8, 12, 16: n = 65536 BF16 444, 458, 467, 451 and FP8
  // 537, 552, 561, 565 TFLOP/s; n = 131072 FP8 891, 952, 966, 975
  // The pivoted low-precision form runs the grouped unpivoted schedule as
  // long as its panels show that partial pivoting would keep every diagonal
  // row (below; its unpivoted panels take nb <= 2048)
  const bool spec_req = lp && !nopiv && nb <= 2048 && speculate_enabled();
  const int G = (early || spec_req) ? group_steps(fp8 ? 16 : 12) : 1;
  const bool grouped = G > 1;
  const bool spec = grouped && !nopiv;
  const size_t nbuf = grouped ? 2 * G : early ? 2 : 1;
  // K of the low-precision operands, zero-padded (64 bytes per LDS stage row)
  const int Kp = fp8 ? (nb + 63) / 64 * 64 : (nb + 31) / 32 * 32;
  const size_t pk_bytes = (size_t)(fp8 ? 1 : 2) * n * Kp;  // one packed operand
  const size_t esc_ints = 2 * (size_t)(n / 128 + 2) * (1 + FP8_PART_PER_TILE);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if ((rc = ensure_runs(ws, nruns))) return rc;
    if ((rc = ensure_bytes(&ws.winv32, &ws.winv32_b, sizeof(float) * 2 * nb * nb))) return rc;
    if ((rc = ensure_bytes(&ws.ubuf32, &ws.ubuf32_b, sizeof(float) * nb * n))) return rc;
    if ((rc = ensure_bytes(&ws.tinv, &ws.tinv_b, sizeof(float) * nb * nb))) return rc;
    if ((rc = ensure_bytes(&ws.pinv, &ws.pinv_b, sizeof(float) * nb * nb / 4))) return rc;
    if ((rc = ensure_bytes(&ws.ptmp, &ws.ptmp_b, sizeof(float) * nb * nb / 4))) return rc;
    if (lp) {
      if ((rc = ensure_bytes(&ws.pa16, &ws.pa16_b, nbuf * pk_bytes))) return rc;
      if ((rc = ensure_bytes(&ws.pu16, &ws.pu16_b, nbuf * pk_bytes))) return rc;
    }
    // per operand: one exponent and FP8_PART_PER_TILE partial maxima per
    // 128-row tile
    if (fp8 && (rc = ensure_bytes(&ws.esc, &ws.esc_b, sizeof(int) * nbuf * esc_ints))) return rc;
    if (nopiv || spec) {
      if ((rc = ensure_bytes(&ws.uinv32, &ws.uinv32_b, sizeof(float) * nb * nb))) return rc;
      if ((rc = ensure_bytes(&ws.l21buf, &ws.l21buf_b, sizeof(float) * nb * n))) return rc;
    }
    if (spec && (rc = ensure_bytes(&ws.specbak, &ws.specbak_b, sizeof(float) * 4 * nb * nb)))
      return rc;
  }
  float *ubuf = (float *)ws.ubuf32;
  // inv(L11) of step t, by step parity: the panel of step t+1 writes its
  // own while the rest of step t's U row still reads this one
  auto winv_of = [&](int t) { return (float *)ws.winv32 + (size_t)(t & 1) * nb * nb; };
  ws.pw.info = dinfo;
  if ((rc = fill_int((int *)ws.abortw, 0, 1, su))) return rc;
  CLU_HIP(hipEventRecord(ws.ev_start, su));
  CLU_HIP(hipStreamWaitEvent(sp, ws.ev_start, 0));
  CLU_HIP(hipStreamWaitEvent(sm, ws.ev_start, 0));

  const int NB = n / nb;
  auto moves_of = [&](int t) { return ws.smoves + (size_t)(t & 1) * nruns * 2 * RUN; };
  auto counts_of = [&](int t) { return ws.scount + (size_t)(t & 1) * nruns; };
  auto panel = [&](int t, bool unpiv = false) -> int {
    const int r0 = t * nb;
    const int pid = prof_begin(PK_PANEL, (double)(n - r0) * nb, sp);
    int r;
    if (nopiv || unpiv) {
      // without interchanges the panel splits into the LU of the diagonal
      // block (here) and L21 = A21 inv(U11), one GEMM over all rows below
      // that the update stream runs with all its CUs at the start of step t
      float *A11 = dA + (long)r0 * lda + r0;
      if ((r = factor_panel<float>(dA, lda, r0 + nb, r0, nb, r0, nb, dipiv, ws, sp, false)))
        return r;
      if (r0 + nb < n &&
          (r = tri_inv<float>(true, false, nb, 1, A11, lda, 0, (float *)ws.uinv32, nb, 0,
                              (float *)ws.tinv, sp)))
        return r;
    } else {
      if ((r = factor_panel<float>(dA, lda, n, r0, nb, r0, nb, dipiv, ws, sp))) return r;
      if ((r = build_moves(dipiv, r0, r0 + nb, moves_of(t), counts_of(t), sp))) return r;
    }
    if (r0 + nb < n)
      r = tri_inv<float>(false, true, nb, 1, dA + (long)r0 * lda + r0, lda, 0, winv_of(t), nb, 0,
                         (float *)ws.tinv, sp);
    prof_end(pid, sp);
    return r;
  };
  // C (M x N) -= L21 (M x nb) * U (nb x N) for the rows from r_off of L21
  // (C already offset); U column c0 is column c0 of the packed transposed
  // operand; the 16-bit L21 is packed negated, so the low-precision kernel
  // accumulates onto C (C += (-L21) U).  FP8 row offsets are multiples of
  // 128 (one scale per 128-row tile).
  // packed operands and FP8 exponents (then partial maxima) of step t
  auto pa_of = [&](int t) { return (char *)ws.pa16 + (size_t)(t & 1) * (nbuf - 1) * pk_bytes; };
  auto pu_of = [&](int t) { return (char *)ws.pu16 + (size_t)(t & 1) * (nbuf - 1) * pk_bytes; };
  auto esc_of = [&](int t) { return (int *)ws.esc + (size_t)(t & 1) * (nbuf - 1) * esc_ints; };
  auto update = [&](int t, int M, int N, int r_off, const float *L, long ldl, int c0, float *C,
                    int Nt, hipStream_t st) -> int {
    if (M <= 0 || N <= 0) return CLU_OK;
    const int pid = prof_begin(PK_GEMM_BIG, 2.0 * M * N * nb, st);
    const size_t ro = (size_t)r_off * Kp;
    int r;
    if (!lp)
      r = sgemm(false, false, M, N, nb, -1.0f, L + (long)r_off * ldl, ldl, ubuf + c0, Nt, 1.0f, C,
                lda, st);
    else if (fp8)
      r = fgemm_nt(M, N, Kp, 1.0f, (const uint8_t *)pa_of(t) + ro, Kp, esc_of(t) + r_off / 128,
                   (const uint8_t *)pu_of(t) + (size_t)c0 * Kp, Kp,
                   esc_of(t) + (n / 128 + 2) + c0 / 128, 1.0f, C, lda, st);
    else if (prec == CHARMLU_PREC_BF16)
      r = hgemm_nt<__bf16>(M, N, Kp, 1.0f, (const __bf16 *)pa_of(t) + ro, Kp,
                           (const __bf16 *)pu_of(t) + (size_t)c0 * Kp, Kp, 1.0f, C, lda, st);
    else
      r = hgemm_nt<_Float16>(M, N, Kp, 1.0f, (const _Float16 *)pa_of(t) + ro, Kp,
                             (const _Float16 *)pu_of(t) + (size_t)c0 * Kp, Kp, 1.0f, C, lda, st);
    prof_end(pid, st);
    return r;
  };
  // Rows of the trailing update (after the look-ahead column) that run on
  // the update stream beside the panel of the next step; the rest waits for
  // that panel and runs on every CU.  Sized as in the FP64 form: the panel's
  // duration in the previous call of the same kind and shape, else a model
  // (FP32 1024-column panels: 1.5 + 14 m/65536 ms pivoted, 3.5 ms
  // unpivoted), times the update rate of the stream's CUs, plus 10 %.
  const bool split = ws.panel_cus > 0;
  const bool force_split = split_forced();
  bool full_pending = false;
  // the update's rate inside the LU (flop per ms), as measured: FP8 1.0
  // PFLOP/s at block 2048 and 0.73 at 1024, BF16 and FP16 0.51, FP32 0.6 of
  // the FP32 peak.  (With 0.53 for FP8 the first part ended long before the
  // panel and the second, which the next look-ahead update waits for in the
  // pivoted form, grew: C5 pivoted 547-559 against 602-607 TFLOP/s.  The
  // margin of 1.1 over the panel: 0.9, 1.3 and 1.6 measured within 1 %.)
  const double chip_rate =
      !lp ? 0.6 * 157.3e9 : fp8 ? (nb >= 2048 ? 1000e9 : 730e9) : 510e9;
  auto split_rows = [&](int t, int M, int N) -> int {
    if (!split) return M;
    if (force_split) return forced_split(M);
    const int m = n - (t + 1) * nb;
    double p_ms = (t + 1 < (int)ws.histmx.ms.size()) ? ws.histmx.ms[t + 1] : -1.0;
    if (p_ms <= 0.0)
      p_ms = nb / 1024.0 * (nopiv ? 3.5 : 1.5 + 14.0 * (m < 65536 ? m / 65536.0 : 1.0));
    const double rate = chip_rate * (ws.ncu - ws.panel_cus) / ws.ncu;
    const double rows = 1.1 * p_ms * rate / (2.0 * N * nb);
    if (rows >= M) return M;
    const int Ma = ((int)rows + 127) / 128 * 128;
    return Ma > M ? M : Ma;
  };
  // the first part always holds the next step's block row when that step's
  // U row is computed beside the second part
  auto split_rows_of = [&](int t, int M, int N) -> int {
    const int Ma = split_rows(t, M, N);
    return early && Ma < nb ? (nb < M ? nb : M) : Ma;
  };
  // (the speculative schedule's panels are the unpivoted ones: their
  // durations, not those of the exchanging panels, size its updates)
  if ((rc = history_prepare(ws.histmx, n, nb, 1 + 2 * prec + ((nopiv || spec) ? 1 : 0), NB))) {
    fence_streams(ws, su);
    return rc;
  }

  // the steps from t0 on, one trailing update per step (panel(t0) enqueued
  // and ev_panel recorded)
  auto per_step = [&](int t0) -> int {
  int rc;
  for (int t = t0; t < NB; ++t) {
    const int r0 = t * nb, r1 = r0 + nb, Nt = n - r1;
    CLU_HIP(hipStreamWaitEvent(sm, ws.ev_panel, 0));
    if (full_pending && !early) {  // the previous update's second part
      CLU_HIP(hipStreamWaitEvent(sm, ws.ev_full, 0));
      full_pending = false;
    }
    const int la = Nt < nb ? Nt : nb;
    // Interchanges, U row and packed operands: first for the look-ahead
    // block column (whose update the next panel waits for), the rest after
    // that update, while the panel runs (as in the FP64 form)
    auto swap_cols = [&](int c0, int c1, int skip0, int skip1) -> int {
      if (nopiv) return CLU_OK;
      const int p = prof_begin(PK_SWAP, (double)(c1 - c0 - (skip1 - skip0)), sm);
      int e = apply_runs<float>(nruns, moves_of(t), counts_of(t), dA, lda, c0, c1, skip0, skip1,
                                false, sm);
      prof_end(p, sm);
      return e;
    };
    if (Nt <= 0) {
      if ((rc = swap_cols(0, n, r0, r1))) return rc;
      break;
    }
    // early: L21 and the U row of the look-ahead block column (with their
    // packed operands) need the panel and the first part of the previous
    // update (which holds this block row), not its second part, which they
    // run beside
    if ((rc = swap_cols(r1, r1 + la, r1 + la, r1 + la))) return rc;
    float *A11 = dA + (long)r0 * lda + r0;
    // L21: in the matrix, or (unpivoted form) A21 inv(U11) in l21buf
    const float *L21 = A11 + (long)nb * lda;
    long ldl21 = lda;
    int pid;
    if (nopiv) {
      float *l21 = (float *)ws.l21buf;
      pid = prof_begin(PK_GEMM_OTHER, 2.0 * Nt * nb * nb, sm);
      if ((rc = gemm_upper_right<float>(Nt, nb, A11 + (long)nb * lda, lda,
                                        (const float *)ws.uinv32, nb, l21, nb, sm)))
        return rc;
      prof_end(pid, sm);
      L21 = l21;
      ldl21 = nb;
    }
    // U row columns [j0, j1) and their packed operand (the packed L21 with
    // the first part)
    auto u_row = [&](int j0, int j1) -> int {
      if (j1 <= j0) return CLU_OK;
      const int p = prof_begin(PK_TRSM, (double)(j1 - j0) * nb * (nb - 1), sm);
      int e = gemm_lower_left<float>(nb, j1 - j0, winv_of(t), nb, A11 + nb + j0, lda, ubuf + j0,
                                     Nt, sm);
      // the pass that reads the U row (and, unpivoted, L21 from its
      // buffer) for the packed operand also stores it into the factor
      float *ucp = A11 + nb + j0, *lcp = nopiv ? A11 + (long)nb * lda : nullptr;
      if (!e && fp8) {
        // FP8 operands with power-of-two scales per 128-row tile of L21 and
        // per 128-column tile of U12
        const int T = n / 128 + 2;
        int *escL = esc_of(t), *escU = escL + T;
        float *partL = (float *)(escU + T), *partU = partL + (size_t)FP8_PART_PER_TILE * T;
        if (j0 == 0)
          e = quant_rows_fp8(Nt, nb, Kp, -1.0f, L21, ldl21, (uint8_t *)pa_of(t), escL, partL, sm,
                             lcp, lda);
        if (!e)
          e = quant_cols_fp8(nb, j1 - j0, Kp, ubuf + j0, Nt, (uint8_t *)pu_of(t) + (size_t)j0 * Kp,
                             escU + j0 / 128, partU + (size_t)FP8_PART_PER_TILE * (j0 / 128), sm,
                             ucp, lda);
      } else if (!e && lp) {
        // operands of the trailing update, rounded to 16 bits once per step
        if (prec == CHARMLU_PREC_BF16) {
          if (j0 == 0)
            e = pack_rows<__bf16>(Nt, nb, Kp, -1.0f, L21, ldl21, (__bf16 *)pa_of(t), sm, lcp, lda);
          if (!e)
            e = pack_cols<__bf16>(nb, j1 - j0, Kp, ubuf + j0, Nt,
                                  (__bf16 *)pu_of(t) + (size_t)j0 * Kp, sm, ucp, lda);
        } else {
          if (j0 == 0)
            e = pack_rows<_Float16>(Nt, nb, Kp, -1.0f, L21, ldl21, (_Float16 *)pa_of(t), sm, lcp,
                                    lda);
          if (!e)
            e = pack_cols<_Float16>(nb, j1 - j0, Kp, ubuf + j0, Nt,
                                    (_Float16 *)pu_of(t) + (size_t)j0 * Kp, sm, ucp, lda);
        }
      }
      prof_end(p, sm);
      return e;
    };
    if ((rc = u_row(0, la))) return rc;
    if (full_pending) {  // early: the previous update's second part, now
      CLU_HIP(hipStreamWaitEvent(sm, ws.ev_full, 0));
      full_pending = false;
    }
    float *C = A11 + (long)nb * lda + nb;
    if ((rc = update(t, Nt, la, 0, L21, ldl21, 0, C, Nt, sm))) return rc;
    CLU_HIP(hipEventRecord(ws.ev_la, sm));
    if ((rc = swap_cols(0, n, r0, r1 + la))) return rc;
    if ((rc = u_row(la, Nt))) return rc;
    CLU_HIP(hipEventRecord(ws.ev_urow, sm));  // operands of this step's update ready
    // part 1 beside the panel (enqueued before the panel's launches), part 2
    // on every CU after it
    const int Ma = Nt > la ? split_rows_of(t, Nt, Nt - la) : 0;
    if ((rc = update(t, Ma, Nt - la, 0, L21, ldl21, la, C + la, Nt, sm))) return rc;
    CLU_HIP(hipStreamWaitEvent(sp, ws.ev_la, 0));
    CLU_HIP(hipEventRecord(ws.histmx.ev[2 * (t + 1)], sp));
    if ((rc = panel(t + 1))) return rc;
    CLU_HIP(hipEventRecord(ws.histmx.ev[2 * (t + 1) + 1], sp));
    CLU_HIP(hipEventRecord(ws.ev_panel, sp));
    if (Nt > la && Ma < Nt) {
      hipStream_t sf = ws.s_full;
      CLU_HIP(hipStreamWaitEvent(sf, ws.ev_urow, 0));
      CLU_HIP(hipStreamWaitEvent(sf, ws.ev_panel, 0));
      if ((rc = update(t, Nt - Ma, Nt - la, Ma, L21, ldl21, la, C + (long)Ma * lda + la, Nt,
                       sf)))
        return rc;
      CLU_HIP(hipEventRecord(ws.ev_full, sf));
      full_pending = true;
    }
    // the FP32 update's U row and L21 into the factor storage (the
    // low-precision forms stored them while packing)
    if (!lp && (rc = convert<float, float>(nb, Nt, ubuf, Nt, A11 + nb, lda, sm))) return rc;
    if (!lp && nopiv &&
        (rc = convert<float, float>(Nt, nb, L21, ldl21, A11 + (long)nb * lda, lda, sm)))
      return rc;
  }
  if (full_pending) CLU_HIP(hipStreamWaitEvent(sm, ws.ev_full, 0));
  if ((rc = finish_info(dinfo, ws.abortw, sm))) return rc;
  CLU_HIP(hipEventRecord(ws.ev_end, sm));
  CLU_HIP(hipStreamWaitEvent(su, ws.ev_end, 0));
  return CLU_OK;
  };

  // ---- unpivoted low-precision form in groups of steps --------------------
  // The G steps a = G p .. a + G - 1 of group p update the rest of the matrix
  // (rows and columns from (a + G) nb on) together, K = G nb: their packed
  // operands lie side by side in one buffer per group (K segment i: step
  // a + i), so the FP32 C is read and written once per G steps (the update's
  // C traffic takes as long as its K loop at K = nb).  Inside the group each
  // step first brings the next block column and block row up to date with
  // the group's steps so far (K = (i + 1) nb), which the next panel and U
  // row need.  The group's update does the next group's G block columns and
  // block rows first; the rest of it runs beside the next group's panels on
  // the update stream's CUs and after them on every CU, split by rows, and
  // every element of it is updated by the same launch shape wherever the
  // split falls.
  const long ldg = (long)G * Kp;
  const int T = n / 128 + 2;
  auto PA = [&](int p) { return (char *)ws.pa16 + (size_t)(p & 1) * G * pk_bytes; };
  auto PU = [&](int p) { return (char *)ws.pu16 + (size_t)(p & 1) * G * pk_bytes; };
  auto ESC = [&](int p, int sg) { return (int *)ws.esc + (size_t)(G * (p & 1) + sg) * esc_ints; };
  // C (M x N from global row gr, column gc) += segments [0, ns) of group p
  auto updg = [&](int p, int ns, int gr, int gc, int M, int N, hipStream_t st) -> int {
    if (M <= 0 || N <= 0) return CLU_OK;
    const int pid = prof_begin(PK_GEMM_BIG, 2.0 * M * N * nb * ns, st);
    float *C = dA + (long)gr * lda + gc;
    const size_t ao = (size_t)gr * ldg, bo = (size_t)gc * ldg;
    const int K = ns * Kp;
    int r;
    if (fp8)
      r = fgemm_nt(M, N, K, 1.0f, (const uint8_t *)PA(p) + ao, ldg, ESC(p, 0) + gr / 128,
                   (const uint8_t *)PU(p) + bo, ldg, ESC(p, 0) + T + gc / 128, 1.0f, C, lda, st,
                   Kp, ns, (long)esc_ints);
    else if (prec == CHARMLU_PREC_BF16)
      r = hgemm_nt<__bf16>(M, N, K, 1.0f, (const __bf16 *)PA(p) + ao, ldg,
                           (const __bf16 *)PU(p) + bo, ldg, 1.0f, C, lda, st);
    else
      r = hgemm_nt<_Float16>(M, N, K, 1.0f, (const _Float16 *)PA(p) + ao, ldg,
                             (const _Float16 *)PU(p) + bo, ldg, 1.0f, C, lda, st);
    prof_end(pid, st);
    return r;
  };
  // rows [row0, row0 + rows) of L21 = A21 inv(U11) of step t into the same
  // rows of l21buf
  auto l21_rows = [&](int t, int row0, int rows) -> int {
    if (rows <= 0) return CLU_OK;
    const int r0 = t * nb;
    float *A11 = dA + (long)r0 * lda + r0;
    const int pid = prof_begin(PK_GEMM_OTHER, 2.0 * rows * nb * nb, sm);
    int e = gemm_upper_right<float>(rows, nb, A11 + (long)(nb + row0) * lda, lda,
                                    (const float *)ws.uinv32, nb,
                                    (float *)ws.l21buf + (size_t)row0 * nb, nb, sm);
    prof_end(pid, sm);
    return e;
  };
  // U row of step t, columns [j0, j1) after its block, into ubuf
  auto urow_gemm = [&](int t, int j0, int j1) -> int {
    const int r0 = t * nb, Nt = n - r0 - nb;
    float *A11 = dA + (long)r0 * lda + r0;
    return gemm_lower_left<float>(nb, j1 - j0, winv_of(t), nb, A11 + nb + j0, lda, ubuf + j0, Nt,
                                  sm);
  };
  // those columns packed into K segment t % G of group t / G (global columns;
  // FP8 scales per global 128-column tile), with copy also stored into the
  // factor by the same pass
  auto pack_u = [&](int t, int j0, int j1, bool copy) -> int {
    const int p = t / G, sg = t % G, r0 = t * nb, r1 = r0 + nb, Nt = n - r1;
    if (j1 <= j0) return CLU_OK;
    float *ucp = copy ? dA + (long)r0 * lda + r0 + nb + j0 : nullptr;
    const size_t co = (size_t)(r1 + j0) * ldg + (size_t)sg * Kp;
    if (fp8) {
      int *escU = ESC(p, sg) + T;
      float *partU = (float *)(escU + T) + (size_t)FP8_PART_PER_TILE * T;
      return quant_cols_fp8(nb, j1 - j0, Kp, ubuf + j0, Nt, (uint8_t *)PU(p) + co,
                            escU + (r1 + j0) / 128,
                            partU + (size_t)FP8_PART_PER_TILE * ((r1 + j0) / 128), sm, ucp, lda,
                            ldg);
    }
    if (prec == CHARMLU_PREC_BF16)
      return pack_cols<__bf16>(nb, j1 - j0, Kp, ubuf + j0, Nt, (__bf16 *)PU(p) + co, sm, ucp, lda,
                               ldg);
    return pack_cols<_Float16>(nb, j1 - j0, Kp, ubuf + j0, Nt, (_Float16 *)PU(p) + co, sm, ucp,
                               lda, ldg);
  };
  // rows [row0, row0 + rows) of step t's L21 (l21buf) packed into its K
  // segment (global rows; FP8 scales per global 128-row tile), with copy
  // also stored into the factor
  auto pack_l = [&](int t, int row0, int rows, bool copy) -> int {
    const int p = t / G, sg = t % G, r0 = t * nb, r1 = r0 + nb;
    if (rows <= 0) return CLU_OK;
    const float *l21 = (const float *)ws.l21buf + (size_t)row0 * nb;
    float *lcp = copy ? dA + (long)(r1 + row0) * lda + r0 : nullptr;
    const size_t ro = (size_t)(r1 + row0) * ldg + (size_t)sg * Kp;
    if (fp8) {
      int *escL = ESC(p, sg);
      float *partL = (float *)(escL + 2 * T);
      return quant_rows_fp8(rows, nb, Kp, -1.0f, l21, nb, (uint8_t *)PA(p) + ro,
                            escL + (r1 + row0) / 128,
                            partL + (size_t)FP8_PART_PER_TILE * ((r1 + row0) / 128), sm, lcp, lda,
                            ldg);
    }
    if (prec == CHARMLU_PREC_BF16)
      return pack_rows<__bf16>(rows, nb, Kp, -1.0f, l21, nb, (__bf16 *)PA(p) + ro, sm, lcp, lda,
                               ldg);
    return pack_rows<_Float16>(rows, nb, Kp, -1.0f, l21, nb, (_Float16 *)PA(p) + ro, sm, lcp, lda,
                               ldg);
  };
  // U row columns [j0, j1) of step t, packed and stored into the factor
  auto u_row2 = [&](int t, int j0, int j1) -> int {
    if (j1 <= j0) return CLU_OK;
    const int pid = prof_begin(PK_TRSM, (double)(j1 - j0) * nb * (nb - 1), sm);
    int e = urow_gemm(t, j0, j1);
    if (!e) e = pack_u(t, j0, j1, true);
    prof_end(pid, sm);
    return e;
  };
  // rows of a K = ns nb update of N columns that fill the duration of the
  // panel of step s on the update stream's CUs.  The rate at K: 0.8 of the
  // standalone kernel's, which at 64512^2 spends a fixed 6.9 (FP8) or 6.5
  // (16-bit) ms on C plus 3.9 or 7.3 ms per 1024 of K (measured inside the
  // LU: FP8 0.93-0.96 PFLOP/s at K = 2048 and 1.13-1.23 at 4096, BF16 0.64
  // at 2048)
  auto fill_rows = [&](int s, int ns, int M, int N) -> int {
    if (M <= 0) return 0;
    if (!split) return M;
    if (force_split) return forced_split(M);
    double p_ms = s < (int)ws.histmx.ms.size() ? ws.histmx.ms[s] : -1.0;
    if (p_ms <= 0.0) p_ms = nb / 1024.0 * 3.5;
    const double kk = (double)ns * nb / 1024.0;
    const double k_rate = 0.8 * 2.0 * 64512.0 * 64512.0 * 1024.0 * kk /
                          (fp8 ? 6.9 + 3.9 * kk : 6.5 + 7.3 * kk);  // flop per ms
    const double rate = k_rate * (ws.ncu - ws.panel_cus) / ws.ncu;
    const double rows = 1.1 * p_ms * rate / (2.0 * N * ns * nb);
    if (rows >= M) return M;
    const int r = ((int)rows + 127) / 128 * 128;
    return r > M ? M : r;
  };
  // Speculative identity pivots (the pivoted form on this schedule): each
  // panel is eliminated without interchanges, from a copy of its diagonal
  // block, and checked before anything uses it.  If every multiplier of the
  // elimination is below 1/2 in magnitude (and no pivot is tiny), partial
  // pivoting keeps every diagonal row: the diagonal candidate exceeds every
  // other one in its column at least twice over, and the pivoted schedule
  // computes the same candidates with other roundings of the same
  // operations, which move them by a small multiple of the update's unit
  // roundoff.  Otherwise the step falls back: the rest of the matrix gets
  // what the group's schedule still owes it, the diagonal block is restored,
  // and the exchanging panel and one update per step take over from that
  // step on.  The check costs one host round trip per step (the host waits
  // for its flag before it enqueues the step's U row).
  constexpr float SPEC_THR = 0.5f;
  const float tiny32 = 100.0f * 1.1920929e-07f;  // the panel's singular-pivot bound
  // copies of diagonal blocks, by step parity: before the block's
  // speculative panel (slot 0) and before its first update of the step
  // before (slot 2; a failed check of that step restores it)
  auto bak = [&](int slot, int s) {
    return (float *)ws.specbak + (size_t)(slot + (s & 1)) * nb * nb;
  };
  auto gpanel = [&](int s) -> int {
    if (spec) {
      const int r0 = s * nb;
      int e = convert<float, float>(nb, nb, dA + (long)r0 * lda + r0, lda, bak(0, s), nb, sp);
      if (e) return e;
    }
    return panel(s, true);
  };
  auto grouped_steps = [&]() -> int {
    int rc;
    if ((rc = gpanel(0))) return rc;
    CLU_HIP(hipEventRecord(ws.ev_panel, sp));
    // the rest of group yp's update (yg segments): rows and columns [y0, n),
    // ydone rows of it enqueued; y_full: its last rows run on the all-CU
    // stream (ev_full)
    int yp = -1, yg = 0, y0 = n, ydone = 0;
    bool y_full = false;
    auto y_part = [&](int rows, hipStream_t st) -> int {
      const int M = n - y0 - ydone;
      rows = rows < M ? rows : M;
      if (rows <= 0) return CLU_OK;
      const int e = updg(yp, yg, y0 + ydone, y0, rows, n - y0, st);
      ydone += rows;
      return e;
    };
    // the panel of step s, after what ev_la marks (the work that runs beside
    // it is enqueued before its launches, so that the host's enqueueing of
    // the panel's many small launches does not hold it back)
    auto panel_step = [&](int s) -> int {
      CLU_HIP(hipStreamWaitEvent(sp, ws.ev_la, 0));
      CLU_HIP(hipEventRecord(ws.histmx.ev[2 * s], sp));
      const int e = gpanel(s);
      if (e) return e;
      CLU_HIP(hipEventRecord(ws.histmx.ev[2 * s + 1], sp));
      CLU_HIP(hipEventRecord(ws.ev_panel, sp));
      return CLU_OK;
    };
    // the fall back at step s (index i of group p)
    auto fallback = [&](int s, int i, int p, bool next_ran) -> int {
      int e;
      const int r0 = s * nb, r1 = r0 + nb, Nt = n - r1;
      if (next_ran) {
        // panel(s + 1) eliminated its diagonal block after that block's
        // update by step s: back to the block before that update
        CLU_HIP(hipStreamWaitEvent(sm, ws.ev_panel, 0));
        float *A22 = dA + (long)r1 * lda + r1;
        if ((e = convert<float, float>(nb, nb, bak(2, s + 1), nb, A22, lda, sm))) return e;
      }
      if (y_full) {
        CLU_HIP(hipStreamWaitEvent(sm, ws.ev_full, 0));
        y_full = false;
      }
      if (yp >= 0 && ydone < n - y0 && (e = y_part(n - y0 - ydone, sm))) return e;
      if (i > 0 && Nt > 0 && (e = updg(p, i, r1, r1, Nt, Nt, sm))) return e;
      float *A11 = dA + (long)r0 * lda + r0;
      if ((e = convert<float, float>(nb, nb, bak(0, s), nb, A11, lda, sm))) return e;
      if ((e = fill_int(dinfo, 0, 1, sm))) return e;  // a tiny pivot of the elimination
      CLU_HIP(hipEventRecord(ws.ev_la, sm));
      CLU_HIP(hipStreamWaitEvent(sp, ws.ev_la, 0));
      if ((e = panel(s))) return e;
      CLU_HIP(hipEventRecord(ws.ev_panel, sp));
      full_pending = false;
      return per_step(s);
    };
    // the check of step s's elimination (L21 in l21buf unless it is the
    // last), enqueued; then its verdict (1 keep, 0 fall back, -1 error)
    auto spec_enqueue = [&](int s, int Nt) -> int {
      const int r0 = s * nb;
      int e = spec_check(Nt, nb, (const float *)ws.l21buf, nb, dA + (long)r0 * lda + r0, lda,
                         SPEC_THR, tiny32, ws.spec_flag, sm);
      if (e) return e;
      CLU_HIP(hipMemcpyAsync(ws.spec_hflag, ws.spec_flag, sizeof(int), hipMemcpyDeviceToHost, sm));
      CLU_HIP(hipEventRecord(ws.ev_chk, sm));
      return CLU_OK;
    };
    auto spec_verdict = [&]() -> int {
      if (hipEventSynchronize(ws.ev_chk) != hipSuccess) return -1;
      return *ws.spec_hflag == 0 ? 1 : 0;
    };
    auto spec_ok = [&](int s, int Nt) -> int {
      if (spec_enqueue(s, Nt)) return -1;
      return spec_verdict();
    };
    for (int a = 0; a < NB; a += G) {
      const int p = a / G, gl = NB - a < G ? NB - a : G;
      for (int i = 0; i < gl; ++i) {
        const int s = a + i, r1 = (s + 1) * nb, Nt = n - r1;
        CLU_HIP(hipStreamWaitEvent(sm, ws.ev_panel, 0));  // panel(s)
        if (Nt <= 0) {
          if (spec) {
            const int ok = spec_ok(s, 0);
            if (ok < 0) {
              set_error("the check of a speculative panel failed to run");
              return CLU_ERR_HIP;
            }
            if (!ok) return fallback(s, i, p, false);
          }
          break;
        }
        // the previous group's remaining rows on every CU once the group's
        // last panel is done (no panel runs until the next update waits for
        // them), beside that step's L21 and U row
        if (i == gl - 1 && yp >= 0 && ydone < n - y0) {
          hipStream_t sf = ws.s_full;
          CLU_HIP(hipStreamWaitEvent(sf, ws.ev_panel, 0));
          if ((rc = y_part(n - y0 - ydone, sf))) return rc;
          CLU_HIP(hipEventRecord(ws.ev_full, sf));
          y_full = true;
        }
        const bool last = i == gl - 1;
        const int ns = last ? gl : i + 1;  // K segments of block column s + 1's update
        // Between panel(s) and panel(s + 1) only what the next panel needs:
        // the first block row of L21, the U row of block column s + 1 (both
        // packed, their copies into the factor after the check) and the
        // diagonal block of s + 1 brought up to date (its copy first, for a
        // failed check)
        int pid = prof_begin(PK_TRSM, (double)nb * nb * (nb - 1), sm);
        if ((rc = l21_rows(s, 0, nb))) return rc;
        if ((rc = urow_gemm(s, 0, nb))) return rc;
        if ((rc = pack_u(s, 0, nb, false))) return rc;
        if ((rc = pack_l(s, 0, nb, false))) return rc;
        prof_end(pid, sm);
        if (last && y_full) {  // the previous group's rest holds the next block
          CLU_HIP(hipStreamWaitEvent(sm, ws.ev_full, 0));
          y_full = false;
        }
        if (spec && (rc = convert<float, float>(nb, nb, dA + (long)r1 * lda + r1, lda,
                                                bak(2, s + 1), nb, sm)))
          return rc;
        if ((rc = updg(p, ns, r1, r1, nb, nb, sm))) return rc;
        CLU_HIP(hipEventRecord(ws.ev_la, sm));
        // beside panel(s + 1): the rest of L21, then (speculative) the check
        if ((rc = l21_rows(s, nb, Nt - nb))) return rc;
        bool panel_done = false;
        if (spec) {
          if ((rc = spec_enqueue(s, Nt))) return rc;
          if (s + 1 < NB && (rc = panel_step(s + 1))) return rc;
          panel_done = true;
          const int ok = spec_verdict();
          if (ok < 0) {
            set_error("the check of a speculative panel failed to run");
            return CLU_ERR_HIP;
          }
          if (!ok) return fallback(s, i, p, s + 1 < NB);
        }
        // the step's factor copies, the rest of its L21 and of block column
        // s + 1, the rest of its U row
        pid = prof_begin(PK_TRSM, (double)(Nt - nb) * nb * nb, sm);
        if ((rc = convert<float, float>(nb, nb, (const float *)ws.l21buf, nb,
                                        dA + (long)r1 * lda + s * nb, lda, sm)))
          return rc;
        if ((rc = convert<float, float>(nb, nb, ubuf, Nt, dA + (long)(s * nb) * lda + r1, lda,
                                        sm)))
          return rc;
        if ((rc = pack_l(s, nb, Nt - nb, true))) return rc;
        prof_end(pid, sm);
        if ((rc = updg(p, ns, r1 + nb, r1, Nt - nb, nb, sm))) return rc;
        if ((rc = u_row2(s, nb, Nt))) return rc;
        if (!last) {
          // block row s + 1 with the group's steps a .. s, then rows of the
          // previous group's rest
          if ((rc = updg(p, i + 1, r1, r1 + nb, nb, Nt - nb, sm))) return rc;
          if (yp >= 0 && (rc = y_part(fill_rows(s + 1, yg, n - y0 - ydone, n - y0), sm)))
            return rc;
        } else {
          // the group's update of the rest (gl segments): the next group's
          // other block columns and its block rows, then the first rows of
          // the rest
          const int W = Nt < G * nb ? Nt : G * nb;
          if ((rc = updg(p, gl, r1, r1 + nb, Nt, W - nb, sm))) return rc;
          if ((rc = updg(p, gl, r1, r1 + W, W, Nt - W, sm))) return rc;
          yp = p;
          yg = gl;
          y0 = r1 + W;
          ydone = 0;
          if (y0 < n && (rc = y_part(fill_rows(s + 1, yg, n - y0, n - y0), sm))) return rc;
        }
        if (!panel_done && s + 1 < NB && (rc = panel_step(s + 1))) return rc;
      }
    }
    if (y_full) CLU_HIP(hipStreamWaitEvent(sm, ws.ev_full, 0));
    if (yp >= 0 && ydone < n - y0 && (rc = y_part(n - y0 - ydone, sm))) return rc;
    if ((rc = finish_info(dinfo, ws.abortw, sm))) return rc;
    CLU_HIP(hipEventRecord(ws.ev_end, sm));
    CLU_HIP(hipStreamWaitEvent(su, ws.ev_end, 0));
    return CLU_OK;
  };

  rc = [&]() -> int {  // the steps; an error returns from here to the fence below
  int rc;
  if (grouped) return grouped_steps();
  if ((rc = panel(0))) return rc;
  CLU_HIP(hipEventRecord(ws.ev_panel, sp));
  return per_step(0);
  }();
  if (rc) fence_streams(ws, su);
  return rc;
}

// FP64 -> FP32 copy of a matrix (the working matrix of the mixed path)
int charmlu_dlag2s(long m, long n, const double *dA, long lda, float *dS, long lds,
                   void *stream) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (!dA && m > 0 && n > 0) return -3;
  if (lda < (n > 1 ? n : 1)) return -4;
  if (!dS && m > 0 && n > 0) return -5;
  if (lds < (n > 1 ? n : 1)) return -6;
  return convert<float, double>(m, n, dA, lda, dS, lds, (hipStream_t)stream);
}

}  // extern "C"

// ---- FP64 GMRES-based iterative refinement ------------------------------------
// x0 = op(LU)^{-1} b in FP32; then, while the reference's scaled residual
// max|b - op(A) x| / ((max|a_ij| max|x| + max|b|) n eps) (benchmark.C:385-400)
// exceeds tol: r = b - op(A) x in FP64, one GMRES cycle on
// op(LU)^{-1} op(A) d = op(LU)^{-1} r (FP64 Arnoldi basis, classical
// Gram-Schmidt applied twice, Givens rotations on the host), x += d.  op is
// the identity or the transpose (the adjoint solve A^T lambda = g of the
// differentiable solve, from the same stored factors).  The FP32 factors
// only apply the preconditioner; residuals, the basis, inner products and x
// stay in FP64.  (Starting from x0 = 0 instead saves one solve and one
// residual pass but measured one more cycle at n = 65536.)
// b and x: contiguous n-vectors; stats[4] as charmlu_dgesv_ir.
namespace clu {
static int gmres_ir(bool trans, int n, const double *dA, long lda, const float *dLU, long ldlu,
                    const int *dipiv, const double *db, double *dx, double tol, int max_iter,
                    int restart, double *stats, double *irbuf, bool &inv_ready, hipStream_t s) {
  const int m = restart;
  // irbuf: [V: (m+1) x n][r: n][w: n][small: 2*(m+1)+16][z32: n floats]
  const size_t small = 2 * (size_t)(m + 1) + 16;
  const size_t dbl = (size_t)(m + 3) * n + small;
  double *V = irbuf;
  double *r = V + (size_t)(m + 1) * n;
  double *w = r + n;
  double *dh = w + n;                // h (m+1), then h2 (m+1)
  double *dnorm = dh + 2 * (m + 1);  // 2 norms (then 2 more for the transposed check)
  double *dres = dnorm + 4;          // residual maxima (5), column maxima (5)
  float *z32 = (float *)(V + dbl);
  const double eps = 0x1.0p-53;
  const int tr = trans ? CHARMLU_TRANS : CHARMLU_NO_TRANS;
  int rc;
  auto precond = [&](const double *src, double scale, double *dst) -> int {
    // dst = op(LU)^{-1} src, with src scaled by 1/scale into FP32 range
    int e = vec_d2s(n, 1.0 / scale, src, z32, s);
    if (e) return e;
    e = getrs_impl<float>(tr, n, 1, dLU, ldlu, dipiv, z32, 1, s, inv_ready,
                          n >= 8192 ? 1024 : 256);
    if (e) return e;
    inv_ready = true;
    return vec_s2d(n, scale, z32, dst, s);
  };
  // op(A)^T's fixed maxima for the transposed check: max|a_ij|, ||A^T||_inf, max|b|
  double amax = 0.0, anorm_t = 0.0, bmax = 0.0;
  if (trans) {
    double h3[3];
    if ((rc = abs_colsum(n, dA, lda, dres + 5, s))) return rc;
    if ((rc = vec_norms(n, db, dnorm + 2, s))) return rc;
    CLU_HIP(hipMemcpyAsync(h3, dres + 5, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    CLU_HIP(hipMemcpyAsync(h3 + 2, dnorm + 3, sizeof(double), hipMemcpyDeviceToHost, s));
    CLU_HIP(hipStreamSynchronize(s));
    amax = h3[0];
    anorm_t = h3[1];
    bmax = h3[2];
  }
  double hres[5];
  // r = b - op(A) x and the two scaled residuals
  auto check = [&](double *ref, double *ns, double *rmax) -> int {
    double xmax;
    if (!trans) {
      int e = residual(n, dA, lda, dx, db, r, dres, s);
      if (e) return e;
      CLU_HIP(hipMemcpyAsync(hres, dres, 5 * sizeof(double), hipMemcpyDeviceToHost, s));
      CLU_HIP(hipStreamSynchronize(s));
      amax = hres[0];
      bmax = hres[1];
      xmax = hres[2];
      *rmax = hres[3];
      anorm_t = hres[4];
    } else {
      int e = vec_scale(n, 1.0, db, r, s);
      if (!e) e = dgemv(true, n, n, -1.0, dA, lda, dx, 1, 1.0, r, 1, s);
      if (!e) e = vec_norms(n, r, dnorm, s);
      if (!e) e = vec_norms(n, dx, dnorm + 2, s);
      if (e) return e;
      CLU_HIP(hipMemcpyAsync(hres, dnorm, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
      CLU_HIP(hipStreamSynchronize(s));
      *rmax = hres[1];
      xmax = hres[3];
    }
    const double den = (amax * xmax + bmax) * n * eps;
    *ref = den > 0 ? *rmax / den : (*rmax > 0 ? INFINITY : 0.0);
    const double den2 = anorm_t * xmax * n * eps;
    *ns = den2 > 0 ? *rmax / den2 : (*rmax > 0 ? INFINITY : 0.0);
    return CLU_OK;
  };
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), hh(2 * (m + 1) + 2), y(m);
  int outer = 0, inner = 0;
  double ref = 0, ns = 0, rmax = 0;
  // x0 = op(LU)^{-1} b
  if ((rc = precond(db, 1.0, dx))) return rc;
  if ((rc = check(&ref, &ns, &rmax))) return rc;
  for (;;) {
    if (ref <= tol || outer >= max_iter || !(rmax > 0)) break;
    ++outer;
    // v0 = op(LU)^{-1} r / beta
    if ((rc = precond(r, rmax, V))) return rc;
    if ((rc = vec_norms(n, V, dnorm, s))) return rc;
    CLU_HIP(hipMemcpyAsync(hh.data(), dnorm, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    CLU_HIP(hipStreamSynchronize(s));
    const double beta = sqrt(hh[0]);
    if (!(beta > 0) || !std::isfinite(beta)) {
      set_error("refinement broke down: preconditioned residual norm %g", beta);
      return CLU_ERR_HIP;
    }
    if ((rc = vec_scale(n, 1.0 / beta, V, V, s))) return rc;
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    for (int j = 0; j < m; ++j) {
      double *vj = V + (size_t)j * n, *vn = V + (size_t)(j + 1) * n;
      // w = op(A) v_j in FP64, then v_{j+1} = op(LU)^{-1} w
      if ((rc = dgemv(trans, n, n, 1.0, dA, lda, vj, 1, 0.0, w, 1, s))) return rc;
      if ((rc = precond(w, 1.0, vn))) return rc;
      // classical Gram-Schmidt, twice
      if ((rc = basis_dot(j + 1, n, V, n, vn, dh, s))) return rc;
      if ((rc = basis_axpy(j + 1, n, -1.0, V, n, dh, vn, s))) return rc;
      if ((rc = basis_dot(j + 1, n, V, n, vn, dh + (m + 1), s))) return rc;
      if ((rc = basis_axpy(j + 1, n, -1.0, V, n, dh + (m + 1), vn, s))) return rc;
      if ((rc = vec_norms(n, vn, dnorm, s))) return rc;
      CLU_HIP(hipMemcpyAsync(hh.data(), dh, (2 * (m + 1) + 2) * sizeof(double),
                             hipMemcpyDeviceToHost, s));
      CLU_HIP(hipStreamSynchronize(s));
      ++inner;
      // column j of the Hessenberg matrix (stored by rows, H[i*m + j])
      for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = hh[i] + hh[(m + 1) + i];
      const double hn = sqrt(hh[2 * (m + 1)]);
      H[(size_t)(j + 1) * m + j] = hn;
      // earlier rotations, then the one that zeroes H[j+1][j]
      for (int i = 0; i < j; ++i) {
        const double a = H[(size_t)i * m + j], b2 = H[(size_t)(i + 1) * m + j];
        H[(size_t)i * m + j] = cs[i] * a + sn[i] * b2;
        H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * b2;
      }
      const double a = H[(size_t)j * m + j], b2 = H[(size_t)(j + 1) * m + j];
      const double d = hypot(a, b2);
      cs[j] = d > 0 ? a / d : 1.0;
      sn[j] = d > 0 ? b2 / d : 0.0;
      H[(size_t)j * m + j] = d;
      H[(size_t)(j + 1) * m + j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      k = j + 1;
      // stop the cycle once the preconditioned residual is at FP64 level
      if (!(hn > 0) || fabs(g[j + 1]) <= 1e-15 * beta) break;
      if ((rc = vec_scale(n, 1.0 / hn, vn, vn, s))) return rc;
    }
    // y = H(1:k,1:k)^{-1} g(1:k); x += V(:,1:k) y
    for (int i = k - 1; i >= 0; --i) {
      double t = g[i];
      for (int c = i + 1; c < k; ++c) t -= H[(size_t)i * m + c] * y[c];
      const double dgi = H[(size_t)i * m + i];
      y[i] = dgi != 0 ? t / dgi : 0.0;
    }
    // the correction coefficients go through the h buffer (free again here)
    CLU_HIP(hipMemcpyAsync(dh, y.data(), k * sizeof(double), hipMemcpyHostToDevice, s));
    if ((rc = basis_axpy(k, n, 1.0, V, n, dh, dx, s))) return rc;
    if ((rc = check(&ref, &ns, &rmax))) return rc;
  }
  {
    // a one-launch block solve that gave up a wait left NaN in its rows and
    // its sequence number in the abort word (check() has synchronized s)
    Workspace *wsp;
    if ((rc = get_ws(&wsp))) return rc;
    unsigned ab = 0;
    CLU_HIP(hipMemcpy(&ab, (unsigned *)wsp->abortw + 8, sizeof(ab), hipMemcpyDeviceToHost));
    if (ab != 0) {
      CLU_HIP(hipMemset((unsigned *)wsp->abortw + 8, 0, sizeof(ab)));
      set_error("block solve: a wait for published values gave up (call %u)", ab);
      return CLU_ERR_HIP;
    }
  }
  if (stats) {
    stats[0] = outer;
    stats[1] = inner;
    stats[2] = ref;
    stats[3] = ns;
  }
  return CLU_OK;
}

// bytes of the refinement's buffer (plus two n-vectors for a column of a
// multi-RHS call)
static size_t irbuf_bytes(int n, int restart) {
  const size_t small = 2 * (size_t)(restart + 1) + 16;
  return ((size_t)(restart + 3) * n + small + 2 * (size_t)n) * sizeof(double) +
         sizeof(float) * n + 256;
}
}  // namespace clu

extern "C" {

int charmlu_dgesv_ir(int n, const double *dA, long lda, const float *dLU, long ldlu,
                     const int *dipiv, const double *db, double *dx, double tol, int max_iter,
                     int restart, double *stats, void *stream) {
  if (n < 0) return -1;
  if (!dA && n > 0) return -2;
  if (lda < (n > 1 ? n : 1)) return -3;
  if (!dLU && n > 0) return -4;
  if (ldlu < (n > 1 ? n : 1)) return -5;
  if (!dipiv && n > 0) return -6;
  if (!db && n > 0) return -7;
  if (!dx && n > 0) return -8;
  if (!(tol > 0.0)) return -9;
  if (max_iter < 0) return -10;
  if (restart < 1 || restart > 200) return -11;
  hipStream_t s = (hipStream_t)stream;
  if (stats) for (int i = 0; i < 4; ++i) stats[i] = 0.0;
  if (n == 0) return CLU_OK;
  Workspace *wsp;
  int rc;
  if ((rc = get_ws(&wsp))) return rc;
  Workspace &ws = *wsp;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if ((rc = ensure_bytes(&ws.irbuf, &ws.irbuf_b, irbuf_bytes(n, restart)))) return rc;
  }
  bool inv_ready = false;
  return gmres_ir(false, n, dA, lda, dLU, ldlu, dipiv, db, dx, tol, max_iter, restart, stats,
                  (double *)ws.irbuf, inv_ready, s);
}

int charmlu_dgesv_ir_ex(int trans, int n, int nrhs, const double *dA, long lda, const float *dLU,
                        long ldlu, const int *dipiv, const double *dB, long ldb, double *dX,
                        long ldx, double tol, int max_iter, int restart, double *stats,
                        void *stream) {
  if (trans != CHARMLU_NO_TRANS && trans != CHARMLU_TRANS) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (!dA && n > 0) return -4;
  if (lda < (n > 1 ? n : 1)) return -5;
  if (!dLU && n > 0) return -6;
  if (ldlu < (n > 1 ? n : 1)) return -7;
  if (!dipiv && n > 0) return -8;
  if (!dB && n > 0 && nrhs > 0) return -9;
  if (ldb < (nrhs > 1 ? nrhs : 1)) return -10;
  if (!dX && n > 0 && nrhs > 0) return -11;
  if (ldx < (nrhs > 1 ? nrhs : 1)) return -12;
  if (!(tol > 0.0)) return -13;
  if (max_iter < 0) return -14;
  if (restart < 1 || restart > 200) return -15;
  hipStream_t s = (hipStream_t)stream;
  if (stats) for (int i = 0; i < 4; ++i) stats[i] = 0.0;
  if (n == 0 || nrhs == 0) return CLU_OK;
  Workspace *wsp;
  int rc;
  if ((rc = get_ws(&wsp))) return rc;
  Workspace &ws = *wsp;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if ((rc = ensure_bytes(&ws.irbuf, &ws.irbuf_b, irbuf_bytes(n, restart)))) return rc;
  }
  double *buf = (double *)ws.irbuf;
  // a column's right-hand side and solution, contiguous, after the
  // refinement's own vectors
  const size_t small = 2 * (size_t)(restart + 1) + 16;
  double *bcol = buf + (size_t)(restart + 3) * n + small + (size_t)(n + 1) / 2 + 1;
  double *xcol = bcol + n;
  // the single-vector layout puts n floats right after the doubles: keep the
  // columns past them (the float area rounded up to whole doubles, plus one)
  bool inv_ready = false;
  const bool tr = trans == CHARMLU_TRANS;
  for (int k = 0; k < nrhs; ++k) {
    double st[4] = {0, 0, 0, 0};
    if ((rc = convert<double, double>(n, 1, dB + k, ldb, bcol, 1, s))) return rc;
    if ((rc = gmres_ir(tr, n, dA, lda, dLU, ldlu, dipiv, bcol, xcol, tol, max_iter, restart, st,
                       buf, inv_ready, s)))
      return rc;
    if ((rc = convert<double, double>(n, 1, xcol, 1, dX + k, ldx, s))) return rc;
    if (stats) {
      stats[0] = std::max(stats[0], st[0]);
      stats[1] += st[1];
      stats[2] = std::max(stats[2], st[2]);
      stats[3] = std::max(stats[3], st[3]);
    }
  }
  return CLU_OK;
}

}  // extern "C"

extern "C" {

// ---- generation and validation ---------------------------------------------------
int charmlu_dgen_matrix(int kind, uint64_t seed, long n, double *dA, long lda, void *stream) {
  if (kind != 0 && kind != 1) return -1;
  if (n < 0) return -3;
  if (!dA && n > 0) return -4;
  if (lda < (n > 1 ? n : 1)) return -5;
  if (kind == 0) return gen_uniform(seed, 0, 0, n, n, dA, lda, (hipStream_t)stream);
  return gen_diagdom(seed, n, dA, lda, (hipStream_t)stream);
}

int charmlu_dgen_rhs(int seed, long skip, long n, double *db, void *stream) {
  if (skip < 0) return -2;
  if (n < 0) return -3;
  if (!db && n > 0) return -4;
  return gen_rhs(seed, skip, n, db, (hipStream_t)stream);
}

int charmlu_dresidual(int n, const double *dA, long lda, const double *dx, const double *db,
                      double *dr, double *dout, void *stream) {
  if (n < 0) return -1;
  if (!dA && n > 0) return -2;
  if (lda < (n > 1 ? n : 1)) return -3;
  if (!dout) return -7;
  return residual(n, dA, lda, dx, db, dr, dout, (hipStream_t)stream);
}

}  // extern "C"
