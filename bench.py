"""Benchmark: effective TFLOP/s of the device LU (factor + solve) at residual < 16.

Default workload: plain FP64 blocked LU with partial pivoting plus the
solve, n = 32768, block 256, A and b uniform [-0.5, 0.5) from the seeds of
the reference driver (benchmark.C:436-442), one right-hand side -- the
single-GPU FP64 configuration.  A step is one factor + solve of a fresh copy
of A (the copy restores the input in HBM; it is inside the timed region).
Effective TFLOP/s = (2/3 n^3) / (seconds per step) (the reference reports the
HPL form (2/3 n^3 + 3/2 n^2)/t, benchmark.C:344-349; both are printed).
Validation (the reference's scaled residual, benchmark.C:385-400) runs after
the timed region, as in the reference.

The JSON line also carries:
  roofline      the trailing-update GEMM (dgemm_nn_lds_kernel, 128x128 tiles):
                algorithmic flops per launch / average launch duration, both
                from HIP events the library records around each launch on
                the stream it runs on (charmlu_profile), against the FP64
                matrix peak of 78.6 TFLOP/s; beside it the same flops over
                the union of the launch intervals (the two launches of a
                step overlap).
  cpu_baseline  the CPU oracle (a restatement of the reference's block
                algorithm with the vendor BLAS per block) timed on this host
                on a bounded sample of the same workload (its first block
                steps), in the same unit, with C1 in full and the vendor
                LAPACK's LU beside it; core count and CPU model included.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--nb B]
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "effective TFLOP/s (2/3n^3) at residual<16: FP64 LU vs mixed-prec LU+FP64 IR"
FP64_PEAK = 78.6  # TFLOP/s, MI355X FP64 matrix (dense)
HBM_PEAK = 8000.0  # GB/s


def host_cpu():
    """(threads to use, nproc, CPU model): the threads this process may run
    on, capped by OMP_NUM_THREADS where the host sets it (a GPU host grants
    each GPU a share of its cores)."""
    nproc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = nproc
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(avail, int(omp)) if omp.isdigit() and int(omp) > 0 else avail
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, nproc, model


# The restatement's run on the host: C1 in full, and the first SAMPLE_STEPS
# block steps of the single-GPU configuration (the most expensive ones) as
# its bounded sample; then, in a second process (the vendor library's
# threading is chosen once per process), the vendor LAPACK's multithreaded
# LU + solve at C1 and at n = 16384.
_PORT_CODE = """
import json, sys, oracle
T, n, nb, steps = (int(a) for a in sys.argv[1:5])
mkl = oracle.use_mkl()
c1 = oracle.lu_solve_problem(4096, 128, 0, nthreads=T)
A = oracle.gen_uniform(oracle.SEED_A, n)
oracle.set_max_steps(steps)
t0 = oracle.wtime()
oracle.getrf(A, nb, 0, None, T)
sec = oracle.wtime() - t0
print(json.dumps(dict(mkl=mkl, c1_seconds=c1['seconds'], c1_residual=c1['residual'],
                      c1_info=c1['info'], sample_seconds=sec)))
"""
_LAPACK_CODE = """
import json, sys, oracle
T = int(sys.argv[1])
out = {}
for n in (4096, 16384):
    A = oracle.gen_uniform(oracle.SEED_A, n)
    b = oracle.gen_rhs(n)
    info, x, sec = oracle.lapack_gesv(A, b, T)
    res = oracle.residual(A, x, b)[0] if n <= 4096 else None
    out[str(n)] = dict(info=info, seconds=sec, residual=res)
print(json.dumps(out))
"""
SAMPLE_STEPS = 8


def _child(code, args, timeout):
    out = subprocess.run([sys.executable, "-c", code, *map(str, args)], cwd=ROOT,
                         capture_output=True, text=True, timeout=timeout)
    if out.returncode != 0:
        raise RuntimeError(out.stderr[-500:])
    return json.loads(out.stdout.strip().splitlines()[-1])


def cpu_baseline(n: int, nb: int, timeout: float = 300.0):
    """The CPU baselines on this host's cores (reported beside the GPU
    figure, not its target): the oracle (a restatement of the reference's
    block algorithm with the vendor BLAS per block, the work of a step
    spread over the threads as the reference spreads it over one PE per
    core) at C1 (n = 4096, block 128) and on the first block steps of the
    n x n factorization, plus the vendor LAPACK's LU; returns the
    cpu_baseline object."""
    threads, nproc, model = host_cpu()
    base = {"unit": "TFLOP/s", "cores": threads, "kind": "port", "nproc": nproc,
            "cpu_model": model}
    try:
        rec = _child(_PORT_CODE, [threads, n, nb, SAMPLE_STEPS], timeout)
    except Exception as e:  # the baseline is reported, never required
        return dict(base, value=None, sample=f"failed: {e}")
    flops = 2.0 / 3.0 * (n ** 3 - (n - SAMPLE_STEPS * nb) ** 3)
    tf = flops / rec["sample_seconds"] / 1e12
    blas = "MKL sequential dgemm/dtrsm/dger per block" if rec["mkl"] else "plain C loops"
    out = dict(base, value=round(tf, 5), sample=(
        f"oracle restatement of the reference block LU ({blas}, {threads} threads over the "
        f"blocks of a step): the first {SAMPLE_STEPS} of {n // nb} block steps of n={n}, b={nb} "
        f"(2/3 n^3 - 2/3 (n - {SAMPLE_STEPS}*{nb})^3 = {flops:.3e} flop) in "
        f"{rec['sample_seconds']:.2f} s"))
    c1_tf = (2.0 / 3.0 * 4096 ** 3) / rec["c1_seconds"] / 1e12
    out["c1"] = {"n": 4096, "block": 128, "seconds": round(rec["c1_seconds"], 3),
                 "tflops": round(c1_tf, 5), "residual": rec["c1_residual"],
                 "info": rec["c1_info"]}
    try:
        lap = _child(_LAPACK_CODE, [threads], timeout)
        out["library"] = {"kind": "vendor LAPACK dgetrf + dgetrs (MKL, multithreaded)",
                          "cores": threads}
        for k, v in lap.items():
            m = int(k)
            out["library"][f"n{m}"] = {"seconds": round(v["seconds"], 3),
                                       "tflops": round(2.0 / 3.0 * m ** 3 / v["seconds"] / 1e12, 5),
                                       "info": v["info"], "residual": v["residual"]}
    except Exception as e:
        out["library"] = {"failed": str(e)}
    return out


def block_dgemm(lib, _lib, torch, nb, lu_tflops, trials=10):
    """The reference's single-block check (testdgemm and "Percent of DGEMM",
    benchmark.C:297-341,374-381): C -= A B for one b x b block, the mean of
    10 timed calls after one untimed.  On the device through
    charmlu_dgemm_dev (operands resident, each call timed to its
    completion, as the reference times one call), and through the
    host-pointer CBLAS shim charmlu_cblas_dgemm (the platformBlas.h
    drop-in: its copies to and from the device included).  percent_of_dgemm
    is the reference's ratio: the LU's rate over the device block rate (a
    single b x b product is far too small to fill the GPU, so the ratio
    exceeds 100 %; the trailing update's rate is the roofline object)."""
    import ctypes
    import numpy as np
    g = torch.Generator(device="cpu").manual_seed(0)
    Ah, Bh, Ch = ((torch.rand(nb, nb, generator=g, dtype=torch.float64) - 0.5) for _ in range(3))
    A, B, C = Ah.cuda(), Bh.cuda(), Ch.cuda()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    flops = 2.0 * nb ** 3

    def dev_call():
        _lib.check(lib.charmlu_dgemm_dev(_lib.NO_TRANS, _lib.NO_TRANS, nb, nb, nb, -1.0, p(A), nb,
                                         p(B), nb, 1.0, p(C), nb, st), "charmlu_dgemm_dev")
        torch.cuda.synchronize()

    dev_call()
    t = []
    for _ in range(trials):
        t0 = time.perf_counter()
        dev_call()
        t.append(time.perf_counter() - t0)
    dev_s = sum(t) / trials
    An, Bn, Cn = (np.ascontiguousarray(x.numpy()) for x in (Ah, Bh, Ch))
    dp = lambda a: a.ctypes.data_as(ctypes.c_void_p)

    def host_call():
        rc = lib.charmlu_cblas_dgemm(_lib.ROW_MAJOR, _lib.NO_TRANS, _lib.NO_TRANS, nb, nb, nb,
                                     -1.0, dp(An), nb, dp(Bn), nb, 1.0, dp(Cn), nb)
        if rc != 0:
            raise RuntimeError(f"charmlu_cblas_dgemm returned {rc}")

    host_call()
    t = []
    for _ in range(trials):
        t0 = time.perf_counter()
        host_call()
        t.append(time.perf_counter() - t0)
    host_s = sum(t) / trials
    dev_tf = flops / dev_s / 1e12
    return {"b": nb, "device_ms": round(dev_s * 1e3, 4), "device_tflops": round(dev_tf, 3),
            "host_shim_ms": round(host_s * 1e3, 4),
            "host_shim_tflops": round(flops / host_s / 1e12, 4),
            "percent_of_dgemm": round(lu_tflops / dev_tf * 100.0, 1),
            "note": "one b x b x b dgemm (C -= A B) per call, mean of 10 calls"}


def gemm_only(lib, _lib, torch, n, nb, steps, warmup):
    """The step-0 trailing update C -= L U (M = N = n - nb, K = nb) on the
    factorization's own row-major layout, timed with events on the stream."""
    import ctypes
    m = n - nb
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    A.uniform_(-0.5, 0.5)
    base = A.data_ptr()
    L = ctypes.c_void_p(base + (nb * n) * 8)
    U = ctypes.c_void_p(base + nb * 8)
    C = ctypes.c_void_p(base + (nb * n + nb) * 8)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def run():
        _lib.check(lib.charmlu_dgemm_dev(_lib.NO_TRANS, _lib.NO_TRANS, m, m, nb, -1.0, L, n, U, n,
                                         1.0, C, n, stream), "charmlu_dgemm_dev")
    for _ in range(warmup):
        run()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        run()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    flops = 2.0 * m * m * nb
    print(json.dumps({"gemm": f"{m}x{m}x{nb}", "ms": round(ms, 4),
                      "tflops": round(flops / ms / 1e9, 3),
                      "frac_fp64_peak": round(flops / ms / 1e9 / FP64_PEAK, 4),
                      "algorithmic_bytes": 8 * (2 * m * m + 2 * m * nb)}), flush=True)


def lp_gemm_only(lib, _lib, torch, n, nb, steps, warmup, prec="bf16"):
    """The step-0 low-precision trailing update C += A B^T (M = N = n - nb,
    K = nb; FP32 C, BF16/FP16 operands), timed with events on the stream."""
    import ctypes
    m = n - nb
    fp8 = prec == "fp8"
    dt = torch.float32 if fp8 else torch.bfloat16 if prec == "bf16" else torch.float16
    A = (torch.rand((m, nb), device="cuda") - 0.5).to(dt)
    B = (torch.rand((nb, m) if fp8 else (m, nb), device="cuda") - 0.5).to(dt)
    C = torch.rand((m, m), device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = _lib.PREC_BF16 if prec == "bf16" else _lib.PREC_FP16
    res = {}
    for alpha, tag in ((1.0, "accumulate"), (-1.0, "alpha_beta")):
        def run():
            if fp8:  # quantizes both operands (e4m3, per-tile scales), then the product
                _lib.check(lib.charmlu_gemm_fp8_dev(m, m, nb, alpha, ctypes.c_void_p(A.data_ptr()),
                                                    nb, ctypes.c_void_p(B.data_ptr()), m, 1.0,
                                                    ctypes.c_void_p(C.data_ptr()), m, stream),
                           "charmlu_gemm_fp8_dev")
                return
            _lib.check(lib.charmlu_gemm_lp_dev(p, m, m, nb, alpha, ctypes.c_void_p(A.data_ptr()),
                                               nb, ctypes.c_void_p(B.data_ptr()), nb, 1.0,
                                               ctypes.c_void_p(C.data_ptr()), m, stream),
                       "charmlu_gemm_lp_dev")
        for _ in range(warmup):
            run()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(steps):
            run()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        flops = 2.0 * m * m * nb
        res[tag] = {"ms": round(ms, 4), "tflops": round(flops / ms / 1e9, 1),
                    "c_gbs": round(8.0 * m * m / ms / 1e6, 1)}
    # reference point for the C stream: an elementwise read-modify-write of
    # the same FP32 matrix (8 bytes per element)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    C.add_(1.0)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        C.add_(1.0)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    res["rmw_reference_gbs"] = round(8.0 * m * m / ms / 1e6, 1)
    print(json.dumps({"lp_gemm": f"{m}x{m}x{nb} {prec}", **res}), flush=True)


def fp64_in_place(clu, lib, _lib, torch, A, b, nb64, n, g=None):
    """One FP64 LU + solve factoring A itself (no copy: at n = 131072 the
    FP64 matrix and a copy do not both fit beside the other buffers), then A
    regenerated from its seed for the residual.  With g, also the
    transposed solve A^T lam = g from the same factors (untimed).  Returns
    (record, x64, lam64)."""
    import ctypes
    flops = 2.0 / 3.0 * n ** 3
    x64 = b.clone()
    torch.cuda.synchronize()
    lib.charmlu_profile(1)
    t0 = time.perf_counter()
    ipiv, info = clu.getrf(A, nb64)
    clu.getrs(A, ipiv, x64)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    prof = (ctypes.c_double * 20)()
    _lib.check(lib.charmlu_profile_read(prof, 20), "charmlu_profile_read")
    lib.charmlu_profile(0)
    lam64 = None
    if g is not None:
        lam64 = g.clone()
        clu.getrs(A, ipiv, lam64, trans=True)
    _lib.check(lib.charmlu_dgen_matrix(1, clu.lu.SEED_A, n, ctypes.c_void_p(A.data_ptr()), n,
                                       clu.lu._stream_ptr()), "charmlu_dgen_matrix")
    res = clu.residual(A, x64, b)
    upd = prof[1] / (prof[15] * 1e-3) / 1e12 if prof[15] > 0 else 0.0
    return {"tflops": round(flops / sec / 1e12, 3), "ms": round(sec * 1e3, 1), "block": nb64,
            "residual": res["residual"], "info": int(info.item()), "steps": 1,
            "trailing_update_tflops": round(upd, 3),
            "trailing_update_frac": round(upd / FP64_PEAK, 4),
            "note": "one step factoring A in place (no copy of A), after the mixed runs"}, x64, lam64


def mixed_bench(clu, lib, _lib, torch, n, nb, precs, steps, warmup, fp64_steps, fp64_nb=None,
                fp64_after=False, adjoint=False, nb_of=None):
    """Mixed-precision LU + FP64 GMRES-IR against the plain FP64 LU on the
    same diagonally dominant system (A U[-0.5,0.5) with a_ii = sum_j |a_ij|,
    b from MatGen).  A mixed step converts A to FP32, factors it with the
    low-precision trailing update and refines x in FP64 to the reference's
    scaled residual <= 1; a FP64 step copies A, factors and solves
    (fp64_after: one step after the mixed runs, factoring A in place).
    adjoint: after each form's timed run, the adjoint solve of the
    differentiable solve, A^T lam = g by FP64 GMRES-IR from the same stored
    low-precision factors (no refactoring; g = b reversed), timed, with its
    error against the FP64 LU's transposed solve when that runs after.
    Returns the "mixed" object of the JSON line."""
    import ctypes
    flops = 2.0 / 3.0 * n ** 3
    A = clu.gen_matrix(n, kind="diagdom")
    b = clu.gen_rhs(n)
    g = torch.flip(b, dims=[0]).contiguous() if adjoint else None
    lams = {}
    out = {"n": n, "block": nb,
           "workload": f"A U[-0.5,0.5) with a_ii = sum_j |a_ij| (diagonally dominant), n={n}"}
    # plain FP64 LU on the same system: the speed reference and the solution
    # the mixed results are compared with
    x64 = None
    xs = {}
    if fp64_steps > 0 and not fp64_after:
        # the FP64 LU with its own best block size (1024 measured fastest at
        # n = 65536: 57.0 TFLOP/s against 56.4 with 512 and 52.8 with 256)
        nb64 = fp64_nb or nb
        W = torch.empty_like(A)
        x64 = torch.empty_like(b)

        def step64():
            W.copy_(A)
            x64.copy_(b)
            ipiv, info = clu.getrf(W, nb64)
            clu.getrs(W, ipiv, x64)
            return info
        step64()  # warmup (also records the panel durations the update split uses)
        torch.cuda.synchronize()
        lib.charmlu_profile(1)
        t0 = time.perf_counter()
        for _ in range(fp64_steps):
            info = step64()
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / fp64_steps
        prof = (ctypes.c_double * 20)()
        _lib.check(lib.charmlu_profile_read(prof, 20), "charmlu_profile_read")
        lib.charmlu_profile(0)
        res = clu.residual(A, x64, b)
        g_work, g_busy = prof[1], prof[15]
        upd = g_work / (g_busy * 1e-3) / 1e12 if g_busy > 0 else 0.0
        out["fp64"] = {"tflops": round(flops / sec / 1e12, 3), "ms": round(sec * 1e3, 1),
                       "block": nb64, "residual": res["residual"], "info": int(info.item()),
                       "steps": fp64_steps,
                       "trailing_update_tflops": round(upd, 3),
                       "trailing_update_frac": round(upd / FP64_PEAK, 4)}
        del W
        torch.cuda.empty_cache()
    A32 = torch.empty((n, n), dtype=torch.float32, device="cuda")
    x = torch.empty_like(b)
    for key in precs:
        # "<prec>" factors with partial pivoting (low-precision updates:
        # speculative identity pivots, kept while every panel's multipliers
        # stay below 1/2), "<prec>-xchg" with partial pivoting exchanging
        # candidate rows at every panel (CHARMLU_SPECULATE=0), "<prec>-nopiv"
        # without interchanges (the matrix is diagonally dominant); nb_of:
        # a block size of its own for some of them
        nbk = (nb_of or {}).get(key, nb)
        prec, _, mode = key.partition("-")
        pivot = mode != "nopiv"
        env = {"CHARMLU_SPECULATE": "0"} if mode == "xchg" else {}
        split = [0.0, 0.0]

        last = {}

        def step():
            t_a = time.perf_counter()
            clu.lag2s(A, A32)
            saved = {k: os.environ.get(k) for k in env}
            os.environ.update(env)
            try:
                ipiv, info = clu.getrf_mixed(A32, nbk, prec, pivot=pivot)
                torch.cuda.synchronize()
            finally:
                for k, v in saved.items():
                    if v is None:
                        os.environ.pop(k, None)
                    else:
                        os.environ[k] = v
            t_b = time.perf_counter()
            _, st = clu.gesv_ir(A, A32, ipiv, b, x, tol=1.0)
            split[0] += t_b - t_a
            split[1] += time.perf_counter() - t_b
            last["ipiv"] = ipiv
            return info, st
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        lib.charmlu_profile(1)
        split[0] = split[1] = 0.0
        t0 = time.perf_counter()
        for _ in range(steps):
            info, st = step()
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / steps
        prof = (ctypes.c_double * 15)()
        _lib.check(lib.charmlu_profile_read(prof, 15), "charmlu_profile_read")
        lib.charmlu_profile(0)
        p = list(prof)
        res = clu.residual(A, x, b)
        rec = {"block": nbk, "tflops": round(flops / sec / 1e12, 3), "ms": round(sec * 1e3, 1),
               "residual": res["residual"], "residual_ns": res["residual_ns"],
               "info": int(info.item()), "passed": bool(res["residual"] < 16),
               "ir_cycles": st["cycles"], "ir_iterations": st["iterations"],
               "factor_ms": round(split[0] / steps * 1e3, 1),
               "refine_ms": round(split[1] / steps * 1e3, 1),
               "update_ms": round(p[2] / steps, 1),
               "update_tflops": round(p[1] / (p[2] * 1e-3) / 1e12, 1) if p[2] > 0 else None,
               "panel_stream_ms": round(p[8] / steps, 1),
               "interchange_ms": round(p[11] / steps, 1),
               "u_row_ms": round(p[14] / steps, 1)}
        # partial pivoting picks the diagonal pivots iff every computed
        # multiplier is below 1 in magnitude
        mx = 0.0
        for r in range(0, n, 4096):  # row blocks (one kernel over n^2 entries overflows)
            mx = max(mx, float(A32[r:r + 4096].tril(r - 1).abs().max().item()))
        rec["max_abs_l"] = mx
        identity = bool((last["ipiv"] == torch.arange(n, device="cuda",
                                                       dtype=last["ipiv"].dtype)).all().item())
        if not pivot:
            rec["pivoting"] = "none (max |l_ij| < 1: partial pivoting would agree)" \
                if mx < 1.0 else "none"
        elif mode == "xchg":
            rec["pivoting"] = "partial (candidate rows exchanged at every panel)"
        elif prec != "fp32" and identity:
            rec["pivoting"] = ("partial (speculative: every panel eliminated without "
                               "interchanges had all multipliers below 1/2, so partial "
                               "pivoting keeps every diagonal row)")
        else:
            rec["pivoting"] = "partial"
        if adjoint:
            # the timed run's factors are still in A32 (and ipiv)
            torch.cuda.synchronize()
            t_a = time.perf_counter()
            lam, sta = clu.gesv_ir_multi(A, A32, last["ipiv"], g, trans=True, tol=1.0)
            torch.cuda.synchronize()
            rec["adjoint"] = {"ms": round((time.perf_counter() - t_a) * 1e3, 1),
                              "ir_cycles": sta["cycles"], "ir_iterations": sta["iterations"],
                              "residual": sta["residual"],
                              "note": "A^T lam = g from the stored factors by FP64 GMRES-IR "
                                      "(the backward pass of the mixed differentiable solve; "
                                      "dA = -lam x^T is rank one per right-hand side and not "
                                      "formed: at this n it would take 137 GB)"}
            lams[key] = lam
        if x64 is not None:
            rec["error_vs_fp64"] = float(((x - x64).abs().max() / x64.abs().max()).item())
            rec["speedup_vs_fp64"] = round(rec["tflops"] / out["fp64"]["tflops"], 3)
        elif fp64_after:
            xs[key] = x.clone()
        out[key] = rec
    if fp64_after and fp64_steps > 0:
        del A32
        torch.cuda.empty_cache()
        out["fp64"], x64, lam64 = fp64_in_place(clu, lib, _lib, torch, A, b, fp64_nb or nb, n,
                                                g=g)
        for key, xk in xs.items():
            out[key]["error_vs_fp64"] = float(((xk - x64).abs().max() / x64.abs().max()).item())
            out[key]["speedup_vs_fp64"] = round(out[key]["tflops"] / out["fp64"]["tflops"], 3)
        for key, lk in lams.items():
            out[key]["adjoint"]["error_vs_fp64"] = float(
                ((lk - lam64).abs().max() / lam64.abs().max()).item())
    return out


def small_case(clu, torch, n, nb, steps=5):
    """The reference's own CPU-runnable case on the GPU: FP64 LU + solve of
    n = 4096, block 128 (the same generators), one untimed step, then
    ``steps`` timed ones; effective TFLOP/s over factor + solve, beside the
    CPU figure of the same case in cpu_baseline.c1."""
    A0 = clu.gen_matrix(n)
    b0 = clu.gen_rhs(n)
    A = torch.empty_like(A0)
    x = torch.empty_like(b0)

    def step():
        A.copy_(A0)
        x.copy_(b0)
        ipiv, info = clu.getrf(A, nb)
        clu.getrs(A, ipiv, x)
        return info

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        info = step()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / steps
    res = clu.residual(A0, x, b0)
    return {"n": n, "block": nb, "steps": steps, "ms": round(sec * 1e3, 3),
            "tflops": round(2.0 / 3.0 * n ** 3 / sec / 1e12, 3),
            "residual": res["residual"], "info": int(info.item())}


def c5_bench(clu, lib, _lib, torch, nb, fp64_nb, xchg_nb=2048):
    """C5: the FP8 form at n = 131072 (FP64 A 137 GB + FP32 factors 69 GB in
    HBM), unpivoted, pivoted (speculative identity pivots) and pivoted with
    exchanging panels, one step each after no warmup, then one
    FP64 LU + solve of the same system in place for the error and speed
    comparison.  Block 1024 for the grouped forms (groups of 4 steps: one
    update per group at K = 4096; 893 against 833 TFLOP/s with block 2048,
    829 with 512), 2048 for the exchanging form (one update per step, whose
    FP32 C traffic per flop halves with the block: 614 against 550), 1024
    for the FP64 LU."""
    rec = mixed_bench(clu, lib, _lib, torch, 131072, nb, ["fp8-nopiv", "fp8", "fp8-xchg"], 1, 0,
                      1, fp64_nb, fp64_after=True, adjoint=True, nb_of={"fp8-xchg": xchg_nb})
    rec["steps_note"] = "one timed step per form, no warmup (the first call of its shape)"
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--nb", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--gemm-only", action="store_true",
                    help="time only the trailing-update GEMM of step 0 (M=N=n-nb, K=nb)")
    ap.add_argument("--lp-gemm-only", action="store_true",
                    help="time only the low-precision trailing update of step 0 at --mixed-n")
    ap.add_argument("--lp-prec", default="bf16", choices=["bf16", "fp16", "fp8"],
                    help="operand format of --lp-gemm-only (fp8: quantization included)")
    ap.add_argument("--mixed-n", type=int, default=65536)
    ap.add_argument("--mixed-nb", type=int, default=1024)
    ap.add_argument("--mixed-fp64-nb", type=int, default=1024)
    ap.add_argument("--mixed-precs", default="bf16,fp16,bf16-xchg,bf16-nopiv")
    ap.add_argument("--mixed-steps", type=int, default=2)
    ap.add_argument("--mixed-fp64-steps", type=int, default=2)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-c5", action="store_true",
                    help="skip the FP8 form at n=131072 (C5) in the default run")
    ap.add_argument("--c5-nb", type=int, default=1024,
                    help="block size of the grouped FP8 forms at n=131072 (the FP64 LU uses "
                         "--mixed-fp64-nb)")
    ap.add_argument("--c5-xchg-nb", type=int, default=2048,
                    help="block size of the exchanging FP8 form at n=131072")
    ap.add_argument("--c5", action="store_true",
                    help="only the FP8 form at n=131072 (block --c5-nb, diagonally dominant; "
                         "FP64 A 137 GB + FP32 factors 69 GB in HBM), one step each")
    ap.add_argument("--mixed-only", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run a plain device copy of MIB MiB (calibrates PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the LU runs "
              "device-resident on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    import ctypes

    import torch

    import charmlu_amd as clu
    from charmlu_amd import _lib

    lib = _lib.load()
    _lib.check(lib.charmlu_init(0), "charmlu_init")
    n, nb = args.n, args.nb
    if args.copy_calibration:
        src = torch.ones(args.copy_calibration * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 8}))
        return
    if args.gemm_only:
        gemm_only(lib, _lib, torch, n, nb, args.steps, args.warmup)
        return
    if args.lp_gemm_only:
        lp_gemm_only(lib, _lib, torch, args.mixed_n, args.mixed_nb, args.steps, args.warmup,
                     args.lp_prec)
        return
    if args.c5:
        print(json.dumps({"c5": c5_bench(clu, lib, _lib, torch, args.c5_nb, args.mixed_fp64_nb,
                            args.c5_xchg_nb)}),
              flush=True)
        return
    if args.mixed_only:
        print(json.dumps(mixed_bench(clu, lib, _lib, torch, args.mixed_n, args.mixed_nb,
                                     args.mixed_precs.split(","), args.mixed_steps, args.warmup,
                                     args.mixed_fp64_steps, args.mixed_fp64_nb)), flush=True)
        return
    A0 = clu.gen_matrix(n)
    b0 = clu.gen_rhs(n)
    A = torch.empty_like(A0)
    x = torch.empty_like(b0)

    def step():
        A.copy_(A0)
        x.copy_(b0)
        ipiv, info = clu.getrf(A, nb)
        clu.getrs(A, ipiv, x)
        return ipiv, info

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lib.charmlu_profile(1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ipiv, info = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = (ctypes.c_double * 20)()
    _lib.check(lib.charmlu_profile_read(prof, 20), "charmlu_profile_read")
    lib.charmlu_profile(0)
    res = clu.residual(A0, x, b0)
    info_v = int(info.item())

    sec = (t1 - t0) / args.steps
    flops = 2.0 / 3.0 * n ** 3
    value = flops / sec / 1e12
    p = list(prof)
    g_cnt, g_work, g_ms = p[0:3]
    pan_cnt, pan_work, pan_ms = p[6:9]
    sw_cnt, sw_cols, sw_ms = p[9:12]
    tr_cnt, tr_work, tr_ms = p[12:15]
    # per launch: algorithmic flops of a launch over its average duration
    # (HIP events on the stream each launch runs on); the two parts of a
    # split update overlap, so the same flops over the time during which at
    # least one launch was running is reported beside it
    g_busy = p[15]
    achieved = (g_work / g_cnt) / (g_ms / g_cnt * 1e-3) / 1e12 if g_cnt and g_ms > 0 else 0.0
    achieved_union = g_work / (g_busy * 1e-3) / 1e12 if g_busy > 0 else 0.0
    # row interchanges: algorithmic bytes = 16 * (rows moved) * (columns moved)
    ip = ipiv.cpu().numpy()
    moved = 0
    for t in range(n // nb):
        seg = list(range(t * nb, (t + 1) * nb))
        pos = {}
        for j in seg:
            pj = int(ip[j])
            if pj != j:
                a, b_ = pos.get(j, j), pos.get(pj, pj)
                pos[j], pos[pj] = b_, a
        moved += sum(1 for k, v in pos.items() if k != v)
    swap_bytes = 16.0 * moved * (n - nb) * args.steps
    # HBM traffic of one trailing-update launch, from the committed PMC
    # summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated) of the step-0
    # launch of this configuration
    traffic = None
    traffic_note = None
    tpath = os.path.join(ROOT, "profiles", "dgemm_nn_lds_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            trec = json.load(f)
        if trec.get("launch", "").startswith(f"{n - nb}x{n - nb}x{nb} "):
            traffic = trec["traffic_bytes"]
            traffic_note = (f"step-0 launch {n - nb}x{n - nb}x{nb}: {trec['traffic_bytes'] / 1e9:.2f} GB "
                            f"HBM vs {trec['algorithmic_bytes'] / 1e9:.2f} GB algorithmic "
                            "(profiles/dgemm_nn_lds_traffic.md)")
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (A: counter generator U[-0.5,0.5) seed 2998389; b: MatGen seed 9934835)",
        "config": {"workload": f"FP64 LU with partial pivoting + solve, n={n}, block {nb}, 1 RHS",
                   "n": n, "block": nb, "nrhs": 1, "precision": "fp64"},
        "residual": res["residual"],
        "residual_ns": res["residual_ns"],
        "info": info_v,
        "passed": bool(res["residual"] < 16 and info_v == 0),
        "hpl_gflops": round((flops + 1.5 * n * n) / sec / 1e9, 2),
        "roofline": {
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK, 4),
            "traffic": traffic,
            "traffic_note": traffic_note,
            "kernel": "dgemm_nn_lds_kernel<3, 1, true> (trailing update, 128x128 tiles)",
            "launches": int(g_cnt),
            "avg_launch_ms": round(g_ms / g_cnt, 4) if g_cnt else None,
            "flops_per_launch": g_work / g_cnt if g_cnt else None,
            "busy_ms": round(g_busy, 2),
            "achieved_union": round(achieved_union, 3),
            "frac_union": round(achieved_union / FP64_PEAK, 4),
            "achieved_note": ("achieved = flops_per_launch / avg_launch_ms.  Each step's update "
                              "runs as two launches: the rows beside the next panel on the "
                              "update stream's 224 CUs, the rest on all 256 CUs once that panel "
                              "is done; the two overlap for a while, so achieved_union (all "
                              "flops / busy_ms, the time with at least one launch running) "
                              "counts that overlap once"),
        },
        "phases_ms_per_step": {
            "gemm_trailing_lds": round(g_ms / args.steps, 2),
            "gemm_other": round(p[5] / args.steps, 2),
            "panel_stream": round(pan_ms / args.steps, 2),
            "row_interchanges": round(sw_ms / args.steps, 2),
            "u_row_trsm": round(tr_ms / args.steps, 2),
        },
        "row_interchange_gbs": round(swap_bytes / (sw_ms * 1e-3) / 1e9, 1) if sw_ms > 0 else None,
    }
    out["c1_gpu"] = small_case(clu, torch, 4096, 128)
    out["block_dgemm"] = block_dgemm(lib, _lib, torch, nb, out["value"])
    if not args.no_mixed:
        del A, A0
        torch.cuda.empty_cache()
        out["mixed"] = mixed_bench(clu, lib, _lib, torch, args.mixed_n, args.mixed_nb,
                                   args.mixed_precs.split(","), args.mixed_steps, args.warmup,
                                   args.mixed_fp64_steps, args.mixed_fp64_nb)
    if not args.no_mixed and not args.no_c5:
        torch.cuda.empty_cache()
        out["c5"] = c5_bench(clu, lib, _lib, torch, args.c5_nb, args.mixed_fp64_nb,
                            args.c5_xchg_nb)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, nb)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
