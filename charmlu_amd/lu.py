"""Python host side over the C ABI: device-resident LU, solve, generation and
the reference driver's validation, on torch CUDA tensors (torch supplies the
device memory and the stream; all arithmetic runs in libcharmlu_amd.so).

Mirrors the reference driver (benchmark.C): ``LUConfig`` holds its argv
(benchmark.C:79-149, luConfig.h:8-62); ``run`` generates A and b, factors,
solves and validates exactly as LUSolver/BenchmarkLUBlk do
(benchmark.C:231-400, 430-573) and reports the same figures.
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np

from . import _lib
from ._lib import check

SEED_A = 2998389   # benchmark.C:436
SEED_B = 9934835   # benchmark.C:441
EPS = float(np.finfo(np.float64).eps)


def _torch():
    import torch
    return torch


def _stream_ptr(stream=None):
    torch = _torch()
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes_void(s.cuda_stream)


def ctypes_void(v):
    import ctypes
    return ctypes.c_void_p(int(v)) if v else None


def _ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else None


def _require(t, dtype, name, shape=None):
    """Check a tensor argument before its device pointer goes to the C ABI:
    a CUDA tensor of ``dtype``, row-major (unit stride along rows; a vector
    contiguous), of ``shape`` when given."""
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a CUDA tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if t.dim() == 2 and t.stride(1) != 1:
        raise ValueError(f"{name} must be row-major (stride(1) == 1)")
    if t.dim() == 1 and t.numel() > 1 and t.stride(0) != 1:
        raise ValueError(f"{name} must be contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    return t


def _require_ipiv(ipiv, n):
    """Pivot vectors are int32 CUDA tensors of n entries."""
    torch = _torch()
    return _require(ipiv, torch.int32, "ipiv", (n,))


def check_info(info, what: str = "factorization"):
    """Raise on a nonzero info word: j+1 > 0 a near-singular pivot at column
    j (the reference aborts there, lu.ci:339-341), -1 a panel exchange that
    timed out (the factors are invalid)."""
    torch = _torch()
    v = int(info.item()) if isinstance(info, torch.Tensor) else int(info)
    if v > 0:
        raise torch.linalg.LinAlgError(
            f"{what}: singular pivot at column {v - 1} (|u_jj| <= 100 eps)")
    if v < 0:
        raise _lib.CharmluError(
            f"{what}: a panel strip's cross-workgroup exchange timed out (info {v}); "
            "the factors are invalid")
    return v


# ---------------------------------------------------------------------------
# generation (benchmark.C:27-74, 430-460)
# ---------------------------------------------------------------------------
def gen_matrix(n: int, seed: int = SEED_A, kind: str = "uniform", device="cuda"):
    """A (n x n FP64) on the device: 'uniform' U[-0.5,0.5) or 'diagdom'
    (a_ii = sum_j |a_ij|)."""
    torch = _torch()
    A = torch.empty((n, n), dtype=torch.float64, device=device)
    k = {"uniform": 0, "diagdom": 1}[kind]
    check(_lib.load().charmlu_dgen_matrix(k, seed, n, _ptr(A), max(n, 1), _stream_ptr()),
          "charmlu_dgen_matrix")
    return A


def gen_rhs(n: int, seed: int = SEED_B, skip: int = 0, device="cuda"):
    torch = _torch()
    b = torch.empty(n, dtype=torch.float64, device=device)
    check(_lib.load().charmlu_dgen_rhs(seed, skip, n, _ptr(b), _stream_ptr()), "charmlu_dgen_rhs")
    return b


# ---------------------------------------------------------------------------
# factor / solve
# ---------------------------------------------------------------------------
def getrf(A, block: int):
    """In-place LU with partial pivoting of the row-major FP64 CUDA tensor A.
    Returns (ipiv int32 tensor, info int32 tensor of one element)."""
    torch = _torch()
    _require(A, torch.float64, "A")
    n = A.shape[0]
    if A.shape[1] != n:
        raise ValueError("A must be square")
    ipiv = torch.empty(max(n, 1), dtype=torch.int32, device=A.device)
    info = torch.zeros(1, dtype=torch.int32, device=A.device)
    check(_lib.load().charmlu_dgetrf(n, block, _ptr(A), A.stride(0), _ptr(ipiv), _ptr(info),
                                     _stream_ptr()), "charmlu_dgetrf")
    return ipiv[:n], info


def getrs(LU, ipiv, B, trans: bool = False):
    """Solve op(A) X = B in place (B: n or n x nrhs FP64 CUDA tensor)."""
    torch = _torch()
    _require(LU, torch.float64, "LU")
    n = LU.shape[0]
    _require(LU, torch.float64, "LU", (n, n))
    _require_ipiv(ipiv, n)
    _require(B, torch.float64, "B")
    if B.shape[0] != n or B.dim() not in (1, 2):
        raise ValueError(f"B must be ({n},) or ({n}, nrhs)")
    B2 = B.view(n, 1) if B.dim() == 1 else B
    nrhs = B2.shape[1]
    check(_lib.load().charmlu_dgetrs(_lib.TRANS if trans else _lib.NO_TRANS, n, nrhs, _ptr(LU),
                                     LU.stride(0), _ptr(ipiv), _ptr(B2), B2.stride(0),
                                     _stream_ptr()), "charmlu_dgetrs")
    return B


_PREC = {"fp32": _lib.PREC_FP32, "bf16": _lib.PREC_BF16, "fp16": _lib.PREC_FP16,
         "fp8": _lib.PREC_FP8}


def getrf_mixed(A32, block: int, prec: str = "bf16", pivot: bool = True):
    """In-place LU of the row-major FP32 CUDA tensor A32 with an FP32 panel
    and the trailing update in ``prec`` ('fp32', 'bf16', 'fp16', or 'fp8':
    e4m3 with power-of-two scales per 128-row tile, block a multiple of 128;
    FP32 accumulation).  pivot=False factors without interchanges (for matrices
    that need none, e.g. diagonally dominant ones).  Returns (ipiv, info)
    like ``getrf``."""
    torch = _torch()
    _require(A32, torch.float32, "A32")
    n = A32.shape[0]
    if A32.shape[1] != n:
        raise ValueError("A32 must be square")
    if prec not in _PREC:
        raise ValueError(f"prec must be one of {sorted(_PREC)}")
    ipiv = torch.empty(max(n, 1), dtype=torch.int32, device=A32.device)
    info = torch.zeros(1, dtype=torch.int32, device=A32.device)
    code = _PREC[prec] | (0 if pivot else _lib.MIXED_NOPIV)
    check(_lib.load().charmlu_sgetrf_mixed(n, block, code, _ptr(A32), A32.stride(0),
                                           _ptr(ipiv), _ptr(info), _stream_ptr()),
          "charmlu_sgetrf_mixed")
    return ipiv[:n], info


def lag2s(A, out=None):
    """FP32 copy of the FP64 CUDA matrix A (the mixed path's working matrix)."""
    torch = _torch()
    _require(A, torch.float64, "A")
    m, n = A.shape
    if out is None:
        out = torch.empty((m, n), dtype=torch.float32, device=A.device)
    _require(out, torch.float32, "out")
    check(_lib.load().charmlu_dlag2s(m, n, _ptr(A), max(A.stride(0), 1), _ptr(out),
                                     max(out.stride(0), 1), _stream_ptr()), "charmlu_dlag2s")
    return out


def getrs32(LU32, ipiv, B, trans: bool = False):
    """``getrs`` on FP32 factors and an FP32 right-hand side."""
    torch = _torch()
    _require(LU32, torch.float32, "LU32")
    n = LU32.shape[0]
    _require(LU32, torch.float32, "LU32", (n, n))
    _require_ipiv(ipiv, n)
    _require(B, torch.float32, "B")
    if B.shape[0] != n or B.dim() not in (1, 2):
        raise ValueError(f"B must be ({n},) or ({n}, nrhs)")
    B2 = B.view(n, 1) if B.dim() == 1 else B
    check(_lib.load().charmlu_sgetrs(_lib.TRANS if trans else _lib.NO_TRANS, n, B2.shape[1],
                                     _ptr(LU32), LU32.stride(0), _ptr(ipiv), _ptr(B2),
                                     B2.stride(0), _stream_ptr()), "charmlu_sgetrs")
    return B


def gesv_ir(A, LU32, ipiv, b, x=None, tol: float = 1.0, max_iter: int = 30,
            restart: int = 50):
    """FP64 GMRES-IR: x with A x = b (FP64 A, b) preconditioned by the FP32
    factors of ``getrf_mixed``.  Returns (x, stats dict)."""
    torch = _torch()
    _require(A, torch.float64, "A")
    n = A.shape[0]
    _require(A, torch.float64, "A", (n, n))
    _require(LU32, torch.float32, "LU32", (n, n))
    _require_ipiv(ipiv, n)
    _require(b, torch.float64, "b", (n,))
    if x is None:
        x = torch.empty(n, dtype=torch.float64, device=A.device)
    _require(x, torch.float64, "x", (n,))
    st = np.zeros(4, dtype=np.float64)
    check(_lib.load().charmlu_dgesv_ir(n, _ptr(A), A.stride(0), _ptr(LU32), LU32.stride(0),
                                       _ptr(ipiv), _ptr(b), _ptr(x), float(tol), int(max_iter),
                                       int(restart), ctypes_void(st.ctypes.data), _stream_ptr()),
          "charmlu_dgesv_ir")
    return x, dict(cycles=int(st[0]), iterations=int(st[1]), residual=float(st[2]),
                   residual_ns=float(st[3]))


def gesv_ir_multi(A, LU32, ipiv, B, X=None, trans: bool = False, tol: float = 1.0,
                  max_iter: int = 30, restart: int = 50):
    """FP64 GMRES-IR for op(A) X = B with several right-hand sides (B: n x
    nrhs FP64), preconditioned by the FP32 factors of ``getrf_mixed``;
    trans=True solves A^T X = B from the same factors.  Returns (X, stats
    dict: most cycles of any column, inner iterations of all columns,
    largest final residuals)."""
    torch = _torch()
    _require(A, torch.float64, "A")
    n = A.shape[0]
    _require(A, torch.float64, "A", (n, n))
    _require(LU32, torch.float32, "LU32", (n, n))
    _require_ipiv(ipiv, n)
    _require(B, torch.float64, "B")
    if B.shape[0] != n or B.dim() not in (1, 2):
        raise ValueError(f"B must be ({n},) or ({n}, nrhs)")
    B2 = B.view(n, 1) if B.dim() == 1 else B
    if X is None:
        X = torch.empty_like(B, memory_format=torch.contiguous_format)
    _require(X, torch.float64, "X", tuple(B.shape))
    X2 = X.view(n, 1) if X.dim() == 1 else X
    st = np.zeros(4, dtype=np.float64)
    check(_lib.load().charmlu_dgesv_ir_ex(_lib.TRANS if trans else _lib.NO_TRANS, n,
                                          B2.shape[1], _ptr(A), A.stride(0), _ptr(LU32),
                                          LU32.stride(0), _ptr(ipiv), _ptr(B2),
                                          max(B2.stride(0), 1), _ptr(X2), max(X2.stride(0), 1),
                                          float(tol), int(max_iter), int(restart),
                                          ctypes_void(st.ctypes.data), _stream_ptr()),
          "charmlu_dgesv_ir_ex")
    return X, dict(cycles=int(st[0]), iterations=int(st[1]), residual=float(st[2]),
                   residual_ns=float(st[3]))


def residual(A, x, b):
    """Reference scaled residual (benchmark.C:385-400) and the normwise form
    ||Ax-b||_inf / (||A||_inf ||x||_inf n eps).  Returns a dict."""
    torch = _torch()
    _require(A, torch.float64, "A")
    n = A.shape[0]
    _require(A, torch.float64, "A", (n, n))
    _require(x, torch.float64, "x", (n,))
    _require(b, torch.float64, "b", (n,))
    out = torch.zeros(5, dtype=torch.float64, device=A.device)
    check(_lib.load().charmlu_dresidual(n, _ptr(A), A.stride(0), _ptr(x), _ptr(b), None,
                                        _ptr(out), _stream_ptr()), "charmlu_dresidual")
    amax, bmax, xmax, rmax, arow = out.cpu().tolist()
    ref = rmax / ((amax * xmax + bmax) * n * EPS) if n else 0.0
    ns = rmax / (arow * xmax * n * EPS) if n and arow * xmax > 0 else 0.0
    return dict(residual=ref, residual_ns=ns, a_max=amax, b_max=bmax, x_max=xmax,
                r_max=rmax, a_inf=arow)


# ---------------------------------------------------------------------------
# the reference driver
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class LUConfig:
    """argv of the reference benchmark: <n> <b> <memMB> [<pivotBatch>
    <mapping> [<tileRows> <tileCols> <rotate> <stride>]] (benchmark.C:79-134).
    Only n and b change the computation here; memThreshold, mapping and the
    PE tiling steer the reference's multi-process scheduler and are kept for
    interface compatibility (one device holds the whole matrix).  The pivot
    batch size groups interchanges in the reference (lu.C:300-302); the row
    permutation it produces is the same for any batch size."""
    matrix_size: int
    block_size: int
    mem_threshold: int = 500
    pivot_batch_size: int = 0
    mapping_scheme: int = 2
    pe_tile_rows: int = 0
    pe_tile_cols: int = 0
    pe_tile_rotate: int = 0
    pe_tile_stride: int = 0

    def __post_init__(self):
        if self.pivot_batch_size <= 0:
            self.pivot_batch_size = max(1, self.block_size // 4)  # benchmark.C:95-98

    @classmethod
    def from_argv(cls, argv):
        if len(argv) < 3:
            raise SystemExit("Usage: <matrix size> <block size> <mem threshold> "
                             "[<pivot batch size> <mapping scheme> [<peTileRows> <peTileCols> "
                             "<peRotate> <peStride>] ]")
        vals = [int(a) for a in argv]
        names = ["matrix_size", "block_size", "mem_threshold", "pivot_batch_size",
                 "mapping_scheme", "pe_tile_rows", "pe_tile_cols", "pe_tile_rotate",
                 "pe_tile_stride"]
        return cls(**dict(zip(names, vals)))

    def validate(self):
        if self.block_size <= 0 or self.matrix_size % self.block_size != 0:
            raise ValueError(f"The matrix size {self.matrix_size} should be a multiple of block "
                             f"size {self.block_size}!")  # benchmark.C:143-147


def hpl_gflops(n: int, seconds: float) -> float:
    """(2/3 n^3 + 3/2 n^2) / t  (benchmark.C:344-349), in GFlop/s."""
    return (2.0 / 3.0 * n ** 3 + 1.5 * n ** 2) / seconds / 1e9


def testdgemm(b: int, trials: int = 10) -> float:
    """Seconds of one b x b x b block product C -= A B on the device
    (charmlu_dgemm_dev), the mean of ``trials`` calls each timed to its
    completion after one untimed call: the reference's testdgemm
    (benchmark.C:297-341), which times one cblas_dgemm of the block size."""
    torch = _torch()
    g = torch.Generator(device="cpu").manual_seed(0)
    A, B, C = ((torch.rand(b, b, generator=g, dtype=torch.float64) - 0.5).cuda() for _ in range(3))
    lib = _lib.load()

    def call():
        check(lib.charmlu_dgemm_dev(_lib.NO_TRANS, _lib.NO_TRANS, b, b, b, -1.0, _ptr(A), b,
                                    _ptr(B), b, 1.0, _ptr(C), b, _stream_ptr()),
              "charmlu_dgemm_dev")
        torch.cuda.synchronize()

    call()
    total = 0.0
    for _ in range(trials):
        t0 = time.perf_counter()
        call()
        total += time.perf_counter() - t0
    return total / trials


def run(cfg: LUConfig, seed_a: int = SEED_A, seed_b: int = SEED_B, kind: str = "uniform",
        verbose: bool = True):
    """Generate, factor + solve (timed, benchmark.C:242-245 -> 343-349),
    validate (benchmark.C:462-573).  Returns a dict."""
    torch = _torch()
    cfg.validate()
    n, b = cfg.matrix_size, cfg.block_size
    A = gen_matrix(n, seed_a, kind)
    rhs = gen_rhs(n, seed_b)
    LU = A.clone()
    x = rhs.clone()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    ipiv, info = getrf(LU, b)
    getrs(LU, ipiv, x)
    stop.record()
    torch.cuda.synchronize()
    seconds = start.elapsed_time(stop) / 1e3
    if int(info.item()) < 0:
        check_info(info, "getrf")  # a timed-out panel exchange: the factors are invalid
    res = residual(A, x, rhs)
    out = dict(n=n, block=b, seconds=seconds, info=int(info.item()),
               gflops_hpl=hpl_gflops(n, seconds),
               tflops_eff=(2.0 / 3.0 * n ** 3) / seconds / 1e12 if seconds > 0 else math.inf,
               ipiv=ipiv, x=x, **res)
    # the block-dgemm line and the LU's rate against it (benchmark.C:374-381;
    # one device stands for the reference's cores)
    dgemm_s = testdgemm(b)
    out["dgemm_ms"] = dgemm_s * 1e3
    out["dgemm_gflops"] = 2.0 * b ** 3 / dgemm_s / 1e9
    out["percent_of_dgemm"] = out["gflops_hpl"] / out["dgemm_gflops"] * 100.0
    if verbose:
        print(f"RESULT procs: 1 device\tblock size: {b}\tTime(s): {seconds:f}")
        print(f"HPL flop rate: {out['gflops_hpl']:f} GFlops")
        print(f"The dgemm {b} x {b} takes {out['dgemm_ms']:g} ms and achieves "
              f"{out['dgemm_gflops']:g} GFlop/sec")
        print(f"Percent of DGEMM is: {out['percent_of_dgemm']:g}%")
        print(f"epsilon = {EPS:e}\nresidual = {res['residual']:f}")
        if res["residual"] > 16:
            print("=== WARNING: Scaled residual is greater than 16 - OUT OF SPEC ===")
    return out
