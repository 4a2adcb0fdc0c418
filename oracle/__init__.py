"""ctypes view of the CPU oracle (``liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``, always as the test oracle or the
timed CPU baseline, never as the product path.  The oracle restates
charmlu's block LU (lu.C, lu.ci, benchmark.C of the reference); see
``lu_oracle.h`` for what pins each function.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

MKL_PATH = "/opt/conda/lib/libmkl_rt.so"

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build() -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_matgen_doubles.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_long, _dp]
        L.orc_matgen_nth_int.argtypes = [ctypes.c_int, ctypes.c_long]
        L.orc_matgen_nth_int.restype = ctypes.c_int
        L.orc_gen_value.argtypes = [ctypes.c_uint64, ctypes.c_long, ctypes.c_long]
        L.orc_gen_value.restype = ctypes.c_double
        L.orc_gen_uniform.argtypes = [ctypes.c_uint64, ctypes.c_long, ctypes.c_long,
                                      ctypes.c_long, ctypes.c_long, _dp, ctypes.c_long]
        L.orc_gen_diagdom.argtypes = [ctypes.c_uint64, ctypes.c_long, _dp, ctypes.c_long]
        L.orc_dgemm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                _dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double,
                                _dp, ctypes.c_int]
        L.orc_dtrsm_llnu.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp,
                                     ctypes.c_int]
        L.orc_dger.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int,
                               _dp, ctypes.c_int, _dp, ctypes.c_int]
        L.orc_dgemv.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int,
                                _dp, ctypes.c_double, _dp]
        L.orc_use_cblas.argtypes = [ctypes.c_char_p]
        L.orc_use_cblas.restype = ctypes.c_int
        L.orc_using_cblas.restype = ctypes.c_int
        L.orc_swaps_to_cycles.argtypes = [ctypes.c_int, _ip, _ip, _ip, _ip]
        L.orc_swaps_to_cycles.restype = ctypes.c_int
        L.orc_apply_cycles.argtypes = [ctypes.c_int, _ip, _ip, _dp, ctypes.c_long,
                                       ctypes.c_int, ctypes.c_int]
        L.orc_getrf.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_long,
                                _dp, _ip, ctypes.c_int]
        L.orc_getrf.restype = ctypes.c_int
        L.orc_set_tie_log.argtypes = [_ip, _dp, _dp]
        L.orc_getf2.argtypes = [ctypes.c_int, _dp, ctypes.c_long, _ip]
        L.orc_getf2.restype = ctypes.c_int
        L.orc_solve.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_long, _dp]
        L.orc_apply_ipiv.argtypes = [ctypes.c_int, _ip, _dp]
        L.orc_residual.argtypes = [ctypes.c_int, _dp, ctypes.c_long, _dp, _dp, _dp, _dp]
        L.orc_residual.restype = ctypes.c_double
        L.orc_wtime.restype = ctypes.c_double
        L.orc_set_max_steps.argtypes = [ctypes.c_int]
        L.orc_lapack_gesv.argtypes = [ctypes.c_char_p, ctypes.c_int, _dp, ctypes.c_long, _dp,
                                      _ip, ctypes.c_int, _dp]
        L.orc_lapack_gesv.restype = ctypes.c_int
        _lib = L
    return _lib


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


# ---- generators -------------------------------------------------------------
SEED_A = 2998389   # benchmark.C:436
SEED_B = 9934835   # benchmark.C:441


def matgen(seed: int, count: int, skip: int = 0) -> np.ndarray:
    out = np.empty(count, np.float64)
    lib().orc_matgen_doubles(seed, skip, count, _d(out))
    return out


def matgen_nth_int(seed: int, k: int) -> int:
    return lib().orc_matgen_nth_int(seed, k)


def gen_value(seed: int, i: int, j: int) -> float:
    return lib().orc_gen_value(seed, i, j)


def gen_uniform(seed: int, m: int, n: int | None = None, row0: int = 0, col0: int = 0) -> np.ndarray:
    """m x n block of the counter generator starting at (row0, col0)."""
    n = m if n is None else n
    A = np.empty((m, n), np.float64)
    lib().orc_gen_uniform(seed, row0, col0, m, n, _d(A), n)
    return A


def gen_diagdom(seed: int, n: int) -> np.ndarray:
    A = np.empty((n, n), np.float64)
    lib().orc_gen_diagdom(seed, n, _d(A), n)
    return A


def gen_rhs(n: int, seed: int = SEED_B) -> np.ndarray:
    return matgen(seed, n)


# ---- BLAS surface -----------------------------------------------------------
def use_mkl(path: str = MKL_PATH) -> bool:
    if not os.path.exists(path):
        return False
    return lib().orc_use_cblas(path.encode()) == 0


def using_cblas() -> bool:
    return bool(lib().orc_using_cblas())


def dgemm(alpha, A, B, beta, C):
    M, K = A.shape
    N = B.shape[1]
    assert B.shape[0] == K and C.shape == (M, N)
    lib().orc_dgemm(M, N, K, alpha, _d(A), A.shape[1], _d(B), B.shape[1], beta, _d(C), C.shape[1])
    return C


def dtrsm_llnu(L, B):
    M, N = B.shape
    assert L.shape[0] >= M and L.shape[1] >= M
    lib().orc_dtrsm_llnu(M, N, _d(L), L.shape[1], _d(B), B.shape[1])
    return B


def dger(alpha, x, y, A):
    M, N = A.shape
    assert x.shape == (M,) and y.shape == (N,)
    lib().orc_dger(M, N, alpha, _d(x), 1, _d(y), 1, _d(A), A.shape[1])
    return A


def dgemv(alpha, A, x, beta, y):
    M, N = A.shape
    assert x.shape == (N,) and y.shape == (M,)
    lib().orc_dgemv(M, N, alpha, _d(A), A.shape[1], _d(x), beta, _d(y))
    return y


# ---- pivots -----------------------------------------------------------------
def swaps_to_cycles(pos, row):
    pos = np.ascontiguousarray(pos, np.int32)
    row = np.ascontiguousarray(row, np.int32)
    cnt = len(pos)
    seq_index = np.zeros(cnt + 1, np.int32)
    seq = np.zeros(max(2 * cnt, 1), np.int32)
    nc = lib().orc_swaps_to_cycles(cnt, _i(pos), _i(row), _i(seq_index), _i(seq))
    return nc, seq_index, seq


def apply_cycles(nc, seq_index, seq, A, c0=0, c1=None):
    c1 = A.shape[1] if c1 is None else c1
    lib().orc_apply_cycles(nc, _i(seq_index), _i(seq), _d(A), A.shape[1], c0, c1)
    return A


# ---- factorization / solve --------------------------------------------------
def getrf(A: np.ndarray, b: int, batch: int = 0, rhs: np.ndarray | None = None,
          nthreads: int = 1, ties: dict | None = None):
    """In-place blocked LU; returns (info, ipiv).  With a dict ``ties``, fills
    it with the near-tie log: ``mag1`` (|pivot| per column), ``row2`` and
    ``mag2`` (the best other candidate row per column and its |value|)."""
    n = A.shape[0]
    ipiv = np.zeros(max(n, 1), np.int32)
    if ties is not None:
        row2 = np.full(max(n, 1), -1, np.int32)
        mag2 = np.zeros(max(n, 1))
        mag1 = np.zeros(max(n, 1))
        lib().orc_set_tie_log(_i(row2), _d(mag2), _d(mag1))
    try:
        info = lib().orc_getrf(n, b, batch, _d(A), A.shape[1],
                               _d(rhs) if rhs is not None else None, _i(ipiv), nthreads)
    finally:
        if ties is not None:
            lib().orc_set_tie_log(None, None, None)
    if ties is not None:
        ties.update(row2=row2[:n], mag2=mag2[:n], mag1=mag1[:n])
    return info, ipiv[:n]


def near_ties(ties: dict, ulps: float = 4.0) -> np.ndarray:
    """Columns whose two best pivot candidates lie within ``ulps`` ulp of
    each other (the pivot choice there is decided by rounding)."""
    m1, m2 = ties["mag1"], ties["mag2"]
    return np.nonzero((ties["row2"] >= 0) & (m1 - m2 <= ulps * np.spacing(m1)))[0]


def is_near_tie(ties: dict, j: int, row: int, ulps: float = 4.0) -> bool:
    """Is choosing ``row`` instead of the oracle's pivot at column ``j`` a
    near-tie: the oracle's runner-up row, within ``ulps`` ulp of the pivot?"""
    if int(ties["row2"][j]) != int(row):
        return False
    m1, m2 = float(ties["mag1"][j]), float(ties["mag2"][j])
    return m1 - m2 <= ulps * float(np.spacing(m1))


def getf2(A: np.ndarray):
    n = A.shape[0]
    ipiv = np.zeros(max(n, 1), np.int32)
    info = lib().orc_getf2(n, _d(A), A.shape[1], _i(ipiv))
    return info, ipiv[:n]


def solve(LU: np.ndarray, b: int, rhs: np.ndarray) -> np.ndarray:
    lib().orc_solve(LU.shape[0], b, _d(LU), LU.shape[1], _d(rhs))
    return rhs


def apply_ipiv(ipiv, v):
    ipiv = np.ascontiguousarray(ipiv, np.int32)
    lib().orc_apply_ipiv(len(ipiv), _i(ipiv), _d(v))
    return v


def residual(A, x, b):
    """Returns (reference scaled residual, normwise residual
    ||b-Ax||_inf / (||A||_inf ||x||_inf n eps), maxima)."""
    out4 = np.zeros(4)
    ns = np.zeros(1)
    n = A.shape[0]
    r = lib().orc_residual(n, _d(A), A.shape[1], _d(x), _d(b), _d(out4), _d(ns))
    return r, float(ns[0]), out4


def wtime() -> float:
    return lib().orc_wtime()


def set_max_steps(k: int) -> None:
    """Run only the first k block steps in later ``getrf`` calls (0: all)."""
    lib().orc_set_max_steps(int(k))


def sample_flops(n: int, b: int, steps: int) -> float:
    """LU flops of the first ``steps`` block steps of an n x n factorization
    with block b: 2/3 n^3 minus the 2/3 m^3 of the m = n - steps*b trailing
    matrix left (the leading term of each)."""
    m = max(n - steps * b, 0)
    return 2.0 / 3.0 * (float(n) ** 3 - float(m) ** 3)


def lapack_gesv(A: np.ndarray, bvec: np.ndarray, nthreads: int, path: str = MKL_PATH):
    """The vendor LAPACK's multithreaded LU + solve of A x = b (in place on
    copies); returns (info, x, seconds).  Call it in a process where the
    oracle's sequential BLAS (``use_mkl``) is not loaded: the library's
    threading can be chosen once per process."""
    Aw = np.ascontiguousarray(A).copy()
    x = np.ascontiguousarray(bvec).copy()
    n = Aw.shape[0]
    ipiv = np.zeros(max(n, 1), np.int32)
    sec = np.zeros(1)
    info = lib().orc_lapack_gesv(path.encode(), n, _d(Aw), Aw.shape[1], _d(x), _i(ipiv),
                                 int(nthreads), _d(sec))
    return info, x, float(sec[0])


def lu_solve_problem(n: int, b: int, batch: int = 0, nthreads: int = 1,
                     seed_a: int = SEED_A, seed_b: int = SEED_B, A: np.ndarray | None = None,
                     ties: bool = False, validate: bool = True):
    """The reference driver's whole run on the CPU: generate, factor, solve,
    validate (benchmark.C:231-400).  Returns a dict (with ``ties``, the
    near-tie log of ``getrf`` under the key "ties")."""
    A0 = gen_uniform(seed_a, n) if A is None else A
    bvec = gen_rhs(n, seed_b)
    LU = A0.copy()
    beta = bvec.copy()
    tlog = {} if ties else None
    t0 = wtime()
    info, ipiv = getrf(LU, b, batch, beta, nthreads, ties=tlog)
    solve(LU, b, beta)
    t1 = wtime()
    r, ns, m4 = residual(A0, beta, bvec) if validate else (None, None, None)
    out = dict(info=info, ipiv=ipiv, x=beta, LU=LU, residual=r, residual_ns=ns,
               maxima=m4, seconds=t1 - t0, A=A0, b=bvec)
    if ties:
        out["ties"] = tlog
    return out
