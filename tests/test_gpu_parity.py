"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Integer/byte work (generators, pivots, row moves) is compared bit for bit;
floating-point results within stated tolerances:
  * block ops: |C - C_ref| <= 8 * K * eps * (|A||B| + |C|)  (MFMA and the
    oracle sum in different orders)
  * LU: pivot sequence index-for-index with the oracle (a difference is
    accepted only at a near-tie, |a_p| and |a_q| within 4 ulp, and logged);
    x within 1e-10 relative (normwise, ||dx||_inf / ||x||_inf); the
    reference's scaled residual < 16 (test_lu.pl:75-76).
"""
import ctypes
import math
import os

import numpy as np
import pytest
import scipy.linalg as sl

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "oracle_fixtures.npz")
LARGE = os.path.join(os.path.dirname(__file__), "golden", "large_fixtures.npz")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _sync():
    import torch
    torch.cuda.synchronize()


# ---------------------------------------------------------------------------
# generators: bit-exact
# ---------------------------------------------------------------------------
def test_gen_matrix_bit_exact(dev, orc):
    A = dev.gen_matrix(777, seed=orc.SEED_A).cpu().numpy()
    assert np.array_equal(A, orc.gen_uniform(orc.SEED_A, 777))


def test_gen_rhs_bit_exact(dev, orc):
    g = np.load(GOLDEN)
    b = dev.gen_rhs(4096).cpu().numpy()
    assert np.array_equal(b, g["rhs_4096"])
    assert np.array_equal(dev.gen_rhs(100, skip=1000).cpu().numpy(),
                          orc.matgen(orc.SEED_B, 100, skip=1000))


def test_gen_diagdom(dev, orc):
    A = dev.gen_matrix(300, seed=9, kind="diagdom").cpu().numpy()
    R = orc.gen_diagdom(9, 300)
    off = ~np.eye(300, dtype=bool)
    assert np.array_equal(A[off], R[off])
    assert np.allclose(np.diag(A), np.diag(R), rtol=1e-14, atol=0)


# ---------------------------------------------------------------------------
# block ops (platformBlas.h surface)
# ---------------------------------------------------------------------------
def _gemm_tol(A, B, C, alpha, beta):
    K = A.shape[1]
    return 8 * max(K, 1) * EPS * (abs(alpha) * np.abs(A) @ np.abs(B) + abs(beta) * np.abs(C)) + 1e-300


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (17, 33, 5), (128, 128, 256), (300, 200, 256),
                                   (1000, 64, 256), (64, 1000, 32), (257, 31, 100),
                                   # narrow products (64 x 32 tiles): ragged M and K
                                   (1000, 16, 16), (4097, 32, 33),
                                   (2048, 2048, 256), (513, 1031, 77),
                                   # >= 512 full 128x128 tiles: the direct-to-LDS path
                                   (4096, 2048, 256), (2816, 3072, 264)])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_dgemm_dev(dev, M, N, K, ta, tb):
    from charmlu_amd import _lib
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = rng.uniform(-0.5, 0.5, (K, M) if ta else (M, K))
    B = rng.uniform(-0.5, 0.5, (N, K) if tb else (K, N))
    C = rng.uniform(-0.5, 0.5, (M, N))
    alpha, beta = -1.0, 1.0
    opA = A.T if ta else A
    opB = B.T if tb else B
    ref = alpha * opA @ opB + beta * C
    dA, dB, dC = _t(A), _t(B), _t(C)
    rc = _lib.load().charmlu_dgemm_dev(_lib.TRANS if ta else _lib.NO_TRANS,
                                       _lib.TRANS if tb else _lib.NO_TRANS, M, N, K, alpha, _p(dA),
                                       A.shape[1], _p(dB), B.shape[1], beta, _p(dC), N, None)
    assert rc == 0
    _sync()
    got = dC.cpu().numpy()
    assert np.all(np.abs(got - ref) <= _gemm_tol(opA, opB, C, alpha, beta))


@pytest.mark.parametrize("alpha,beta", [(0.5, -2.0), (1.0, 0.0), (-1.0, 0.25), (2.0, 0.0),
                                        (1.0, 1.0)])
def test_dgemm_dev_lds_path_alpha_beta(dev, alpha, beta):
    """The direct-to-LDS kernel (2048 x 4096 x 264: 512 full tiles) with
    alpha = +-1 (C in the accumulators, scaled by beta) and other alpha (C
    read after the products), beta = 0 included; C holds NaN where beta = 0
    must not read it."""
    from charmlu_amd import _lib
    M, N, K = 2048, 4096, 264
    rng = np.random.default_rng(int(alpha * 8 + beta * 16) + 99)
    A = rng.uniform(-0.5, 0.5, (M, K))
    B = rng.uniform(-0.5, 0.5, (K, N))
    C = rng.uniform(-0.5, 0.5, (M, N))
    if beta == 0.0:
        C[::7, ::5] = np.nan
    ref = alpha * A @ B + (beta * C if beta != 0.0 else 0.0)
    dA, dB, dC = _t(A), _t(B), _t(C)
    rc = _lib.load().charmlu_dgemm_dev(_lib.NO_TRANS, _lib.NO_TRANS, M, N, K, alpha, _p(dA), K,
                                       _p(dB), N, beta, _p(dC), N, None)
    assert rc == 0
    _sync()
    got = dC.cpu().numpy()
    Cz = np.nan_to_num(C) if beta == 0.0 else C
    assert np.all(np.abs(got - ref) <= _gemm_tol(A, B, Cz, alpha, beta))


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (1000, 16, 16), (4097, 32, 33), (777, 64, 40),
                                   (20, 300, 64), (300, 200, 256), (1030, 1100, 100),
                                   # >= 256 full 128x128 tiles, K a multiple of 16: the
                                   # direct-to-LDS FP32 path (NN only)
                                   (2048, 2048, 512), (1024, 4096, 1024), (2048, 4096, 272)])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_sgemm_dev(dev, M, N, K, ta, tb):
    """charmlu_sgemm_dev on every tile shape it dispatches to (N <= 32,
    N <= 64, M <= 32, small and large grids), against a float64 product of
    the same FP32 inputs: |C - C_ref| <= 8 K u32 (|A||B| + |C|)."""
    from charmlu_amd import _lib
    rng = np.random.default_rng(M * 5 + N * 11 + K)
    A = rng.uniform(-0.5, 0.5, (K, M) if ta else (M, K)).astype(np.float32)
    B = rng.uniform(-0.5, 0.5, (N, K) if tb else (K, N)).astype(np.float32)
    C = rng.uniform(-0.5, 0.5, (M, N)).astype(np.float32)
    alpha, beta = -1.0, 1.0
    opA = (A.T if ta else A).astype(np.float64)
    opB = (B.T if tb else B).astype(np.float64)
    ref = alpha * opA @ opB + beta * C.astype(np.float64)
    dA, dB, dC = _t(A), _t(B), _t(C)
    rc = _lib.load().charmlu_sgemm_dev(_lib.TRANS if ta else _lib.NO_TRANS,
                                       _lib.TRANS if tb else _lib.NO_TRANS, M, N, K, alpha, _p(dA),
                                       A.shape[1], _p(dB), B.shape[1], beta, _p(dC), N, None)
    assert rc == 0
    _sync()
    got = dC.cpu().numpy().astype(np.float64)
    u32 = float(np.finfo(np.float32).eps)
    tol = 8 * max(K, 1) * u32 * (np.abs(opA) @ np.abs(opB) + np.abs(C)) + 1e-30
    assert np.all(np.abs(got - ref) <= tol)


@pytest.mark.parametrize("alpha,beta,K", [(0.5, -2.0, 528), (1.0, 0.0, 528), (-1.0, 0.25, 528),
                                          (2.0, 0.0, 528), (0.5, -2.0, 272), (1.0, 0.0, 272)])
def test_sgemm_dev_lds_path_alpha_beta(dev, alpha, beta, K):
    """The FP32 direct-to-LDS kernel (2048 x 4096 x K: four stages for K >=
    512, three below) with alpha = +-1 (C in the accumulators) and other
    alpha (C read after the products), beta = 0 included with NaN in C;
    bound 8 K u32 (|alpha| |A||B| + |beta||C|)."""
    from charmlu_amd import _lib
    M, N = 2048, 4096
    rng = np.random.default_rng(int(alpha * 8 + beta * 16) + 107)
    A = rng.uniform(-0.5, 0.5, (M, K)).astype(np.float32)
    B = rng.uniform(-0.5, 0.5, (K, N)).astype(np.float32)
    C = rng.uniform(-0.5, 0.5, (M, N)).astype(np.float32)
    if beta == 0.0:
        C[::7, ::5] = np.nan
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    C64 = np.nan_to_num(C.astype(np.float64)) if beta == 0.0 else C.astype(np.float64)
    ref = alpha * A64 @ B64 + beta * C64
    dA, dB, dC = _t(A), _t(B), _t(C)
    rc = _lib.load().charmlu_sgemm_dev(_lib.NO_TRANS, _lib.NO_TRANS, M, N, K, alpha, _p(dA), K,
                                       _p(dB), N, beta, _p(dC), N, None)
    assert rc == 0
    _sync()
    got = dC.cpu().numpy().astype(np.float64)
    u32 = float(np.finfo(np.float32).eps)
    tol = 8 * K * u32 * (abs(alpha) * np.abs(A64) @ np.abs(B64) + abs(beta) * np.abs(C64)) + 1e-30
    assert np.all(np.abs(got - ref) <= tol)


def test_dgemm_beta_zero_ignores_nan(dev):
    from charmlu_amd import _lib
    A = np.ones((20, 8)); B = np.ones((8, 12)); C = np.full((20, 12), np.nan)
    dA, dB, dC = _t(A), _t(B), _t(C)
    assert _lib.load().charmlu_dgemm_dev(_lib.NO_TRANS, _lib.NO_TRANS, 20, 12, 8, 2.0, _p(dA), 8,
                                         _p(dB), 12, 0.0, _p(dC), 12, None) == 0
    _sync()
    assert np.array_equal(dC.cpu().numpy(), np.full((20, 12), 16.0))


def test_cblas_shims_match_oracle_fixtures(dev):
    from charmlu_amd import _lib
    lib = _lib.load()
    g = np.load(GOLDEN)
    dp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    A, B, C = g["gemm_A"], g["gemm_B"], g["gemm_C"].copy()
    M, K = A.shape; N = B.shape[1]
    assert lib.charmlu_cblas_dgemm(_lib.ROW_MAJOR, _lib.NO_TRANS, _lib.NO_TRANS, M, N, K, -1.0,
                                   dp(A), K, dp(B), N, 1.0, dp(C), N) == 0
    assert np.allclose(C, g["gemm_out"], rtol=0, atol=1e-13)
    # the ESSL/ACML form: column-major with the operands swapped (lu.C:30-34)
    C2 = g["gemm_C"].copy()
    assert lib.charmlu_cblas_dgemm(_lib.COL_MAJOR, _lib.NO_TRANS, _lib.NO_TRANS, N, M, K, -1.0,
                                   dp(B), N, dp(A), K, 1.0, dp(C2), N) == 0
    assert np.allclose(C2, g["gemm_out"], rtol=0, atol=1e-13)
    L, X = g["trsm_L"], g["trsm_B"].copy()
    assert lib.charmlu_cblas_dtrsm(_lib.ROW_MAJOR, _lib.LEFT, _lib.LOWER, _lib.NO_TRANS, _lib.UNIT,
                                   X.shape[0], X.shape[1], 1.0, dp(L), L.shape[1], dp(X),
                                   X.shape[1]) == 0
    assert np.allclose(X, g["trsm_out"], rtol=1e-12, atol=1e-12)
    # ESSL form: column-major, Right, Upper (lu.C:185)
    X2 = g["trsm_B"].copy()
    assert lib.charmlu_cblas_dtrsm(_lib.COL_MAJOR, _lib.RIGHT, _lib.UPPER, _lib.NO_TRANS,
                                   _lib.UNIT, X2.shape[1], X2.shape[0], 1.0, dp(L), L.shape[1],
                                   dp(X2), X2.shape[1]) == 0
    assert np.allclose(X2, g["trsm_out"], rtol=1e-12, atol=1e-12)
    x, y = g["ger_x"], g["ger_y"]
    C3 = g["gemm_C"].copy()
    assert lib.charmlu_cblas_dger(_lib.ROW_MAJOR, M, N, -1.0, dp(x), 1, dp(y), 1, dp(C3), N) == 0
    assert np.allclose(C3, g["ger_out"], rtol=0, atol=1e-15)
    v = g["gemv_x"]
    yv = np.zeros(M)
    Cg = g["gemm_C"]
    assert lib.charmlu_cblas_dgemv(_lib.ROW_MAJOR, _lib.NO_TRANS, M, N, 1.0, dp(Cg), N, dp(v), 1,
                                   0.0, dp(yv), 1) == 0
    assert np.allclose(yv, g["gemv_out"], rtol=0, atol=1e-14)


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("unit", [False, True])
@pytest.mark.parametrize("M", [5, 256, 600])
def test_dtrsm_all_modes(dev, upper, trans, unit, M):
    from charmlu_amd import _lib
    rng = np.random.default_rng(M + 2 * upper + 4 * trans + 8 * unit)
    T = rng.uniform(-0.5, 0.5, (M, M)) / M + np.eye(M) * (1.0 + rng.uniform(0, 1, M))
    Bm = rng.uniform(-0.5, 0.5, (M, 70))
    Tri = np.triu(T) if upper else np.tril(T)
    if unit:
        Tri = Tri - np.diag(np.diag(Tri)) + np.eye(M)
    op = Tri.T if trans else Tri
    ref = np.linalg.solve(op, 2.5 * Bm)
    X = Bm.copy()
    dp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = _lib.load().charmlu_cblas_dtrsm(_lib.ROW_MAJOR, _lib.LEFT,
                                         _lib.UPPER if upper else _lib.LOWER,
                                         _lib.TRANS if trans else _lib.NO_TRANS,
                                         _lib.UNIT if unit else _lib.NON_UNIT, M, 70, 2.5, dp(T), M,
                                         dp(X), 70)
    assert rc == 0
    assert np.allclose(X, ref, rtol=1e-11, atol=1e-11)


def test_dger_negative_increment(dev):
    from charmlu_amd import _lib
    dp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rng = np.random.default_rng(5)
    A = rng.standard_normal((6, 9)); x = rng.standard_normal(12); y = rng.standard_normal(9)
    ref = A + 0.5 * np.outer(x[::2][::-1], y)
    assert _lib.load().charmlu_cblas_dger(_lib.ROW_MAJOR, 6, 9, 0.5, dp(x), -2, dp(y), 1, dp(A),
                                          9) == 0
    assert np.allclose(A, ref, atol=1e-14)


def test_dgemv_trans_and_colmajor(dev):
    from charmlu_amd import _lib
    dp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rng = np.random.default_rng(6)
    A = rng.standard_normal((37, 23)); x = rng.standard_normal(37); y = rng.standard_normal(23)
    ref = 2.0 * A.T @ x + 0.5 * y
    assert _lib.load().charmlu_cblas_dgemv(_lib.ROW_MAJOR, _lib.TRANS, 37, 23, 2.0, dp(A), 23,
                                           dp(x), 1, 0.5, dp(y), 1) == 0
    assert np.allclose(y, ref, atol=1e-12)
    # column-major 23 x 37 no-trans == row-major 37 x 23 trans
    y2 = np.zeros(23)
    assert _lib.load().charmlu_cblas_dgemv(_lib.COL_MAJOR, _lib.NO_TRANS, 23, 37, 1.0, dp(A), 23,
                                           dp(x), 1, 0.0, dp(y2), 1) == 0
    assert np.allclose(y2, A.T @ x, atol=1e-12)
    # even shapes take the two-columns-per-lane 16-byte path: a partial last
    # group of 128 columns, rows not a multiple of the 32-row unroll, and a
    # negative x increment
    for M, N, incx in [(203, 262, 1), (1100, 128, -1), (64, 2, 1)]:
        A = rng.standard_normal((M, N)); x = rng.standard_normal(M); y = rng.standard_normal(N)
        xs = x[::-1] if incx < 0 else x
        ref = -1.5 * A.T @ xs + 2.0 * y
        assert _lib.load().charmlu_cblas_dgemv(_lib.ROW_MAJOR, _lib.TRANS, M, N, -1.5, dp(A), N,
                                               dp(x), incx, 2.0, dp(y), 1) == 0
        assert np.allclose(y, ref, atol=1e-12), (M, N, incx)


# ---------------------------------------------------------------------------
# whole factorization + solve vs the oracle
# ---------------------------------------------------------------------------
def check_pivots(gp, op, ties, what, ulps=4.0):
    """Pivots index-for-index with the oracle's.  The first difference is
    accepted only at a near-tie: the device's row is the oracle's runner-up
    at that column and their magnitudes (the oracle's values) lie within
    ``ulps`` ulp; it is logged, and the columns after it are not compared
    (a flipped pivot reorders every later row).  ``ties``: the oracle's
    near-tie log, as a dict of full per-column arrays or of the close
    columns only (cols, row2, mag1, mag2).  Also logs every column the
    oracle itself saw as a near-tie.  Returns the number of columns
    compared."""
    gp = np.asarray(gp)
    op = np.asarray(op)
    if "cols" in ties:
        cols = np.asarray(ties["cols"])
        close = [int(c) for c, r2, m1, m2 in zip(cols, ties["row2"], ties["mag1"], ties["mag2"])
                 if m1 - m2 <= ulps * np.spacing(m1)]
    else:
        close = [int(c) for c in np.nonzero((ties["row2"] >= 0) & (
            ties["mag1"] - ties["mag2"] <= ulps * np.spacing(ties["mag1"])))[0]]
    for c in close:
        print(f"{what}: oracle near-tie within {ulps:g} ulp at column {c}")
    mism = np.nonzero(gp != op)[0]
    if len(mism) == 0:
        return len(gp)
    j = int(mism[0])
    if "cols" in ties:
        k = np.nonzero(np.asarray(ties["cols"]) == j)[0]
        ok = len(k) == 1 and int(ties["row2"][k[0]]) == int(gp[j]) and \
            ties["mag1"][k[0]] - ties["mag2"][k[0]] <= ulps * np.spacing(ties["mag1"][k[0]])
    else:
        ok = int(ties["row2"][j]) == int(gp[j]) and \
            ties["mag1"][j] - ties["mag2"][j] <= ulps * np.spacing(ties["mag1"][j])
    if not ok:
        pytest.fail(f"{what}: pivot of column {j} is row {gp[j]}, the oracle's {op[j]} "
                    "(not a near-tie)")
    print(f"{what}: near-tie flip at column {j} (device row {gp[j]}, oracle row {op[j]}); "
          "later columns not compared")
    return j


def _compare_run(dev, orc, n, b, batch=0, kind="uniform", nthreads=8):
    import torch
    A = dev.gen_matrix(n, seed=orc.SEED_A, kind=kind)
    rhs = dev.gen_rhs(n)
    LU = A.clone()
    ipiv, info = dev.getrf(LU, b)
    x = rhs.clone()
    dev.getrs(LU, ipiv, x)
    res = dev.residual(A, x, rhs)
    _sync()
    A_h = A.cpu().numpy()
    o = orc.lu_solve_problem(n, b, batch, nthreads=nthreads, A=A_h.copy(), ties=True)
    check_pivots(ipiv.cpu().numpy(), o["ipiv"], o["ties"], f"n={n} b={b}")
    xg = x.cpu().numpy()
    rel = np.abs(xg - o["x"]).max() / np.abs(o["x"]).max()
    return dict(info=int(info.item()), oinfo=o["info"], rel=rel, res=res, ores=o["residual"],
                LU=LU, o=o)


# the reference's regression grid (test_lu.pl:9-52) with its large cases
# (4096, 256) and (4096, 512); (4096, 128) is C1, below
GRID = [(48, 48), (48, 24), (48, 16), (48, 12), (32, 8), (32, 4), (512, 128), (512, 64),
        (512, 256), (1024, 256), (1024, 128), (1024, 64), (2048, 256), (2048, 512),
        (4096, 256), (4096, 512)]


@pytest.mark.parametrize("n,b", GRID)
def test_getrf_grid_matches_oracle(dev, orc, n, b):
    r = _compare_run(dev, orc, n, b)
    assert r["info"] == r["oinfo"] == 0
    assert r["rel"] <= 1e-10, r["rel"]
    assert r["res"]["residual"] < 16
    # the oracle's residual formula gives the same on the GPU result
    assert abs(r["res"]["residual"] - r["ores"]) < 16


def test_getrf_matches_golden_fixture_pivots(dev, orc):
    g = np.load(GOLDEN)
    for (n, b, batch) in [(48, 12, 0), (32, 8, 9), (512, 128, 128)]:
        A = dev.gen_matrix(n, seed=orc.SEED_A)
        ipiv, info = dev.getrf(A, b)
        _sync()
        assert np.array_equal(ipiv.cpu().numpy(), g[f"n{n}_b{b}_k{batch}_ipiv"])


def test_small_block_panel_bitwise(dev, orc):
    # with b <= 16 every panel is one 16-column strip factored exactly as
    # the oracle's loop (division, same per-element FMA); only the trailing
    # GEMM's summation order differs
    r = _compare_run(dev, orc, 48, 16)
    LU = r["LU"].cpu().numpy()
    assert np.array_equal(LU[:16, :16], r["o"]["LU"][:16, :16])


def test_c1_n4096_matches_oracle(dev, orc):
    r = _compare_run(dev, orc, 4096, 128)  # C1: n=4096, b=128
    assert r["info"] == 0 and r["rel"] <= 1e-10 and r["res"]["residual"] < 16


def test_singular_info(dev, orc):
    import torch
    A = dev.gen_matrix(64, seed=5)
    A[:, 37] = 0.0
    _, info = dev.getrf(A, 16)
    Ah = orc.gen_uniform(5, 64)
    Ah[:, 37] = 0.0
    oinfo, _ = orc.getrf(Ah, 16)
    assert int(info.item()) == oinfo == 38
    B = dev.gen_matrix(64, seed=6)
    B[10] = B[3]
    _, info2 = dev.getrf(B, 16)
    assert int(info2.item()) > 0
    assert not torch.isnan(A).any()


def test_getrs_transpose_and_multi_rhs(dev, orc):
    import torch
    n = 700
    A = dev.gen_matrix(n, seed=21)
    Ah = A.cpu().numpy()
    LU = A.clone()
    ipiv, info = dev.getrf(LU, 100)
    rng = np.random.default_rng(2)
    Bh = rng.standard_normal((n, 64))
    X = _t(Bh)
    dev.getrs(LU, ipiv, X)
    assert np.allclose(X.cpu().numpy(), np.linalg.solve(Ah, Bh), rtol=1e-9, atol=1e-9)
    Xt = _t(Bh)
    dev.getrs(LU, ipiv, Xt, trans=True)
    assert np.allclose(Xt.cpu().numpy(), np.linalg.solve(Ah.T, Bh), rtol=1e-9, atol=1e-9)
    bt = _t(Bh[:, 0].copy())
    dev.getrs(LU, ipiv, bt, trans=True)
    assert np.allclose(bt.cpu().numpy(), np.linalg.solve(Ah.T, Bh[:, 0]), rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("n", [1000, 999, 37])
def test_residual_kernel_matches_oracle(dev, orc, n):
    A = dev.gen_matrix(n, seed=3)
    x = dev.gen_rhs(n, seed=17)
    b = dev.gen_rhs(n, seed=19)
    r = dev.residual(A, x, b)
    ref, ns, m4 = orc.residual(A.cpu().numpy(), x.cpu().numpy(), b.cpu().numpy())
    assert np.isclose(r["residual"], ref, rtol=1e-10)
    assert np.isclose(r["residual_ns"], ns, rtol=1e-10)
    assert r["a_max"] == m4[0] and r["b_max"] == m4[1] and r["x_max"] == m4[2]


def _fixture_run(dev, n, b):
    """Factor + solve the seeded problem on the device and compare it with
    the oracle's pivots and x stored in the large fixtures (written on the
    CPU by tests/golden/make_golden.py --large)."""
    g = np.load(LARGE)
    key = f"n{n}_b{b}"
    out = dev.run(dev.LUConfig(n, b), verbose=True)
    assert out["info"] == 0 == int(g[key + "_info"][0])
    ties = dict(cols=g[key + "_tie_cols"], row2=g[key + "_tie_row2"], mag1=g[key + "_tie_mag1"],
                mag2=g[key + "_tie_mag2"])
    ncmp = check_pivots(out["ipiv"].cpu().numpy(), g[key + "_ipiv"], ties, key)
    xo = g[key + "_x"]
    rel = float(np.abs(out["x"].cpu().numpy() - xo).max() / np.abs(xo).max())
    print(f"{key}: {ncmp} pivots compared, x vs oracle {rel:.2e}, residual "
          f"{out['residual']:.3f} (oracle {g[key + '_res'][0]:.3f}), "
          f"{out['tflops_eff']:.2f} TFLOP/s")
    assert rel <= 1e-10
    assert out["residual"] < 16 and out["residual_ns"] < 16
    return out


def test_c2_n32768_pivots_x_residual_vs_oracle(dev):
    """C2 at full size (n = 32768, block 256): all 32768 pivots index for
    index with the oracle's (a difference only at a logged near-tie within
    4 ulp), x within 1e-10 of the oracle's, the reference's residual < 16."""
    _fixture_run(dev, 32768, 256)


def test_fp64_block1024_pivoting_matrix_vs_oracle(dev):
    """Block 1024 (the FP64 comparator of the mixed form, whose U row uses
    an explicit inverse of the 1024 x 1024 L11) on the uniform, pivoting
    matrix at n = 8192: pivots and x against the oracle at the same block."""
    _fixture_run(dev, 8192, 1024)


def _subprocess(code, env_extra, out_path, timeout=300):
    import subprocess
    import sys
    env = dict(os.environ, **env_extra)
    print(f"subprocess with {env_extra}", flush=True)
    r = subprocess.run([sys.executable, "-c", code, out_path], env=env, cwd=ROOT,
                       capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-3000:]
    return np.load(out_path)


def test_panel_cus_16_fp64_matches_oracle(dev, orc, tmp_path):
    """16 panel CUs (16 workgroups resident per strip) at n = 9216: the
    strips that would need 18 workgroups run 16 columns wide instead; the
    pivots and x match the oracle."""
    code = ("import sys, numpy as np, torch, charmlu_amd as c\n"
            "n, b = 9216, 256\n"
            "A = c.gen_matrix(n); rhs = c.gen_rhs(n)\n"
            "ip, info = c.getrf(A, b); x = rhs.clone(); c.getrs(A, ip, x)\n"
            "torch.cuda.synchronize()\n"
            "np.savez(sys.argv[1], ip=ip.cpu().numpy(), x=x.cpu().numpy(), info=[int(info.item())])\n")
    z = _subprocess(code, {"CHARMLU_PANEL_CUS": "16"}, str(tmp_path / "c16.npz"))
    o = orc.lu_solve_problem(9216, 256, nthreads=16, ties=True)
    assert int(z["info"][0]) == 0 == o["info"]
    check_pivots(z["ip"], o["ipiv"], o["ties"], "16 panel CUs, n=9216")
    assert np.abs(z["x"] - o["x"]).max() / np.abs(o["x"]).max() <= 1e-10


def _fp64_vs_default(dev, z, n, b):
    """The subprocess's FP64 pivots and x against this process's (default
    layout, 32 panel CUs) run of the same problem."""
    import torch
    A = dev.gen_matrix(n)
    rhs = dev.gen_rhs(n)
    LU = A.clone()
    ip, info = dev.getrf(LU, b)
    x = rhs.clone()
    dev.getrs(LU, ip, x)
    res = dev.residual(A, torch.from_numpy(z["x"]).cuda(), rhs)
    _sync()
    assert np.array_equal(z["ip"], ip.cpu().numpy())
    xr = x.cpu().numpy()
    assert np.abs(z["x"] - xr).max() / np.abs(xr).max() <= 1e-10
    assert res["residual"] < 16


_LAYOUT_CODE = ("import sys, numpy as np, torch, charmlu_amd as c\n"
                "n, b, nm = int(sys.argv[2]), int(sys.argv[3]), int(sys.argv[4])\n"
                "A = c.gen_matrix(n); rhs = c.gen_rhs(n)\n"
                "ip, info = c.getrf(A, b); x = rhs.clone(); c.getrs(A, ip, x)\n"
                "del A\n"
                "D = c.gen_matrix(nm, kind='diagdom'); bd = c.gen_rhs(nm)\n"
                "S = c.lag2s(D); ips, infos = c.getrf_mixed(S, 1024, 'bf16')\n"
                "xd, st = c.gesv_ir(D, S, ips, bd, tol=1.0)\n"
                "torch.cuda.synchronize()\n"
                "np.savez(sys.argv[1], ip=ip.cpu().numpy(), x=x.cpu().numpy(),\n"
                "         info=[int(info.item()), int(infos.item())], ips=ips.cpu().numpy(),\n"
                "         res=[st['residual']])\n")


def _layout_run(tmp_path, cus, n, b, nm):
    import subprocess
    import sys
    out = str(tmp_path / f"cus{cus}.npz")
    env = dict(os.environ, CHARMLU_PANEL_CUS=str(cus))
    print(f"subprocess: CHARMLU_PANEL_CUS={cus}, FP64 n={n} b={b}, mixed BF16 n={nm}", flush=True)
    r = subprocess.run([sys.executable, "-c", _LAYOUT_CODE, out, str(n), str(b), str(nm)],
                       env=env, cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-3000:]
    z = np.load(out)
    assert list(z["info"]) == [0, 0]
    assert np.array_equal(z["ips"], np.arange(nm, dtype=np.int32))
    assert z["res"][0] <= 1.0
    return z


def test_panel_cus_8_uses_8_column_strips(dev, tmp_path):
    """8 panel CUs (8 resident workgroups): FP64 strips taller than 16384
    rows and FP32 strips taller than 32768 rows only fit as 8-column strips
    (8 rows of 8 doubles or 16 rows of 8 floats per thread).  FP64
    n = 24576: pivots and x as in the default layout and residual < 16;
    mixed BF16 n = 40960 (diagonally dominant): identity pivots and FP64
    refinement to the reference's residual <= 1."""
    z = _layout_run(tmp_path, 8, 24576, 256, 40960)
    _fp64_vs_default(dev, z, 24576, 256)


def test_panel_cus_48_stays_within_resident_workgroups(dev, tmp_path):
    """48 panel CUs: a 48-CU mask keeps 32 one-per-CU workgroups resident,
    not 48 (1 or 2 CUs per shader engine, workgroups dealt round-robin over
    the engines), so strips stay at 32 workgroups.  (Strips of 33 to 36
    workgroups on such a mask made a mixed BF16 run stop progressing: a
    workgroup that is not resident leaves the others polling for its
    record.)  FP64 n = 32768 against the default layout, mixed BF16
    n = 65536 refined to residual <= 1."""
    z = _layout_run(tmp_path, 48, 32768, 256, 65536)
    _fp64_vs_default(dev, z, 32768, 256)


def test_fp64_beyond_65536_rows_uses_8_column_strips(dev):
    """n = 73728: the first panels are taller than 32 workgroups of 2048
    rows, so they run as 8-column strips of 4096 rows per workgroup; no
    oracle at this size, so the reference's pass criterion (residual < 16)
    and the normwise form."""
    out = dev.run(dev.LUConfig(73728, 256), verbose=True)
    assert out["info"] == 0
    assert out["residual"] < 16 and out["residual_ns"] < 16


def test_too_tall_first_panel_is_rejected_before_any_work(dev):
    """n = 131328 FP64 (first strip 131328 rows: more than 32 workgroups of
    4096 rows): CHARMLU_ERR_UNSUPPORTED with a message, before any kernel
    touches A or ipiv (so small stand-in buffers are safe here), and the
    info word is 0."""
    import torch
    from charmlu_amd import _lib
    lib = _lib.load()
    A = torch.zeros(16, dtype=torch.float64, device="cuda")
    ip = torch.zeros(16, dtype=torch.int32, device="cuda")
    info = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    n = 131328
    rc = lib.charmlu_dgetrf(n, 256, _p(A), n, _p(ip), _p(info), None)
    _sync()
    assert rc == _lib.ERR_UNSUPPORTED
    assert "resident" in lib.charmlu_last_error().decode()
    assert int(info.item()) == 0 and int(ip.abs().sum().item()) == 0


def test_tall_panel_strips_residual(dev):
    # n = 40960: the first panels are taller than 32768 rows, where the strip
    # kernel uses 512-thread workgroups of 2048 rows (20 of them) instead of
    # 256-thread ones; no oracle at this size, so the reference's pass
    # criterion (residual < 16) and the normwise form
    out = dev.run(dev.LUConfig(40960, 256), verbose=True)
    assert out["info"] == 0
    assert out["residual"] < 16 and out["residual_ns"] < 16


@pytest.mark.parametrize("n,b", [(1024, 128), (2048, 256)])
def test_forced_split_matches_oracle_and_unsplit(dev, orc, n, b):
    """Small problems never split a trailing update (it fits beside the
    panel); forced to half the rows at every step, the factorization is
    bit-identical to the unsplit one and still matches the oracle."""
    import torch
    from charmlu_amd import _lib
    lib = _lib.load()
    prof = (ctypes.c_double * 20)()
    A = dev.gen_matrix(n, seed=orc.SEED_A)
    LU1 = A.clone()
    lib.charmlu_profile(1)
    ip1, _ = dev.getrf(LU1, b)
    assert lib.charmlu_profile_read(prof, 20) == 0
    launches_unsplit = prof[0]  # trailing-update launches of the 128x128 kernel
    LU2 = A.clone()
    os.environ["CHARMLU_FORCE_SPLIT"] = "1"
    try:
        ip2, info = dev.getrf(LU2, b)
        _sync()
        assert lib.charmlu_profile_read(prof, 20) == 0
    finally:
        os.environ.pop("CHARMLU_FORCE_SPLIT", None)
        lib.charmlu_profile(0)
    assert prof[0] > launches_unsplit, (prof[0], launches_unsplit)
    assert int(info.item()) == 0
    assert np.array_equal(ip1.cpu().numpy(), ip2.cpu().numpy())
    assert torch.equal(LU1, LU2)
    o = orc.lu_solve_problem(n, b, 0, nthreads=8, A=A.cpu().numpy().copy())
    assert np.array_equal(ip2.cpu().numpy(), o["ipiv"])


@pytest.mark.parametrize("value", ["72", "4", "20", "-1", "x"])
def test_panel_cus_out_of_range_is_an_error(value):
    """CHARMLU_PANEL_CUS takes 0 or a multiple of 8 (the XCDs) up to 64: any
    other value makes the first call on a device fail with
    CHARMLU_ERR_CONFIG and a message, before any stream or workspace is made
    (a fresh process: the workspace is made once per device)."""
    import subprocess
    import sys
    code = ("import torch, charmlu_amd as c\n"
            "A = torch.eye(64, dtype=torch.float64, device='cuda')\n"
            "try:\n"
            "    c.getrf(A, 32)\n"
            "except c.CharmluError as e:\n"
            "    print('ERR', e)\n")
    env = dict(os.environ, CHARMLU_PANEL_CUS=value)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], env=env, cwd=root, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "ERR" in r.stdout and "-102" in r.stdout and "CHARMLU_PANEL_CUS" in r.stdout, r.stdout


def test_panel_cus_zero_is_bitwise_the_default(dev, orc, tmp_path):
    """CHARMLU_PANEL_CUS=0 puts every stream on all CUs (no CU masks, no
    split update, no in-panel GEMMs on a second stream): the factors are bit
    for bit those of the default layout, for FP64 and the mixed BF16 form
    (a fresh process: the streams are made once per device)."""
    import subprocess
    import sys
    code = ("import sys, numpy as np, torch, charmlu_amd as c\n"
            "A = c.gen_matrix(4096, seed=%d)\n"
            "ip, info = c.getrf(A, 256)\n"
            "S = c.gen_matrix(8192, seed=5).float()\n"
            "ips, infos = c.getrf_mixed(S, 1024, 'bf16')\n"
            "torch.cuda.synchronize()\n"
            "np.savez(sys.argv[1], A=A.cpu().numpy(), ip=ip.cpu().numpy(), S=S.cpu().numpy(),\n"
            "         ips=ips.cpu().numpy(), info=[int(info.item()), int(infos.item())])\n"
            % orc.SEED_A)
    out = str(tmp_path / "zero.npz")
    env = dict(os.environ, CHARMLU_PANEL_CUS="0")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code, out], env=env, cwd=root, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    z = np.load(out)
    A = dev.gen_matrix(4096, seed=orc.SEED_A)
    ip, info = dev.getrf(A, 256)
    S = dev.gen_matrix(8192, seed=5).float()
    ips, infos = dev.getrf_mixed(S, 1024, "bf16")
    _sync()
    assert list(z["info"]) == [int(info.item()), int(infos.item())] == [0, 0]
    assert np.array_equal(z["ip"], ip.cpu().numpy())
    assert np.array_equal(z["A"], A.cpu().numpy())
    assert np.array_equal(z["ips"], ips.cpu().numpy())
    assert np.array_equal(z["S"], S.cpu().numpy())


def _probe(dev, groups, sleeps=400):
    import torch
    from charmlu_amd import _lib
    out = torch.zeros(4 * groups, dtype=torch.int64, device="cuda")
    rc = _lib.load().charmlu_panel_probe(groups, sleeps, _p(out), None)
    assert rc == 0
    _sync()
    h = out.cpu().numpy().reshape(groups, 4).astype(np.uint64)
    t0, t1 = h[:, 2], h[:, 3]
    late = int((t0 >= t1.min()).sum())
    return late, h


def test_panel_stream_holds_the_resident_workgroups(dev):
    """The capacity rule the strip shapes are planned with, on this device:
    the panel stream (32 CUs, the default) holds as many one-per-CU
    workgroups at once as charmlu_panel_plan says (every one starts before
    any ends), and not 4 more (some start only after another one ended:
    workgroups are dealt round-robin to the XCDs and their shader engines,
    whatever the free CUs).  Each workgroup sleeps about 1.4 ms."""
    import ctypes
    from charmlu_amd import _lib
    out = (ctypes.c_int * 6)()
    assert _lib.load().charmlu_panel_plan(_lib.PREC_FP64, 1024, 16, 32, 256, 1, out) == 0
    resident = out[4]
    assert resident == 32
    late, h = _probe(dev, resident)
    print(f"{resident} workgroups: {late} started after the first end; XCCs "
          f"{sorted(set(int(x) for x in h[:, 0]))}")
    assert late == 0
    late, _ = _probe(dev, resident + 4)
    print(f"{resident + 4} workgroups: {late} started after the first end")
    assert late > 0


@pytest.mark.parametrize("cus", [16, 24, 48])
def test_panel_stream_resident_workgroups_other_masks(dev, tmp_path, cus):
    """The same check on other panel CU counts (a fresh process each: the
    streams are made once per device): 16 and 24 CUs hold 16 and 24
    one-per-CU workgroups, 48 CUs (1 or 2 per shader engine) only 32."""
    import subprocess
    import sys
    code = ("import sys, ctypes, numpy as np, torch, charmlu_amd as c\n"
            "from charmlu_amd import _lib\n"
            "L = _lib.load(); o = (ctypes.c_int * 6)()\n"
            "assert L.charmlu_panel_plan(0, 1024, 16, int(sys.argv[2]), 256, 1, o) == 0\n"
            "res = []\n"
            "for g in (o[4], o[4] + 4):\n"
            "    out = torch.zeros(4 * g, dtype=torch.int64, device='cuda')\n"
            "    assert L.charmlu_panel_probe(g, 400, ctypes.c_void_p(out.data_ptr()), None) == 0\n"
            "    torch.cuda.synchronize()\n"
            "    h = out.cpu().numpy().reshape(g, 4).astype(np.uint64)\n"
            "    res.append([g, int((h[:, 2] >= h[:, 3].min()).sum())])\n"
            "np.save(sys.argv[1], np.array(res))\n")
    out = str(tmp_path / "probe.npy")
    env = dict(os.environ, CHARMLU_PANEL_CUS=str(cus))
    print(f"subprocess: CHARMLU_PANEL_CUS={cus}", flush=True)
    r = subprocess.run([sys.executable, "-c", code, out, str(cus)], env=env, cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-3000:]
    (g0, late0), (g1, late1) = np.load(out).tolist()
    print(f"{cus} panel CUs: {g0} workgroups, {late0} late; {g1} workgroups, {late1} late")
    assert g0 == {16: 16, 24: 24, 48: 32}[cus]
    assert late0 == 0 and late1 > 0


def test_run_prints_the_reference_result_lines(dev, capsys):
    """lu.run, the reference driver's run on the device: its result lines
    (benchmark.C:343-349, 374-381, 385-400) with the block-dgemm check
    (testdgemm, benchmark.C:297-341) and the residual under 16."""
    from charmlu_amd import lu
    cfg = dev.LUConfig.from_argv(["1024", "128", "100"])
    out = lu.run(cfg)
    text = capsys.readouterr().out
    assert "RESULT procs: 1 device\tblock size: 128" in text
    assert "The dgemm 128 x 128 takes" in text and "GFlop/sec" in text
    assert "Percent of DGEMM is:" in text
    assert out["dgemm_gflops"] > 0 and out["percent_of_dgemm"] > 0
    assert out["residual"] < 16 and out["info"] == 0
