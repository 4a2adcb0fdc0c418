"""Pins of the CPU oracle (oracle/) against what the mathematics and public
references fix.  The reference ships no golden vectors (its only test,
test_lu.pl:70-84, checks residual < 16), so each oracle function is tied to
an independent answer: the C++ standard's minstd_rand known answer, the
published splitmix64 outputs, LAPACK (scipy / MKL), brute force, closed
forms.  A plausible slip (wrong Schrage constant, dropped term, transposed
operand, wrong pivot tie rule, cycles applied backwards) fails one of these.
"""
import os

import numpy as np
import pytest
import scipy.linalg as sl

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "oracle_fixtures.npz")


# ---- generators ----------------------------------------------------------------
def test_matgen_minstd_known_answer(orc):
    # C++11 [rand.predef]: the 10000th consecutive invocation of a
    # default-constructed minstd_rand (seed 1) produces 399268537.
    assert orc.matgen_nth_int(1, 10000) == 399268537


def test_matgen_first_draws_of_seed_b(orc):
    # 48271^k * 9934835 mod (2^31-1), computed with Python integers
    M = 2147483647
    expect = [(9934835 * pow(48271, k, M)) % M for k in range(1, 6)]
    got = [orc.matgen_nth_int(orc.SEED_B, k) for k in range(1, 6)]
    assert got == expect == [675567004, 755670389, 1955603124, 1979727425, expect[4]]
    d = orc.matgen(orc.SEED_B, 4)
    assert np.array_equal(d, np.array(expect[:4]) / M - 0.5)


def test_matgen_skip_is_block_offset(orc):
    # genVec: block row x skips x*b draws (benchmark.C:454-460)
    full = orc.matgen(orc.SEED_B, 64)
    for x in range(8):
        assert np.array_equal(orc.matgen(orc.SEED_B, 8, skip=8 * x), full[8 * x:8 * x + 8])


def test_gen_value_published_splitmix64(orc):
    # splitmix64 seeded with 0 yields 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
    # 0x06C45D188009454F; key(seed, 0, 0) = seed * phi is exactly its state
    # after `seed` steps.
    pub = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    for s, z in enumerate(pub):
        assert orc.gen_value(s, 0, 0) == (z >> 11) * 2.0 ** -53 - 0.5


def test_gen_uniform_range_and_moments(orc):
    A = orc.gen_uniform(orc.SEED_A, 512)
    assert A.min() >= -0.5 and A.max() < 0.5
    assert abs(A.mean()) < 0.005 and abs(A.var() - 1 / 12) < 0.002
    # block offsets address the same entries
    assert np.array_equal(orc.gen_uniform(orc.SEED_A, 5, 7, row0=100, col0=200),
                          A[100:105, 200:207])


def test_gen_diagdom(orc):
    A = orc.gen_diagdom(7, 64)
    U = orc.gen_uniform(7, 64)
    off = ~np.eye(64, dtype=bool)
    assert np.array_equal(A[off], U[off])
    assert np.allclose(np.diag(A), np.abs(U).sum(1), rtol=1e-15, atol=0)


# ---- BLAS surface --------------------------------------------------------------
def test_blas_loops_closed_forms(orc):
    rng = np.random.default_rng(0)
    A = rng.standard_normal((7, 5)); B = rng.standard_normal((5, 9)); C = rng.standard_normal((7, 9))
    assert np.allclose(orc.dgemm(-1.0, A, B, 1.0, C.copy()), C - A @ B, rtol=1e-14, atol=1e-14)
    assert np.allclose(orc.dgemm(2.0, A, B, 0.0, C.copy()), 2 * A @ B, rtol=1e-14, atol=1e-14)
    # identity operands
    I = np.eye(5)
    assert np.array_equal(orc.dgemm(1.0, A, I, 0.0, np.zeros((7, 5))), A)
    L = np.tril(rng.standard_normal((6, 6)), -1) + np.eye(6)
    X = rng.standard_normal((6, 4))
    assert np.allclose(L @ orc.dtrsm_llnu(L, X.copy()), X, atol=1e-12)
    # the strictly upper part of L is ignored (unit lower)
    L2 = L + np.triu(rng.standard_normal((6, 6)))
    assert np.array_equal(orc.dtrsm_llnu(L2, X.copy()), orc.dtrsm_llnu(L, X.copy()))
    x = rng.standard_normal(7); y = rng.standard_normal(9)
    assert np.allclose(orc.dger(-1.0, x, y, C.copy()), C - np.outer(x, y), atol=1e-14)
    v = rng.standard_normal(9)
    assert np.allclose(orc.dgemv(1.0, C, v, 0.0, np.zeros(7)), C @ v, atol=1e-13)


def test_blas_loops_vs_mkl(orc):
    import subprocess, sys
    # separate process: switching the oracle to MKL is global
    code = r"""
import numpy as np, oracle
rng = np.random.default_rng(1)
A = rng.standard_normal((33, 17)); B = rng.standard_normal((17, 29)); C = rng.standard_normal((33, 29))
L = np.tril(rng.standard_normal((33, 33)), -1) + np.eye(33); X = rng.standard_normal((33, 12))
x = rng.standard_normal(33); y = rng.standard_normal(29); v = rng.standard_normal(29)
r1 = [oracle.dgemm(-1.0, A, B, 1.0, C.copy()), oracle.dtrsm_llnu(L, X.copy()),
      oracle.dger(-1.0, x, y, C.copy()), oracle.dgemv(1.0, C, v, 0.0, np.zeros(33))]
if not oracle.use_mkl():
    print("NOMKL"); raise SystemExit(0)
r2 = [oracle.dgemm(-1.0, A, B, 1.0, C.copy()), oracle.dtrsm_llnu(L, X.copy()),
      oracle.dger(-1.0, x, y, C.copy()), oracle.dgemv(1.0, C, v, 0.0, np.zeros(33))]
for a, b in zip(r1, r2):
    assert np.allclose(a, b, rtol=1e-13, atol=1e-13), abs(a - b).max()
print("OK")
"""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    if "NOMKL" in out.stdout:
        pytest.skip("MKL not present")
    assert "OK" in out.stdout


# ---- pivots -> cycles ---------------------------------------------------------
def test_cycles_equal_sequential_swaps(orc):
    rng = np.random.default_rng(3)
    for trial in range(50):
        n, k = 40, int(rng.integers(1, 12))
        j0 = int(rng.integers(0, n - k))
        pos = np.arange(j0, j0 + k)
        row = np.array([int(rng.integers(p, n)) for p in pos])
        M = rng.standard_normal((n, 5))
        ref = M.copy()
        for p, r in zip(pos, row):
            ref[[p, r]] = ref[[r, p]]
        nc, si, seq = orc.swaps_to_cycles(pos, row)
        got = orc.apply_cycles(nc, si, seq, M.copy())
        assert np.array_equal(got, ref)
        # every cycle has length >= 2 (lu.C:95)
        assert all(si[c + 1] - si[c] >= 2 for c in range(nc))


# ---- factorization ---------------------------------------------------------------
@pytest.mark.parametrize("n,b,batch", [(48, 48, 0), (48, 12, 0), (32, 8, 3), (32, 4, 1),
                                       (64, 16, 5), (512, 128, 0)])
def test_getrf_pivots_match_lapack(orc, n, b, batch):
    A = orc.gen_uniform(orc.SEED_A, n)
    LU = A.copy()
    info, ipiv = orc.getrf(LU, b, batch)
    assert info == 0
    _, piv = sl.lu_factor(A)  # LAPACK dgetrf (partial pivoting, same tie rule on distinct |a|)
    assert np.array_equal(ipiv, piv)
    # P A = L U to within a backward-stable bound
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = A.copy()
    for j, p in enumerate(ipiv):
        PA[[j, p]] = PA[[p, j]]
    assert np.abs(PA - L @ U).max() <= 64 * n * np.finfo(float).eps * np.abs(A).max()
    # multipliers bounded by 1 (partial pivoting)
    assert np.abs(np.tril(LU, -1)).max() <= 1.0


def test_getrf_matches_unblocked_brute_force(orc):
    for n, b in [(12, 4), (16, 8), (30, 10), (48, 16)]:
        A = orc.gen_uniform(99 + n, n)
        L1 = A.copy()
        _, p1 = orc.getrf(L1, b)
        L2 = A.copy()
        _, p2 = orc.getf2(L2)
        assert np.array_equal(p1, p2)
        assert np.allclose(L1, L2, rtol=0, atol=1e-13)


def _getf2_runner_up(A):
    """Unblocked LU with partial pivoting (division) in numpy, recording per
    column the |pivot| and the best other candidate row and |value|."""
    A = A.copy()
    n = A.shape[0]
    row2 = np.full(n, -1)
    mag1 = np.zeros(n)
    mag2 = np.zeros(n)
    for j in range(n):
        col = np.abs(A[j:, j])
        p = j + int(np.argmax(col))  # first maximum
        mag1[j] = col[p - j]
        rest = col.copy()
        rest[p - j] = -1.0
        if n - j > 1:
            q = int(np.argmax(rest))
            row2[j], mag2[j] = j + q, rest[q]
        A[[j, p]] = A[[p, j]]
        A[j + 1:, j] /= A[j, j]
        A[j + 1:, j + 1:] -= np.outer(A[j + 1:, j], A[j, j + 1:])
    return row2, mag1, mag2


def test_near_tie_log_matches_brute_force(orc):
    """The runner-up candidates the oracle logs per column (the data that
    classifies a device pivot differing from the oracle's as a near-tie)
    are those of an unblocked LU on the same matrix."""
    for n, b in [(48, 16), (64, 8), (96, 32)]:
        A = orc.gen_uniform(200 + n, n)
        t = {}
        orc.getrf(A.copy(), b, ties=t)
        row2, mag1, mag2 = _getf2_runner_up(A)
        assert np.array_equal(t["row2"][:-1], row2[:-1])
        assert np.allclose(t["mag1"], mag1, rtol=1e-12, atol=0)
        assert np.allclose(t["mag2"][:-1], mag2[:-1], rtol=1e-12, atol=0)
        assert t["row2"][-1] == -1  # the last column has one candidate


def test_near_tie_log_flags_an_engineered_tie(orc):
    """Column 0 holds 0.5 at row 3 and the next double below 0.5 at row 9:
    a 1-ulp near-tie, classified as such only for row 9; row 3 wins (the
    larger), and an exact tie keeps the lower row (strict '>')."""
    n = 16
    A = orc.gen_uniform(7, n) * 0.1
    A[3, 0] = 0.5
    A[9, 0] = -np.nextafter(0.5, 0.0)
    t = {}
    _, ipiv = orc.getrf(A.copy(), 8, ties=t)
    assert ipiv[0] == 3
    assert 0 in orc.near_ties(t)
    assert orc.is_near_tie(t, 0, 9) and not orc.is_near_tie(t, 0, 5)
    assert not orc.is_near_tie(t, 0, 9, ulps=0.25)  # the gap is half a spacing of 0.5
    A[9, 0] = -0.5  # exact tie: the first (lower) row is the pivot
    t = {}
    _, ipiv = orc.getrf(A.copy(), 8, ties=t)
    assert ipiv[0] == 3 and t["row2"][0] == 9 and t["mag1"][0] == t["mag2"][0]
    # a clear winner is no near-tie
    A[9, 0] = 0.25
    t = {}
    orc.getrf(A.copy(), 8, ties=t)
    assert 0 not in orc.near_ties(t)


def test_batch_size_does_not_change_result(orc):
    # cycles == sequential swaps, so the pivot batch size (test_lu.pl:16-26)
    # changes no bit of the factorization
    A = orc.gen_uniform(orc.SEED_A, 32)
    base = A.copy()
    _, p0 = orc.getrf(base, 8, 1)
    for batch in range(2, 10):
        LU = A.copy()
        _, p = orc.getrf(LU, 8, batch)
        assert np.array_equal(p, p0) and np.array_equal(LU, base)


def test_threads_do_not_change_result(orc):
    A = orc.gen_uniform(orc.SEED_A, 256)
    L1 = A.copy(); _, p1 = orc.getrf(L1, 32, 0, None, 1)
    L2 = A.copy(); _, p2 = orc.getrf(L2, 32, 0, None, 4)
    assert np.array_equal(p1, p2) and np.array_equal(L1, L2)


def test_singular_and_bad_block(orc):
    A = orc.gen_uniform(5, 32)
    A[:, 7] = 0.0  # zero column: the pivot of column 7 is zero
    info, _ = orc.getrf(A.copy(), 8)
    assert info == 8
    A2 = orc.gen_uniform(5, 32)
    A2[5] = A2[3]  # duplicate rows: exactly singular
    info2, _ = orc.getrf(A2.copy(), 8)
    assert info2 > 0
    assert orc.getrf(orc.gen_uniform(5, 30), 8)[0] == -1  # n % b != 0
    assert orc.getrf(np.zeros((0, 0)), 8)[0] == 0           # empty


def test_solve_and_residual(orc):
    for n, b in [(48, 12), (256, 64)]:
        r = orc.lu_solve_problem(n, b)
        x_ref = sl.solve(r["A"], r["b"])
        assert np.abs(r["x"] - x_ref).max() <= 1e-10 * np.abs(x_ref).max()
        assert r["residual"] < 16 and r["residual_ns"] < 16


def test_residual_closed_form(orc):
    # A = diag(2), x = 1, b = 2 + delta: max|b - Ax| = delta
    n = 4
    A = np.diag(np.full(n, 2.0))
    x = np.ones(n)
    b = np.full(n, 2.0)
    b[2] += 1e-12
    r, ns, m4 = orc.residual(A, x, b)
    eps = np.finfo(float).eps
    assert np.isclose(m4[3], 1e-12, rtol=1e-3)
    assert np.isclose(r, m4[3] / ((2.0 * 1.0 + (2.0 + 1e-12)) * n * eps), rtol=1e-12)
    assert np.isclose(ns, m4[3] / (2.0 * 1.0 * n * eps), rtol=1e-12)


def test_golden_fixtures_reproduce(orc):
    g = np.load(GOLDEN)
    from tests.golden.make_golden import GRID
    for n, b, batch in GRID:
        if n > 64:
            continue
        r = orc.lu_solve_problem(n, b, batch)
        key = f"n{n}_b{b}_k{batch}"
        assert np.array_equal(r["ipiv"], g[key + "_ipiv"])
        assert np.array_equal(r["x"], g[key + "_x"])
        assert np.array_equal(r["LU"], g[key + "_lu"])
    assert np.array_equal(orc.gen_rhs(4096), g["rhs_4096"])


# ---- the FP32 strip restatement of the mixed-path strip test ---------------------
def test_fp32_strip_restatement_pinned():
    """The FP32 strip restatement (tests/test_gpu_mixed.py) is itself pinned:
    its pivots are LAPACK's for the same strip (getrf on the FP64 copy of the
    FP32 values; no near-ties at this seed), |l_ij| <= 1, and P S = L U to
    FP32 rounding.  A wrong tie rule, a skipped swap or a dropped term in the
    update fails one of these."""
    from tests.test_gpu_mixed import _strip_lu_fp32
    rng = np.random.default_rng(11)
    S0 = (rng.random((300, 32)) - 0.5).astype(np.float32)
    S, piv = _strip_lu_fp32(S0)
    _, lp = sl.lu_factor(S0.astype(np.float64))
    assert np.array_equal(piv, lp[:32])
    L = np.tril(S.astype(np.float64), -1)[:, :32]
    L[np.arange(32), np.arange(32)] = 1.0
    U = np.triu(S.astype(np.float64)[:32])
    assert np.abs(np.tril(S, -1)).max() <= 1.0
    P = S0.astype(np.float64).copy()
    for k, p in enumerate(piv):
        if p != k:
            P[[k, p]] = P[[p, k]]
    err = np.abs(P - L @ U).max() / np.abs(S0).max()
    assert err < 32 * 2.0 ** -24, err
