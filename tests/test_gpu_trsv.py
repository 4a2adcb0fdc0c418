"""The one-launch triangles of the vector solve against the launch-per-block
form and against solutions computed from the same factors on the CPU.

One right-hand side with n a multiple of the solve's diagonal block (256 for
charmlu_dgetrs / charmlu_sgetrs, 1024 inside the refinement from n = 8192)
runs each triangle as one launch (trsv_tiles); CHARMLU_TRSV_TILES=0 selects
the form with two launches per diagonal block (blk_gemv, gemv_sub).  Both
follow forwardSolve / backwardSolve / offDiagSolve (lu.ci:387-458); they sum
in different orders, so they agree to rounding, not bitwise.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _both(fn):
    """fn() with the one-launch triangles, then with two launches per block."""
    old = os.environ.get("CHARMLU_TRSV_TILES")
    try:
        os.environ["CHARMLU_TRSV_TILES"] = "1"
        a = fn()
        os.environ["CHARMLU_TRSV_TILES"] = "0"
        b = fn()
    finally:
        if old is None:
            os.environ.pop("CHARMLU_TRSV_TILES", None)
        else:
            os.environ["CHARMLU_TRSV_TILES"] = old
    return a, b


def _lu_solve_host(LU, ipiv, b, trans=False):
    """x from the factors P A = L U on the CPU in FP64: A x = b (P b, then
    L y = . , U x = y) or A^T x = b (U^T z = b, L^T w = z, x = P^T w)."""
    from scipy.linalg import solve_triangular
    n = LU.shape[0]
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    x = b.astype(np.float64).copy()
    if trans:
        z = solve_triangular(U, x, trans=1, lower=False)
        x = solve_triangular(L, z, trans=1, lower=True, unit_diagonal=True)
        for j in range(n - 1, -1, -1):
            p = int(ipiv[j])
            if p != j:
                x[j], x[p] = x[p], x[j]
        return x
    for j in range(n):
        p = int(ipiv[j])
        if p != j:
            x[j], x[p] = x[p], x[j]
    y = solve_triangular(L, x, lower=True, unit_diagonal=True)
    return solve_triangular(U, y, lower=False)


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("n,block", [(256, 64), (1024, 128), (4096, 256), (2560, 512)])
def test_dgetrs_tiles_match_per_block_form_and_host(dev, orc, n, block, trans):
    import torch
    print(f"dgetrs tiles n={n} block={block} trans={trans}", flush=True)
    A = dev.gen_matrix(n, seed=5 + n)
    LU = A.clone()
    ipiv, info = dev.getrf(LU, block)
    assert int(info.item()) == 0
    b = dev.gen_rhs(n, seed=23)

    def solve():
        x = b.clone()
        dev.getrs(LU, ipiv, x, trans=trans)
        torch.cuda.synchronize()
        return x.cpu().numpy()

    xt, xb = _both(solve)
    assert np.all(np.isfinite(xt))
    xh = _lu_solve_host(LU.cpu().numpy(), ipiv.cpu().numpy(), b.cpu().numpy(), trans)
    scale = np.abs(xh).max()
    assert np.abs(xt - xh).max() / scale <= 1e-12
    assert np.abs(xt - xb).max() / scale <= 1e-12
    # repeated calls reuse the published-value buffers under new sequence numbers
    x2, _ = _both(solve)
    assert np.array_equal(x2, xt)


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("n", [1024, 8192])
def test_sgetrs_tiles_match_per_block_form(dev, orc, n, trans):
    import torch
    print(f"sgetrs tiles n={n} trans={trans}", flush=True)
    Ad = orc.gen_diagdom(orc.SEED_A, n)
    A32 = dev.lag2s(torch.from_numpy(Ad).cuda())
    ipiv, info = dev.getrf_mixed(A32, 512, "fp32")
    assert int(info.item()) == 0
    rng = np.random.default_rng(n)
    bh = rng.uniform(-0.5, 0.5, n).astype(np.float32)

    def solve():
        x = torch.from_numpy(bh.copy()).cuda()
        dev.getrs32(A32, ipiv, x, trans=trans)
        torch.cuda.synchronize()
        return x.cpu().numpy()

    xt, xb = _both(solve)
    assert np.all(np.isfinite(xt))
    xh = _lu_solve_host(A32.cpu().numpy().astype(np.float64), ipiv.cpu().numpy(), bh, trans)
    scale = np.abs(xh).max()
    # FP32 factors, FP64 sums in both forms, one FP32 rounding per block step
    assert np.abs(xt - xh).max() / scale <= 64 * 2.0 ** -24
    assert np.abs(xt - xb).max() / scale <= 64 * 2.0 ** -24


def test_refinement_with_tiles_matches_oracle_at_block_1024(dev, orc):
    """n = 8192: the refinement's preconditioner solves use 1024-row blocks;
    A x = b against the oracle, A^T x = b against the FP64 solve of A^T."""
    import torch
    n = 8192
    print(f"gesv_ir tiles n={n}", flush=True)
    Ad = orc.gen_diagdom(orc.SEED_A, n)
    o = orc.lu_solve_problem(n, 256, A=Ad.copy())
    dA = torch.from_numpy(Ad).cuda()
    A32 = dev.lag2s(dA)
    ipiv, info = dev.getrf_mixed(A32, 1024, "bf16")
    assert int(info.item()) == 0
    rhs = dev.gen_rhs(n)

    def refine():
        x, st = dev.gesv_ir(dA, A32, ipiv, rhs, tol=1.0)
        return x.cpu().numpy(), st

    (xt, st_t), (xb, st_b) = _both(refine)
    for x, st in ((xt, st_t), (xb, st_b)):
        assert st["residual"] <= 1.0
        assert np.abs(x - o["x"]).max() / np.abs(o["x"]).max() <= 1e-10

    def refine_t():
        X, st = dev.gesv_ir_multi(dA, A32, ipiv, rhs.reshape(n, 1), trans=True, tol=1.0)
        return X.cpu().numpy()[:, 0], st

    (xt, st_t), (xb, st_b) = _both(refine_t)
    want = np.linalg.solve(Ad.T, rhs.cpu().numpy())
    for x in (xt, xb):
        assert np.abs(x - want).max() / np.abs(want).max() <= 1e-10
