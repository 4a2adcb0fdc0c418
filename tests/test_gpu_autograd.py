"""The differentiable solve (charmlu_amd.autograd.LUSolve) on the GPU.

The backward pass must equal
  * the analytic adjoint built from the FP64 oracle's factors:
    A^T Lambda = W (LAPACK getrs with trans = T on the oracle's LU and
    pivots), dA = -Lambda X^T, dB = Lambda;
  * FP64 central finite differences of loss = sum(W * X).
The test matrix is A = Q1 diag(s) Q2^T, each Q a product of 8 Householder
reflectors from a seeded normal stream, s log-spaced 1..1e4 (condition
number 1e4), 64 right-hand sides.
"""
import numpy as np
import pytest
import scipy.linalg as sl

pytestmark = pytest.mark.gpu


def householder_matrix(n, seed, nref=8, smax=1e4):
    """A = Q1 diag(s) Q2^T with Q = H_1 ... H_nref, H = I - 2 v v^T / v^T v."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    s = torch.logspace(0, float(np.log10(smax)), n, dtype=torch.float64)
    A = torch.diag(s).cuda()
    vs = [torch.randn(n, dtype=torch.float64, generator=g).cuda() for _ in range(2 * nref)]
    for v in vs[nref:]:          # right factor: A <- A H (H symmetric)
        A.addr_(A @ v, v, alpha=-2.0 / float(v @ v))
    for v in vs[:nref]:          # left factor: A <- H A
        A.addr_(v, v @ A, alpha=-2.0 / float(v @ v))
    return A


def _loss_and_grads(dev, A, B, W):
    import torch
    Ar = A.clone().requires_grad_(True)
    Br = B.clone().requires_grad_(True)
    X = dev.solve(Ar, Br)
    loss = (W * X).sum()
    loss.backward()
    return X.detach(), Ar.grad, Br.grad


@pytest.mark.parametrize("n,k,block", [(256, 4, 64), (520, 3, 8), (384, 1, 128)])
def test_backward_matches_adjoint_small(dev, n, k, block):
    import torch
    torch.manual_seed(n)
    A = householder_matrix(n, seed=n, nref=4, smax=1e3)
    B = torch.randn(n, k, dtype=torch.float64, device="cuda")
    W = torch.randn(n, k, dtype=torch.float64, device="cuda")
    Ar = A.clone().requires_grad_(True)
    Br = B.clone().requires_grad_(True)
    X = dev.solve(Ar, Br, block=block)
    (W * X).sum().backward()
    An, Bn, Wn = A.cpu().numpy(), B.cpu().numpy(), W.cpu().numpy()
    Xr = np.linalg.solve(An, Bn)
    Lam = np.linalg.solve(An.T, Wn)
    assert np.abs(X.detach().cpu().numpy() - Xr).max() / np.abs(Xr).max() < 1e-11
    dA = Ar.grad.cpu().numpy()
    assert np.abs(dA + Lam @ Xr.T).max() / np.abs(Lam @ Xr.T).max() < 1e-11
    assert np.abs(Br.grad.cpu().numpy() - Lam).max() / np.abs(Lam).max() < 1e-11


def test_backward_vector_rhs(dev):
    import torch
    n = 300
    A = householder_matrix(n, seed=5, nref=2, smax=1e2)
    b = torch.randn(n, dtype=torch.float64, device="cuda")
    w = torch.randn(n, dtype=torch.float64, device="cuda")
    Ar = A.clone().requires_grad_(True)
    br = b.clone().requires_grad_(True)
    x = dev.solve(Ar, br)
    (w * x).sum().backward()
    An = A.cpu().numpy()
    lam = np.linalg.solve(An.T, w.cpu().numpy())
    xr = np.linalg.solve(An, b.cpu().numpy())
    assert np.abs(Ar.grad.cpu().numpy() + np.outer(lam, xr)).max() < 1e-11 * np.abs(
        np.outer(lam, xr)).max()
    assert np.abs(br.grad.cpu().numpy() - lam).max() < 1e-11 * np.abs(lam).max()


def test_singular_matrix_raises(dev):
    import torch
    A = torch.zeros(64, 64, dtype=torch.float64, device="cuda")
    B = torch.ones(64, 2, dtype=torch.float64, device="cuda")
    with pytest.raises(torch.linalg.LinAlgError):
        dev.solve(A, B)


def test_c4_fp64_vs_oracle_adjoint_and_finite_differences(dev, orc):
    """n = 8192, 64 right-hand sides, condition number 1e4."""
    import torch
    n, k, block = 8192, 64, 256
    A = householder_matrix(n, seed=8192)
    g = torch.Generator(device="cpu").manual_seed(64)
    B = torch.randn(n, k, dtype=torch.float64, generator=g).cuda()
    W = torch.randn(n, k, dtype=torch.float64, generator=g).cuda()
    X, dA, dB = _loss_and_grads(dev, A, B, W)
    # analytic adjoint from the oracle's FP64 factors
    Xr, dA_ref, Lam = _oracle_adjoint(orc, A, B, W, block)
    rel = lambda a, b: np.abs(a - b).max() / np.abs(b).max()
    e_x = rel(X.cpu().numpy(), Xr)
    e_a = rel(dA.cpu().numpy(), dA_ref)
    e_b = rel(dB.cpu().numpy(), Lam)
    print(f"C4 fp64: x {e_x:.2e} dA {e_a:.2e} dB {e_b:.2e}")
    assert e_x < 1e-10 and e_a < 1e-10 and e_b < 1e-10
    # central finite differences of the loss at the largest-gradient entries
    rng = np.random.default_rng(1)
    cand = rng.integers(0, n, size=(256, 2))
    vals = np.abs(dA_ref[cand[:, 0], cand[:, 1]])
    pick = cand[np.argsort(vals)[-4:]]

    def loss(Ap, Bp):
        return float((W * dev.solve(Ap, Bp)).sum().item())
    # ||A^-1|| = 1 here, so a step of 1e-3 keeps the central difference's
    # truncation error (~h^2 (A^-1)_ji^2 relative) below 1e-6 while the
    # rounding noise of the loss, ~1e-13 |loss|, stays far below h
    for i, j in pick:
        h = 1e-3
        Ap = A.clone(); Ap[i, j] += h
        Am = A.clone(); Am[i, j] -= h
        fd = (loss(Ap, B) - loss(Am, B)) / (2 * h)
        assert abs(fd - dA_ref[i, j]) <= 1e-5 * abs(dA_ref[i, j]), (i, j, fd, dA_ref[i, j])
    for i, j in [(17, 3), (5000, 63)]:
        h = 1e-3
        Bp = B.clone(); Bp[i, j] += h
        Bm = B.clone(); Bm[i, j] -= h
        fd = (loss(A, Bp) - loss(A, Bm)) / (2 * h)
        assert abs(fd - Lam[i, j]) <= 1e-5 * abs(Lam[i, j]) + 1e-9 * np.abs(Lam).max()


def _oracle_adjoint(orc, A, B, W, block=256):
    """X, dA = -Lambda X^T and dB = Lambda from the oracle's FP64 factors
    (A^T Lambda = W by LAPACK getrs with trans = T on them)."""
    An = A.cpu().numpy()
    LU = An.copy()
    info, ipiv = orc.getrf(LU, block, nthreads=16)
    assert info == 0
    Xr = sl.lu_solve((LU, ipiv), B.cpu().numpy())
    Lam = sl.lu_solve((LU, ipiv), W.cpu().numpy(), trans=1)
    return Xr, -(Lam @ Xr.T), Lam


def _fd_loss(dev, W):
    """loss(A, B) = sum(W * X) with X from the FP64 device LU (itself
    checked index for index against the oracle)."""
    def loss(Ap, Bp):
        return float((W * dev.solve(Ap, Bp)).sum().item())
    return loss


def test_c4_fp32_vs_oracle_adjoint_and_finite_differences(dev, orc):
    """The FP32 form (FP32 factors, solves and gradient) of C4, against the
    adjoint from the oracle's FP64 factors and FP64 central finite
    differences.  Bound: 4e-3 relative, from kappa u32 = 1e4 * 6e-8 = 6e-4
    per solve times the two solves of the gradient and a factor 3 of margin
    (measured about 1e-3)."""
    import torch
    n, k = 8192, 64
    A = householder_matrix(n, seed=8192)
    g = torch.Generator(device="cpu").manual_seed(64)
    B = torch.randn(n, k, dtype=torch.float64, generator=g).cuda()
    W = torch.randn(n, k, dtype=torch.float64, generator=g).cuda()
    X32, dA32, dB32 = _loss_and_grads(dev, A.float(), B.float(), W.float())
    Xr, dA_ref, Lam = _oracle_adjoint(orc, A, B, W)
    rel = lambda a, b: float(np.abs(a.double().cpu().numpy() - b).max() / np.abs(b).max())
    e = (rel(X32, Xr), rel(dA32, dA_ref), rel(dB32, Lam))
    print(f"C4 fp32 vs oracle: x {e[0]:.2e} dA {e[1]:.2e} dB {e[2]:.2e}")
    assert max(e) < 4e-3
    # FP64 central differences at the largest-gradient entries (sampled)
    loss = _fd_loss(dev, W)
    rng = np.random.default_rng(3)
    cand = rng.integers(0, n, size=(256, 2))
    pick = cand[np.argsort(np.abs(dA_ref[cand[:, 0], cand[:, 1]]))[-3:]]
    scale = np.abs(dA_ref).max()
    for i, j in pick:
        h = 1e-3
        Ap = A.clone(); Ap[i, j] += h
        Am = A.clone(); Am[i, j] -= h
        fd = (loss(Ap, B) - loss(Am, B)) / (2 * h)
        got = float(dA32[i, j].item())
        print(f"  dA[{i},{j}]: fp32 {got:.6e} fd {fd:.6e}")
        assert abs(got - fd) <= 4e-3 * scale


# ---------------------------------------------------------------------------
# the mixed-precision form: low-precision factors + FP64 GMRES-IR, forward
# and adjoint (C5's "adjoint via stored low-precision factors + IR")
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec,n", [("bf16", 4096), ("fp8", 4096), ("fp16", 2048),
                                    ("bf16", 8192), ("fp8", 8192)])
def test_mixed_solve_adjoint_vs_oracle_and_finite_differences(dev, orc, prec, n):
    """A diagonally dominant (a_ii = sum_j |a_ij|, as C3 and C5), 64
    right-hand sides.  Forward: low-precision factors + FP64 GMRES-IR;
    backward: A^T Lambda = G by transposed GMRES-IR from the same stored
    factors, dA = -Lambda X^T.  X, dA and dB within 1e-10 of the adjoint
    from the oracle's FP64 factors, and dA, dB within 1e-7 of FP64 central
    finite differences."""
    import torch
    k = 64
    A = dev.gen_matrix(n, seed=77, kind="diagdom")
    g = torch.Generator(device="cpu").manual_seed(n + 1)
    B = torch.randn(n, k, dtype=torch.float64, generator=g).cuda()
    W = torch.randn(n, k, dtype=torch.float64, generator=g).cuda()
    Ar = A.clone().requires_grad_(True)
    Br = B.clone().requires_grad_(True)
    X = dev.solve(Ar, Br, precision=prec)
    (W * X).sum().backward()
    Xr, dA_ref, Lam = _oracle_adjoint(orc, A, B, W)
    rel = lambda a, b: float(np.abs(a.detach().cpu().numpy() - b).max() / np.abs(b).max())
    e = (rel(X, Xr), rel(Ar.grad, dA_ref), rel(Br.grad, Lam))
    print(f"mixed {prec} n={n}: vs oracle x {e[0]:.2e} dA {e[1]:.2e} dB {e[2]:.2e}")
    assert max(e) < 1e-10
    # FP64 central differences.  The entries of A are O(1) off the diagonal
    # and about n/4 on it, so (A^-1)_jj ~ 4/n and the central difference's
    # relative error ~ (h (A^-1)_jj)^2 is 1e-8 at h = 0.05 (0.5 measured
    # 1e-6 on a diagonal entry at n = 2048); the loss's rounding noise over
    # 2h stays near 1e-13 of the gradient's scale
    loss = _fd_loss(dev, W)
    scale_a, scale_b = np.abs(dA_ref).max(), np.abs(Lam).max()
    for i, j in [(3, 3), (17, n - 5), (n // 2, 11)]:
        h = 0.05
        Ap = A.clone(); Ap[i, j] += h
        Am = A.clone(); Am[i, j] -= h
        fd = (loss(Ap, B) - loss(Am, B)) / (2 * h)
        assert abs(float(Ar.grad[i, j].item()) - fd) <= 1e-7 * scale_a, (i, j, fd)
    for i, j in [(5, 0), (n - 1, 63)]:
        h = 0.5
        Bp = B.clone(); Bp[i, j] += h
        Bm = B.clone(); Bm[i, j] -= h
        fd = (loss(A, Bp) - loss(A, Bm)) / (2 * h)
        assert abs(float(Br.grad[i, j].item()) - fd) <= 1e-7 * scale_b, (i, j, fd)


def test_gesv_ir_transposed_multi_rhs_vs_oracle(dev, orc):
    """charmlu_dgesv_ir_ex directly: A^T X = B and A X = B with 5
    right-hand sides (one of them zero) from BF16 factors of a uniform
    (pivoting) matrix of n = 1024, against scipy's solves on the oracle's
    FP64 factors; every column refined to the reference's residual <= 1."""
    import torch
    n, k = 1024, 5
    A = dev.gen_matrix(n, seed=31)
    g = torch.Generator(device="cpu").manual_seed(5)
    B = torch.randn(n, k, dtype=torch.float64, generator=g).cuda()
    B[:, 2] = 0.0
    S = dev.lag2s(A)
    ipiv, info = dev.getrf_mixed(S, 256, "bf16")
    assert int(info.item()) == 0
    An = A.cpu().numpy()
    LU = An.copy()
    assert orc.getrf(LU, 256)[0] == 0
    oip = orc.getrf(An.copy(), 256)[1]
    for trans in (False, True):
        X, st = dev.gesv_ir_multi(A, S, ipiv, B, trans=trans, tol=1.0)
        ref = sl.lu_solve((LU, oip), B.cpu().numpy(), trans=1 if trans else 0)
        err = np.abs(X.cpu().numpy() - ref).max() / np.abs(ref).max()
        print(f"trans={trans}: {st}, vs oracle {err:.2e}")
        assert err < 1e-10 and st["residual"] <= 1.0
        assert np.all(X[:, 2].cpu().numpy() == 0.0)


@pytest.mark.parametrize("n,block", [(1024, 256), (1088, 64), (999, 37)])
def test_gesv_ir_transposed_norms(dev, orc, n, block):
    """The transposed refinement's scaled residuals use max|a_ij|, the
    largest column sum ||A^T||_inf and max|b| (abs_colsum on the device):
    their ratio residual / residual_ns = ||A^T||_inf max|x| /
    (max|a| max|x| + max|b|) is free of the residual's rounding, so it pins
    those maxima against numpy on the same A, b and returned x (even n, a
    partial last group of 128 columns, odd n)."""
    import torch
    A = dev.gen_matrix(n, seed=41)
    b = dev.gen_rhs(n, seed=43)
    S = dev.lag2s(A)
    ipiv, info = dev.getrf_mixed(S, block, "bf16")
    assert int(info.item()) == 0
    x, st = dev.gesv_ir_multi(A, S, ipiv, b, trans=True, tol=1.0)
    An, bn, xn = A.cpu().numpy(), b.cpu().numpy(), x.cpu().numpy()
    xmax = np.abs(xn).max()
    want = np.abs(An).sum(axis=0).max() * xmax / (np.abs(An).max() * xmax + np.abs(bn).max())
    assert st["residual"] > 0 and st["residual"] <= 1.0
    assert np.isclose(st["residual"] / st["residual_ns"], want, rtol=1e-12, atol=0)
    assert np.abs(An.T @ xn - bn).max() <= 16 * n * 2.0 ** -53 * (
        np.abs(An).max() * xmax + np.abs(bn).max())
