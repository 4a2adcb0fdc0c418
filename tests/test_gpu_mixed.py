"""Mixed-precision factorization and FP64 GMRES-based refinement on the GPU.

The reference factors in FP64 only, so the mixed path is checked against the
FP64 oracle's solution of the same seeded system:
  * the refined x within 1e-10 relative (normwise) of the oracle's FP64 x,
    and the reference's scaled residual < 16 (test_lu.pl:75-76);
  * the low-precision factors themselves through their backward error
    ||P A - L U||_max / (max|a_ij| * growth) against the unit roundoff of
    the update precision (a property of any correct LU, so a wrong
    operand, sign or index in the update fails it by orders of magnitude);
  * diagonally dominant matrices need no interchanges: pivots are the
    identity, as in the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

UNIT = {"fp32": 2.0 ** -24, "bf16": 2.0 ** -8, "fp16": 2.0 ** -11}


def _unpack(LU):
    L = np.tril(LU, -1) + np.eye(LU.shape[0])
    U = np.triu(LU)
    return L, U


def _perm_rows(A, ipiv):
    P = A.copy()
    for k, p in enumerate(ipiv):
        if p != k:
            P[[k, p]] = P[[p, k]]
    return P


def _backward_error(A, LU, ipiv):
    L, U = _unpack(LU.astype(np.float64))
    PA = _perm_rows(A, ipiv)
    return np.abs(PA - L @ U).max() / np.abs(A).max()


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
@pytest.mark.parametrize("M,N,K,alpha,beta,ldc_pad", [
    (512, 384, 256, 1.0, 1.0, 0), (300, 200, 64, 1.0, 1.0, 0), (256, 256, 32, -1.0, 1.0, 0),
    (129, 257, 96, 0.5, -2.0, 3), (1000, 130, 160, 1.0, 0.0, 1), (128, 128, 256, 1.0, 1.0, 4)])
def test_gemm_lp_dev_vs_fp32(dev, prec, M, N, K, alpha, beta, ldc_pad):
    """charmlu_gemm_lp_dev against an FP32 product of the same 16-bit
    operands: the matrix cores accumulate in FP32, so the only difference is
    the summation order (|err| <= K * 2^-23 * sum |a||b| + |C| terms)."""
    import ctypes
    import torch
    from charmlu_amd import _lib
    torch.manual_seed(M + N + K)
    dt = torch.bfloat16 if prec == "bf16" else torch.float16
    A = (torch.rand(M, K, device="cuda") - 0.5).to(dt)
    B = (torch.rand(N, K, device="cuda") - 0.5).to(dt)
    ldc = N + ldc_pad
    Cfull = torch.rand(M, ldc, device="cuda")
    C0 = Cfull.clone()
    if beta == 0.0:
        Cfull[:, :N] = float("nan")  # beta = 0 must not read C
    rc = _lib.load().charmlu_gemm_lp_dev(
        _lib.PREC_BF16 if prec == "bf16" else _lib.PREC_FP16, M, N, K, alpha,
        ctypes.c_void_p(A.data_ptr()), K, ctypes.c_void_p(B.data_ptr()), K, beta,
        ctypes.c_void_p(Cfull.data_ptr()), ldc, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    ref = alpha * (A.double() @ B.double().T) + (beta * C0[:, :N].double() if beta != 0 else 0)
    bound = K * 2.0 ** -23 * (abs(alpha) * (A.double().abs() @ B.double().abs().T)) + \
        2.0 ** -23 * (abs(beta) * C0[:, :N].double().abs() + ref.abs()) * 4
    err = (Cfull[:, :N].double() - ref).abs()
    assert bool((err <= bound + 1e-30).all()), float(err.max())
    # columns past N untouched
    assert torch.equal(Cfull[:, N:], C0[:, N:])


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp8"])
def test_large_update_tiles_vs_fp64(dev, prec):
    """The trailing update's large-grid tiles (256 x 256 for 16-bit
    operands, 128 x 256 for FP8; taken from 2^25 C elements with K >= 1024,
    C accumulated in place) on a ragged 6144 x 5900 x 1024 product against
    the FP64 product of the same operands: 16-bit within the FP32
    summation bound, FP8 within the e4m3 rounding bound of its operands."""
    import ctypes
    import torch
    from charmlu_amd import _lib
    lib = _lib.load()
    M, N, K = 6144, 5900, 1024
    torch.manual_seed(7)
    C0 = torch.rand(M, N, device="cuda")
    C = C0.clone()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if prec == "fp8":
        A = torch.rand(M, K, device="cuda") - 0.5
        B = torch.rand(K, N, device="cuda") - 0.5
        rc = lib.charmlu_gemm_fp8_dev(M, N, K, 1.0, ctypes.c_void_p(A.data_ptr()), K,
                                      ctypes.c_void_p(B.data_ptr()), N, 1.0,
                                      ctypes.c_void_p(C.data_ptr()), N, s)
        torch.cuda.synchronize()
        prod = A.double() @ B.double()
        bound = 2 * 2.0 ** -4 * (A.double().abs() @ B.double().abs()) + 2.0 ** -22 * (
            prod.abs() + C0.double().abs()) * 4 + 1e-30
    else:
        dt = torch.bfloat16 if prec == "bf16" else torch.float16
        A = (torch.rand(M, K, device="cuda") - 0.5).to(dt)
        B = (torch.rand(N, K, device="cuda") - 0.5).to(dt)
        rc = lib.charmlu_gemm_lp_dev(_lib.PREC_BF16 if prec == "bf16" else _lib.PREC_FP16, M, N,
                                     K, 1.0, ctypes.c_void_p(A.data_ptr()), K,
                                     ctypes.c_void_p(B.data_ptr()), K, 1.0,
                                     ctypes.c_void_p(C.data_ptr()), N, s)
        torch.cuda.synchronize()
        prod = A.double() @ B.double().T
        bound = K * 2.0 ** -23 * (A.double().abs() @ B.double().abs().T) + 2.0 ** -23 * (
            C0.double().abs() + prod.abs()) * 4 + 1e-30
    assert rc == 0
    err = (C.double() - (prod + C0.double())).abs()
    assert bool((err <= bound).all()), float((err / bound).max())


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
@pytest.mark.parametrize("n,nb", [(512, 128), (1024, 256), (768, 256), (960, 192), (1002, 167)])
def test_mixed_factor_diagdom(dev, orc, prec, n, nb):
    import torch
    A = orc.gen_diagdom(orc.SEED_A, n)
    A32 = torch.from_numpy(A.astype(np.float32)).cuda()
    ipiv, info = dev.getrf_mixed(A32, nb, prec)
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    assert np.array_equal(ipiv.cpu().numpy(), np.arange(n))
    be = _backward_error(A, A32.cpu().numpy(), ipiv.cpu().numpy())
    print(f"{prec} n={n} nb={nb} backward error {be:.3e}")
    # each of the n/nb update stages rounds its operands once
    assert be < 8 * UNIT[prec] * (n // nb + 1)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
@pytest.mark.parametrize("n,nb", [(1024, 128), (1002, 167)])
def test_mixed_factor_uniform(dev, orc, prec, n, nb):
    """General matrices pivot; the factors are still a backward-stable LU of
    the FP32 input at the update precision (edge tiles and an unaligned
    leading dimension with n = 1002)."""
    import torch
    A = orc.gen_uniform(orc.SEED_A, n)
    A32 = torch.from_numpy(A.astype(np.float32)).cuda()
    ipiv, info = dev.getrf_mixed(A32, nb, prec)
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    ip = ipiv.cpu().numpy()
    assert np.all(ip >= np.arange(n)) and np.all(ip < n)
    LU = A32.cpu().numpy()
    # partial pivoting: |l_ij| <= 1
    assert np.abs(np.tril(LU, -1)).max() <= 1.0 + 1e-6
    be = _backward_error(A.astype(np.float32).astype(np.float64), LU, ip)
    print(f"{prec} uniform backward error {be:.3e}")
    assert be < 64 * UNIT[prec] * np.sqrt(n)


def test_sgetrs_both_transposes(dev, orc):
    import torch
    n, nb = 768, 128
    A = orc.gen_uniform(7, n)
    A32 = torch.from_numpy(A.astype(np.float32)).cuda()
    ipiv, _ = dev.getrf_mixed(A32, nb, "fp32")
    rng = np.random.default_rng(3)
    B = rng.standard_normal((n, 5))
    for trans in (False, True):
        X = torch.from_numpy(B.astype(np.float32)).cuda()
        dev.getrs32(A32, ipiv, X, trans=trans)
        x = X.cpu().numpy().astype(np.float64)
        ref = np.linalg.solve(A.T if trans else A, B)
        err = np.abs(x - ref).max() / np.abs(ref).max()
        assert err < 1e-3, (trans, err)
    # one right-hand side as a vector
    xv = torch.from_numpy(B[:, 0].astype(np.float32)).cuda()
    dev.getrs32(A32, ipiv, xv)
    ref = np.linalg.solve(A, B[:, 0])
    assert np.abs(xv.cpu().numpy() - ref).max() / np.abs(ref).max() < 1e-3


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp32"])
@pytest.mark.parametrize("n,nb", [(1024, 256), (2048, 256), (2048, 512)])
def test_gmres_ir_matches_fp64_oracle(dev, orc, prec, n, nb):
    import torch
    A = orc.gen_diagdom(orc.SEED_A, n)
    b = orc.gen_rhs(n)
    x_ref = orc.lu_solve_problem(n, nb, A=A.copy())["x"]
    dA = torch.from_numpy(A).cuda()
    A32 = dA.float()
    ipiv, info = dev.getrf_mixed(A32, nb, prec)
    assert int(info.item()) == 0
    db = torch.from_numpy(b).cuda()
    x, st = dev.gesv_ir(dA, A32, ipiv, db, tol=1.0)
    xh = x.cpu().numpy()
    err = np.abs(xh - x_ref).max() / np.abs(x_ref).max()
    print(f"{prec} n={n}: {st}, error vs FP64 oracle {err:.2e}")
    assert st["residual"] <= 1.0
    res = dev.residual(dA, x, db)
    assert res["residual"] < 16
    assert err < 1e-10
    assert st["cycles"] >= 1


def test_gmres_ir_uniform_fp32_factors(dev, orc):
    """A general (pivoting) matrix refined from FP32 factors."""
    import torch
    n, nb = 1024, 128
    A = orc.gen_uniform(orc.SEED_A, n)
    b = orc.gen_rhs(n)
    x_ref = orc.lu_solve_problem(n, nb, A=A.copy())["x"]
    dA = torch.from_numpy(A).cuda()
    A32 = dA.float()
    ipiv, _ = dev.getrf_mixed(A32, nb, "fp32")
    x, st = dev.gesv_ir(dA, A32, ipiv, torch.from_numpy(b).cuda(), tol=1.0)
    err = np.abs(x.cpu().numpy() - x_ref).max() / np.abs(x_ref).max()
    print(f"uniform fp32: {st}, error {err:.2e}")
    assert st["residual"] <= 1.0
    assert err < 1e-10


def test_gmres_ir_already_converged_and_limits(dev, orc):
    import torch
    n, nb = 512, 128
    A = orc.gen_diagdom(3, n)
    b = orc.gen_rhs(n)
    dA = torch.from_numpy(A).cuda()
    A32 = dA.float()
    ipiv, _ = dev.getrf_mixed(A32, nb, "bf16")
    db = torch.from_numpy(b).cuda()
    # max_iter = 0: x is the low-precision solve only
    x0, st0 = dev.gesv_ir(dA, A32, ipiv, db, tol=1.0, max_iter=0)
    assert st0["cycles"] == 0 and st0["iterations"] == 0
    assert st0["residual"] > 1.0
    x1, st1 = dev.gesv_ir(dA, A32, ipiv, db, tol=1.0)
    assert st1["residual"] <= 1.0 and st1["cycles"] >= 1
    # a loose tolerance needs no more work than a tight one
    _, st2 = dev.gesv_ir(dA, A32, ipiv, db, tol=1e3)
    assert st2["iterations"] <= st1["iterations"]


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp32"])
@pytest.mark.parametrize("n,nb", [(1024, 256), (2048, 512), (1536, 384), (1000, 200),
                                  (4096, 2048)])
def test_unpivoted_mixed_on_diagdom(dev, orc, prec, n, nb):
    """Without interchanges on a diagonally dominant matrix: identity
    pivots, every |l_ij| < 1 (so partial pivoting would choose the same
    pivots), factors close to the pivoted ones, and the refined x within
    1e-10 of the FP64 oracle."""
    import torch
    A = orc.gen_diagdom(orc.SEED_A, n)
    b = orc.gen_rhs(n)
    dA = torch.from_numpy(A).cuda()
    P32 = dA.float()
    ipiv_p, _ = dev.getrf_mixed(P32, nb, prec)
    U32 = dA.float()
    ipiv, info = dev.getrf_mixed(U32, nb, prec, pivot=False)
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    assert np.array_equal(ipiv.cpu().numpy(), np.arange(n))
    assert np.array_equal(ipiv_p.cpu().numpy(), np.arange(n))
    LU = U32.cpu().numpy()
    assert np.abs(np.tril(LU, -1)).max() < 1.0
    be = _backward_error(A, LU, np.arange(n))
    # the unpivoted panel forms inv(U11) of the nb-row diagonal block
    # explicitly (L21 = A21 inv(U11)), whose rounding grows with nb: the
    # bound of the update's roundoff per step, times nb / 512 beyond 512
    # (FP32 at nb = 2048: 5.0e-6 = 83 u)
    assert be < 8 * UNIT[prec] * (n // nb + 1) * max(1, nb // 512)
    diff = np.abs(LU - P32.cpu().numpy()).max() / np.abs(LU).max()
    assert diff < 64 * UNIT[prec]
    x_ref = orc.lu_solve_problem(n, nb, A=A.copy())["x"]
    x, st = dev.gesv_ir(dA, U32, ipiv, torch.from_numpy(b).cuda(), tol=1.0)
    err = np.abs(x.cpu().numpy() - x_ref).max() / np.abs(x_ref).max()
    print(f"nopiv {prec} n={n} nb={nb}: be {be:.2e} vs pivoted {diff:.2e} {st} err {err:.2e}")
    assert st["residual"] <= 1.0 and err < 1e-10


def test_unpivoted_singular_block_reports_info(dev):
    import torch
    n, nb = 512, 128
    A = torch.eye(n, dtype=torch.float32, device="cuda")
    A[200, 200] = 0.0
    ipiv, info = dev.getrf_mixed(A, nb, "bf16", pivot=False)
    assert int(info.item()) == 201


# ---------------------------------------------------------------------------
# FP8 e4m3 update with per-tile power-of-two scales
# ---------------------------------------------------------------------------
def _fp8_gemm(dev, A, B, alpha, beta, C):
    import ctypes
    import torch
    from charmlu_amd import _lib
    M, K = A.shape
    N = B.shape[1]
    rc = _lib.load().charmlu_gemm_fp8_dev(
        M, N, K, alpha, ctypes.c_void_p(A.data_ptr()), K, ctypes.c_void_p(B.data_ptr()), N, beta,
        ctypes.c_void_p(C.data_ptr()), N, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    return C


@pytest.mark.parametrize("M,N,K", [(256, 384, 512), (300, 140, 100), (128, 128, 64)])
def test_fp8_gemm_exact_on_representable_tiles(dev, M, N, K):
    """Entries in {+-0.5, +-1, +-1.5, +-2} with a 2 in every 128-row tile of A
    and every 128-column tile of B: each tile scales by 2^7 onto exact e4m3
    values, every product and partial sum is exact in FP32, so the result
    equals the FP64 product bit for bit (both the alpha/beta and the
    accumulate-into-C forms)."""
    import torch
    g = torch.Generator().manual_seed(M + N + K)
    vals = torch.tensor([-2.0, -1.5, -1.0, -0.5, 0.5, 1.0, 1.5, 2.0])
    A = vals[torch.randint(0, 8, (M, K), generator=g)]
    B = vals[torch.randint(0, 8, (K, N), generator=g)]
    A[::128, 0] = 2.0
    B[0, ::128] = -2.0
    ref = (A.double() @ B.double()).float()
    A, B = A.cuda(), B.cuda()
    C = torch.full((M, N), float("nan"), device="cuda")
    _fp8_gemm(dev, A, B, 1.0, 0.0, C)
    assert torch.equal(C.cpu(), ref)
    C0 = torch.zeros(M, N, device="cuda")
    _fp8_gemm(dev, A, B, 1.0, 1.0, C0)
    assert torch.equal(C0.cpu(), ref)


def test_fp8_gemm_random_within_e4m3_rounding(dev):
    """Random operands: each element within the e4m3 rounding bound
    (2 * 2^-4 relative per product, summed as |a||b|), and the overall
    relative error at the statistical size of that rounding."""
    import torch
    torch.manual_seed(5)
    M, N, K = 512, 256, 256
    A = (torch.rand(M, K, device="cuda") - 0.5) * torch.logspace(-6, 2, M, device="cuda")[:, None]
    B = torch.rand(K, N, device="cuda") - 0.5
    C0 = torch.rand(M, N, device="cuda")
    C = _fp8_gemm(dev, A, B, -1.0, 1.0, C0.clone())
    ref = C0.double() - A.double() @ B.double()
    bound = 2 * 2.0 ** -4 * (A.double().abs() @ B.double().abs()) + 2.0 ** -22 * ref.abs() + 1e-30
    err = (C.double() - ref).abs()
    assert bool((err <= bound).all())
    # per-tile scaling keeps small rows accurate: relative error of A@B per row
    P = C0.double() - C.double()
    Pr = A.double() @ B.double()
    rel = (P - Pr).norm(dim=1) / Pr.norm(dim=1)
    assert float(rel.max()) < 0.08, float(rel.max())


@pytest.mark.parametrize("pivot", [True, False])
@pytest.mark.parametrize("n,nb", [(1024, 128), (2048, 512), (1536, 256), (4096, 2048)])
def test_fp8_mixed_factor_and_refinement(dev, orc, n, nb, pivot):
    import torch
    A = orc.gen_diagdom(orc.SEED_A, n)
    b = orc.gen_rhs(n)
    dA = torch.from_numpy(A).cuda()
    A32 = dA.float()
    ipiv, info = dev.getrf_mixed(A32, nb, "fp8", pivot=pivot)
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    assert np.array_equal(ipiv.cpu().numpy(), np.arange(n))
    be = _backward_error(A, A32.cpu().numpy(), np.arange(n))
    assert be < 8 * 2.0 ** -4 * (n // nb + 1)
    x_ref = orc.lu_solve_problem(n, nb, A=A.copy())["x"]
    x, st = dev.gesv_ir(dA, A32, ipiv, torch.from_numpy(b).cuda(), tol=1.0)
    err = np.abs(x.cpu().numpy() - x_ref).max() / np.abs(x_ref).max()
    print(f"fp8 pivot={pivot} n={n} nb={nb}: backward error {be:.2e}, {st}, err {err:.2e}")
    assert st["residual"] <= 1.0 and err < 1e-10


def test_fp8_rejects_block_not_multiple_of_128(dev):
    import torch
    from charmlu_amd import CharmluError
    A = torch.eye(960, dtype=torch.float32, device="cuda")
    with pytest.raises(CharmluError):
        dev.getrf_mixed(A, 192, "fp8")


def test_tall_panel_mixed_refines_to_fp64(dev):
    # n = 40960, block 1024: FP32 panels taller than 32768 rows run on 512-
    # thread workgroups (20 of them).  The refined x must pass the
    # reference's criterion and agree with the FP64 device LU's x.
    import torch
    n, nb = 40960, 1024
    A = dev.gen_matrix(n, kind="diagdom")
    b = dev.gen_rhs(n)
    A32 = dev.lag2s(A)
    ipiv, info = dev.getrf_mixed(A32, nb, "bf16")
    assert int(info.item()) == 0
    x, st = dev.gesv_ir(A, A32, ipiv, b, tol=1.0)
    del A32
    W = A.clone()
    ip64, info64 = dev.getrf(W, nb)
    x64 = b.clone()
    dev.getrs(W, ip64, x64)
    torch.cuda.synchronize()
    assert int(info64.item()) == 0
    res = dev.residual(A, x, b)
    err = float(((x - x64).abs().max() / x64.abs().max()).item())
    print(f"n={n}: {st}, residual {res['residual']:.3f}, error vs FP64 LU {err:.2e}")
    assert st["residual"] <= 1.0 and res["residual"] < 16
    assert err < 1e-10


def _strip_lu_fp32(S):
    """Partial-pivoting LU of an FP32 m x w strip, column by column, in the
    strip kernel's per-element order (the row restated from the reference's
    panel loop, lu.C:259-288 and lu.C:230-254): the first largest |a| from
    row c down (lowest row on ties), the row exchange, the multipliers by
    FP32 division, then a[i][q] - l[i]*u[q] rounded once as the kernel's FMA
    does (the product of two FP32 values is exact in FP64)."""
    S = S.astype(np.float32).copy()
    m, w = S.shape
    piv = np.zeros(w, dtype=np.int64)
    for c in range(w):
        p = c + int(np.argmax(np.abs(S[c:, c])))
        piv[c] = p
        if p != c:
            S[[c, p]] = S[[p, c]]
        u = S[c].copy()
        if u[c] == 0:
            continue
        l = (S[c + 1:, c] / u[c]).astype(np.float32)
        S[c + 1:, c] = l
        if c + 1 < w:
            prod = l.astype(np.float64)[:, None] * u[c + 1:].astype(np.float64)[None, :]
            S[c + 1:, c + 1:] = (S[c + 1:, c + 1:].astype(np.float64) - prod).astype(np.float32)
    return S, piv


@pytest.mark.parametrize("n", [1024, 2048, 4096])
def test_fp32_first_strip_matches_restatement(dev, orc, n):
    """With block 32 the first panel of the FP32 LU is one 32-column strip
    (one workgroup at n <= 2048, four exchanging through memory at 4096)
    applied to the input itself: its pivots equal the restatement's, and its
    columns -- after the later steps' interchanges, replayed from the GPU's
    ipiv -- are bit-identical."""
    import torch
    A = orc.gen_uniform(orc.SEED_A + n, n).astype(np.float32)
    A32 = torch.from_numpy(A.copy()).cuda()
    ipiv, info = dev.getrf_mixed(A32, 32, "fp32")
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    ip = ipiv.cpu().numpy()
    S, piv = _strip_lu_fp32(A[:, :32])
    assert np.array_equal(ip[:32], piv), np.nonzero(ip[:32] != piv)[0]
    for k in range(32, n):
        if ip[k] != k:
            S[[k, ip[k]]] = S[[ip[k], k]]
    G = A32.cpu().numpy()[:, :32]
    err = np.abs(G.astype(np.float64) - S).max() / np.abs(S).max()
    print(f"n={n}: first strip max relative difference {err:.2e}")
    # same operations in the same order: bit-identical (measured at all
    # three sizes)
    assert np.array_equal(G, S), err


@pytest.mark.parametrize("prec,pivot,n,nb", [("bf16", True, 2048, 512), ("fp8", True, 2048, 512),
                                             ("fp32", True, 2048, 512),
                                             ("bf16", False, 2048, 512),
                                             ("fp8", False, 2048, 512),
                                             ("bf16", False, 4096, 256),
                                             ("fp8", False, 3840, 256)])
def test_forced_split_is_bitwise_unsplit_mixed(dev, orc, prec, pivot, n, nb):
    """The trailing update's split (rows beside the next panel, the rest on
    every CU after it) computes every tile as the unsplit launch does: with
    the split forced to half the rows at every step, the factors and pivots
    are bit-identical (the FP8 half starts at a multiple of 128 rows, so its
    per-tile scales line up).  The unpivoted cases of 15 and 16 steps split
    the rest of each pair's update three ways."""
    import os
    import torch
    A = (orc.gen_uniform if pivot else orc.gen_diagdom)(orc.SEED_A + 5, n)
    runs = []
    for forced in (False, True):
        A32 = torch.from_numpy(A.astype(np.float32)).cuda()
        if forced:
            os.environ["CHARMLU_FORCE_SPLIT"] = "1"
        try:
            ipiv, info = dev.getrf_mixed(A32, nb, prec, pivot=pivot)
            torch.cuda.synchronize()
        finally:
            os.environ.pop("CHARMLU_FORCE_SPLIT", None)
        assert int(info.item()) == 0
        runs.append((A32.cpu().numpy(), ipiv.cpu().numpy()))
    assert np.array_equal(runs[0][1], runs[1][1])
    assert np.array_equal(runs[0][0], runs[1][0])


GROUP_CASES = [(prec, g, n, nb) for prec in ("bf16", "fp8") for g in (2, 3, 4)
               for n, nb in ((4096, 256), (2560, 512))] + [
    ("fp16", 2, 1536, 512), ("fp16", 3, 4096, 256), ("fp8", 2, 3072, 1024),
    ("fp8", 3, 3072, 1024), ("bf16", 2, 1536, 512), ("fp8", 8, 4096, 256),
    ("bf16", 8, 4096, 256), ("fp8", 6, 2560, 512), ("fp8", 7, 3840, 256)]


@pytest.mark.parametrize("prec,g,n,nb", GROUP_CASES)
def test_grouped_steps_match_per_step_form(dev, orc, prec, g, n, nb):
    """The unpivoted low-precision form updates the rest of the matrix once
    per g steps (K = g nb, the steps' packed operands side by side, FP8 with
    a scale set per step; CHARMLU_GROUP=g); CHARMLU_GROUP=1 updates it once
    per step.  Both factor within the update's roundoff bound of the FP64
    matrix, agree to a few units of it, and refine to the FP64 oracle's x
    (3 to 16 steps, groups of 2 to 8, cut short at the end)."""
    import os
    import torch
    A = orc.gen_diagdom(orc.SEED_A + 7, n)
    b = orc.gen_rhs(n)
    dA = torch.from_numpy(A).cuda()
    unit = 2.0 ** -4 if prec == "fp8" else UNIT[prec]
    out = {}
    for group in (g, 1):
        A32 = dA.float()
        os.environ["CHARMLU_GROUP"] = str(group)
        try:
            print(f"{prec} n={n} nb={nb} group={group}", flush=True)
            ipiv, info = dev.getrf_mixed(A32, nb, prec, pivot=False)
            torch.cuda.synchronize()
        finally:
            os.environ.pop("CHARMLU_GROUP", None)
        assert int(info.item()) == 0
        LU = A32.cpu().numpy()
        be = _backward_error(A, LU, np.arange(n))
        assert be < 8 * unit * (n // nb + 1) * max(1, nb // 512), (group, be)
        out[group] = (LU, ipiv, A32, be)
    diff = np.abs(out[g][0] - out[1][0]).max() / np.abs(out[1][0]).max()
    assert diff < 64 * unit, diff
    x_ref = orc.lu_solve_problem(n, nb, A=A.copy())["x"]
    _, ipiv, A32, be = out[g]
    x, st = dev.gesv_ir(dA, A32, ipiv, torch.from_numpy(b).cuda(), tol=1.0)
    err = np.abs(x.cpu().numpy() - x_ref).max() / np.abs(x_ref).max()
    print(f"group {g} {prec} n={n} nb={nb}: be {be:.2e} (per step {out[1][3]:.2e}), "
          f"diff {diff:.2e}, {st}, err {err:.2e}")
    assert st["residual"] <= 1.0 and err < 1e-10


@pytest.mark.parametrize("prec", ["fp8", "bf16", "fp16"])
@pytest.mark.parametrize("M,N,K", [(6144, 5900, 1024), (5950, 6400, 2048), (6400, 6144, 1024)])
def test_wide_tiles_match_narrow(dev, prec, M, N, K):
    """The large accumulate-into-C updates take 256 x 256 tiles of four
    128 x 128 waves for their whole tiles and 128 x 128 tiles for the rows
    below and the columns beside them; CHARMLU_LP_WIDE=0 takes tiles of
    64 x 64 per wave everywhere.  Each element of C receives the same
    products in the same order from both, so C is bit-identical (ragged N,
    ragged M, and no edge)."""
    import ctypes
    import os
    import torch
    from charmlu_amd import _lib
    torch.manual_seed(11)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    C0 = torch.rand(M, N, device="cuda")
    if prec == "fp8":
        A = torch.rand(M, K, device="cuda") - 0.5
        B = torch.rand(K, N, device="cuda") - 0.5
    else:
        dt = torch.bfloat16 if prec == "bf16" else torch.float16
        A = (torch.rand(M, K, device="cuda") - 0.5).to(dt)
        B = (torch.rand(N, K, device="cuda") - 0.5).to(dt)
    out = {}
    for wide in ("1", "0"):
        C = C0.clone()
        os.environ["CHARMLU_LP_WIDE"] = wide
        try:
            print(f"{prec} update {M}x{N}x{K} wide={wide}", flush=True)
            if prec == "fp8":
                _fp8_gemm(dev, A, B, 1.0, 1.0, C)
            else:
                rc = _lib.load().charmlu_gemm_lp_dev(
                    _lib.PREC_BF16 if prec == "bf16" else _lib.PREC_FP16, M, N, K, 1.0,
                    ctypes.c_void_p(A.data_ptr()), K, ctypes.c_void_p(B.data_ptr()), K, 1.0,
                    ctypes.c_void_p(C.data_ptr()), N, s)
                assert rc == 0
                torch.cuda.synchronize()
        finally:
            os.environ.pop("CHARMLU_LP_WIDE", None)
        out[wide] = C
    assert not torch.equal(out["1"], C0)
    assert torch.equal(out["1"], out["0"])


@pytest.mark.parametrize("prec,n,nb,pivot", [("fp8", 8192, 1024, True), ("fp8", 8192, 1024, False),
                                             ("fp8", 7168, 512, True), ("bf16", 8192, 1024, True),
                                             ("fp16", 7168, 512, False)])
def test_wide_tiles_in_the_lu(dev, orc, prec, n, nb, pivot):
    """The low-precision LU with groups of 2 steps (rest updates of K = 2 nb,
    FP8 with one scale pair per step, over up to (n - 2 nb)^2 elements)
    gives bit-identical factors and pivots with the wide update tiles and
    without (ragged edges: test_wide_tiles_match_narrow)."""
    A = orc.gen_diagdom(orc.SEED_A + 19, n)
    W, ipw, infw = _getrf_env(dev, A, nb, prec, {"CHARMLU_GROUP": "2", "CHARMLU_LP_WIDE": "1"},
                              pivot=pivot)
    X, ipx, infx = _getrf_env(dev, A, nb, prec, {"CHARMLU_GROUP": "2", "CHARMLU_LP_WIDE": "0"},
                              pivot=pivot)
    assert infw == infx == 0
    assert np.array_equal(ipw, ipx)
    assert np.array_equal(W.cpu().numpy(), X.cpu().numpy())


def _getrf_env(dev, A, nb, prec, env, pivot=True):
    import os
    import torch
    A32 = torch.from_numpy(np.ascontiguousarray(A, dtype=np.float32)).cuda()
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        ipiv, info = dev.getrf_mixed(A32, nb, prec, pivot=pivot)
        torch.cuda.synchronize()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return A32, ipiv.cpu().numpy(), int(info.item())


@pytest.mark.parametrize("prec,n,nb", [("bf16", 2048, 256), ("fp8", 4096, 512),
                                       ("fp16", 1536, 512), ("bf16", 2560, 512)])
def test_speculative_pivots_on_diagdom(dev, orc, prec, n, nb):
    """The pivoted low-precision form eliminates each panel without
    interchanges and keeps the result when every multiplier is below 1/2
    (partial pivoting then keeps every diagonal row).  On a diagonally
    dominant matrix: the identity pivots, as the exchanging panels find
    (CHARMLU_SPECULATE=0), |l_ij| < 1/2, the update's backward-error bound,
    and the refined x within 1e-10 of the FP64 oracle's."""
    import torch
    A = orc.gen_diagdom(orc.SEED_A + 11, n)
    S, ipiv, info = _getrf_env(dev, A, nb, prec, {})
    X, ipiv_x, info_x = _getrf_env(dev, A, nb, prec, {"CHARMLU_SPECULATE": "0"})
    assert info == info_x == 0
    assert np.array_equal(ipiv, np.arange(n)) and np.array_equal(ipiv_x, np.arange(n))
    LU = S.cpu().numpy()
    assert np.abs(np.tril(LU, -1)).max() < 0.5
    unit = 2.0 ** -4 if prec == "fp8" else UNIT[prec]
    be = _backward_error(A, LU, ipiv)
    assert be < 8 * unit * (n // nb + 1) * max(1, nb // 512), be
    b = orc.gen_rhs(n)
    x_ref = orc.lu_solve_problem(n, nb, A=A.copy())["x"]
    dA = torch.from_numpy(A).cuda()
    x, st = dev.gesv_ir(dA, S, torch.from_numpy(ipiv).cuda(), torch.from_numpy(b).cuda(), tol=1.0)
    err = np.abs(x.cpu().numpy() - x_ref).max() / np.abs(x_ref).max()
    print(f"speculative {prec} n={n} nb={nb}: be {be:.2e} {st} err {err:.2e}")
    assert st["residual"] <= 1.0 and err < 1e-10


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
def test_speculative_falls_back_at_the_first_panel(dev, orc, prec):
    """A uniform matrix needs interchanges in its first column: the
    speculative elimination of panel 0 fails its check, the diagonal block
    is restored and the exchanging schedule runs from step 0, so factors
    and pivots are bit-identical to CHARMLU_SPECULATE=0."""
    n, nb = 2048, 256
    A = orc.gen_uniform(orc.SEED_A + 13, n)
    S, ipiv, info = _getrf_env(dev, A, nb, prec, {})
    X, ipiv_x, info_x = _getrf_env(dev, A, nb, prec, {"CHARMLU_SPECULATE": "0"})
    assert info == info_x == 0
    assert not np.array_equal(ipiv, np.arange(n))
    assert np.array_equal(ipiv, ipiv_x)
    assert np.array_equal(S.cpu().numpy(), X.cpu().numpy())


@pytest.mark.parametrize("prec,k", [("bf16", 3), ("bf16", 5), ("bf16", 2), ("bf16", 7),
                                    ("fp8", 7)])
def test_speculative_falls_back_mid_factorization(dev, orc, prec, k):
    """Diagonally dominant except the trailing block from block k on (its
    diagonal shrunk to 1e-3), whose columns need interchanges: panels 0 ..
    k-1 keep the identity (their multipliers are tiny), panel k fails the
    check and the group's owed updates are applied before the exchanging
    schedule takes over (k = 3: the fourth step of a group of 4, k = 5 and
    2: the second and third of one, k = 7: the last panel of the matrix;
    FP8 only there, whose factors of a random block precondition poorly).
    A partial-pivoting factor: identity pivots before block k, |l_ij| <= 1,
    the update's backward-error bound, and refinement to the FP64 oracle's
    x."""
    import torch
    n, nb = 2048, 256
    A = orc.gen_diagdom(orc.SEED_A + 17, n)
    idx = np.arange(k * nb, n)
    A[idx, idx] = 1e-3
    S, ipiv, info = _getrf_env(dev, A, nb, prec, {})
    assert info == 0
    assert np.array_equal(ipiv[:k * nb], np.arange(k * nb))
    assert not np.array_equal(ipiv[k * nb:], np.arange(k * nb, n))
    LU = S.cpu().numpy()
    assert np.abs(np.tril(LU, -1)).max() <= 1.0 + 1e-6
    unit = 2.0 ** -4 if prec == "fp8" else UNIT[prec]
    be = _backward_error(A, LU, ipiv)
    assert be < 64 * unit * np.sqrt(n), be
    b = orc.gen_rhs(n)
    x_ref = orc.lu_solve_problem(n, nb, A=A.copy())["x"]
    dA = torch.from_numpy(A).cuda()
    x, st = dev.gesv_ir(dA, S, torch.from_numpy(ipiv).cuda(), torch.from_numpy(b).cuda(), tol=1.0)
    err = np.abs(x.cpu().numpy() - x_ref).max() / np.abs(x_ref).max()
    print(f"fallback {prec} k={k}: be {be:.2e} {st} err {err:.2e}")
    assert st["residual"] <= 1.0 and err < 1e-10
