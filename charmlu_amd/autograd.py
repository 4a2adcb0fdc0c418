"""Differentiable linear solve X = A^{-1} B over the C ABI.

``solve(A, B)`` factors A once with the device LU (``charmlu_dgetrf`` for
FP64, ``charmlu_sgetrf_mixed`` with an FP32 update for FP32) and solves
with ``charmlu_dgetrs`` / ``charmlu_sgetrs``.  The backward pass reuses the
stored factors and pivots, without refactoring:

    A^T Lambda = G          (the transposed solve from the same factors)
    dB = Lambda
    dA = -Lambda X^T        (one GEMM with alpha = -1, written directly)

``solve(A, B, precision=p)`` with FP64 A and B is the mixed-precision form:
the forward pass factors an FP32 copy of A with the trailing update in p
('fp32', 'bf16', 'fp16', or 'fp8') and refines X in FP64 by GMRES-IR
(``charmlu_dgesv_ir_ex``); the backward pass solves A^T Lambda = G by the
same refinement, transposed, preconditioned by the same stored
low-precision factors, so both X and the gradients carry FP64 accuracy.

The reference has no derivative; the forward solve is its factor + solve
(lu.ci:31-370, 387-458).  All arithmetic runs in libcharmlu_amd.so; torch
holds the device memory and supplies the stream.
"""
from __future__ import annotations

import torch

from . import _lib
from ._lib import check
from .lu import (_ptr, _stream_ptr, check_info, gesv_ir_multi, getrf, getrf_mixed, getrs, getrs32,
                 lag2s)

MIXED = ("fp32", "bf16", "fp16", "fp8")


def pick_block(n: int) -> int:
    """The largest supported block size dividing n (n % block == 0 is the
    reference's requirement, benchmark.C:143-147)."""
    for b in (256, 192, 128, 96, 64, 48, 32, 16, 8, 4, 2, 1):
        if n % b == 0:
            return b
    return 1


def pick_block_mixed(n: int, precision: str) -> int:
    """Block of the mixed-precision factorization: the largest of 1024, 512,
    256, 128 dividing n (the FP8 update needs a multiple of 128)."""
    for b in (1024, 512, 256, 128):
        if n % b == 0 and n >= b:
            return b
    if precision == "fp8":
        raise ValueError(f"the FP8 form needs n divisible by 128 (n = {n})")
    return pick_block(n)


def _factor(A, block):
    LU = A.detach().clone(memory_format=torch.contiguous_format)
    if LU.dtype == torch.float64:
        ipiv, info = getrf(LU, block)
    else:
        ipiv, info = getrf_mixed(LU, block, "fp32")
    check_info(info)
    return LU, ipiv


def _solve(LU, ipiv, B, trans):
    X = B.detach().clone(memory_format=torch.contiguous_format)
    if LU.dtype == torch.float64:
        getrs(LU, ipiv, X, trans=trans)
    else:
        getrs32(LU, ipiv, X, trans=trans)
    return X


def _neg_outer(Lam, X):
    """-Lam X^T (n x n) from two n x k (or n) operands."""
    n = Lam.shape[0]
    k = 1 if Lam.dim() == 1 else Lam.shape[1]
    dA = torch.empty((n, n), dtype=Lam.dtype, device=Lam.device)
    lib = _lib.load()
    if Lam.dtype == torch.float64:
        fn, name = lib.charmlu_dgemm_dev, "charmlu_dgemm_dev"
    else:
        fn, name = lib.charmlu_sgemm_dev, "charmlu_sgemm_dev"
    check(fn(_lib.NO_TRANS, _lib.TRANS, n, n, k, -1.0, _ptr(Lam), k, _ptr(X), k, 0.0, _ptr(dA),
             n, _stream_ptr()), name)
    return dA


def _check_inputs(A, B):
    if A.dim() != 2 or A.shape[0] != A.shape[1]:
        raise ValueError("A must be square")
    if B.shape[0] != A.shape[0] or B.dim() not in (1, 2):
        raise ValueError("B must be n or n x k")


class LUSolve(torch.autograd.Function):
    """X = A^{-1} B with A (n x n) and B (n or n x k), FP64 or FP32, on the
    GPU; differentiable in A and B."""

    @staticmethod
    def forward(ctx, A, B, block):
        _check_inputs(A, B)
        if A.dtype not in (torch.float64, torch.float32) or B.dtype != A.dtype:
            raise ValueError("A and B must both be float64 or both float32")
        LU, ipiv = _factor(A, block)
        X = _solve(LU, ipiv, B, trans=False)
        ctx.save_for_backward(LU, ipiv, X)
        return X

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, G):
        LU, ipiv, X = ctx.saved_tensors
        Lam = _solve(LU, ipiv, G.contiguous(), trans=True)
        dA = _neg_outer(Lam, X) if ctx.needs_input_grad[0] else None
        dB = Lam if ctx.needs_input_grad[1] else None
        return dA, dB, None


class MixedLUSolve(torch.autograd.Function):
    """X = A^{-1} B for FP64 A (n x n) and B (n or n x k): low-precision
    factors (FP32 panel, trailing update in ``precision``) and FP64 GMRES-IR
    forward; the adjoint A^T Lambda = G by transposed GMRES-IR from the same
    stored factors backward, then dA = -Lambda X^T."""

    @staticmethod
    def forward(ctx, A, B, block, precision, tol, max_iter):
        _check_inputs(A, B)
        if A.dtype != torch.float64 or B.dtype != torch.float64:
            raise ValueError("the mixed-precision solve takes float64 A and B")
        if precision not in MIXED:
            raise ValueError(f"precision must be one of {MIXED}")
        A64 = A.detach().contiguous()
        LU32 = lag2s(A64)
        ipiv, info = getrf_mixed(LU32, block, precision)
        check_info(info, "mixed-precision factorization")
        X, st = gesv_ir_multi(A64, LU32, ipiv, B.detach().contiguous(), tol=tol,
                              max_iter=max_iter)
        if not st["residual"] <= tol:
            raise _lib.CharmluError(f"refinement did not converge: {st}")
        ctx.save_for_backward(A64, LU32, ipiv, X)
        ctx.tol, ctx.max_iter = tol, max_iter
        ctx.stats = st
        return X

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, G):
        A64, LU32, ipiv, X = ctx.saved_tensors
        Lam, st = gesv_ir_multi(A64, LU32, ipiv, G.contiguous(), trans=True, tol=ctx.tol,
                                max_iter=ctx.max_iter)
        if not st["residual"] <= ctx.tol:
            raise _lib.CharmluError(f"adjoint refinement did not converge: {st}")
        dA = _neg_outer(Lam, X) if ctx.needs_input_grad[0] else None
        dB = Lam if ctx.needs_input_grad[1] else None
        return dA, dB, None, None, None, None


def solve(A, B, block: int | None = None, precision: str | None = None, tol: float = 1.0,
          max_iter: int = 30):
    """Differentiable X = A^{-1} B.  precision None: ``LUSolve`` in the dtype
    of A (FP64 or FP32).  precision 'fp32' / 'bf16' / 'fp16' / 'fp8' (FP64 A
    and B): ``MixedLUSolve``, refined to the reference's scaled residual
    <= tol in both directions."""
    if precision is None:
        return LUSolve.apply(A, B, pick_block(A.shape[0]) if block is None else block)
    b = pick_block_mixed(A.shape[0], precision) if block is None else block
    return MixedLUSolve.apply(A, B, b, precision, tol, max_iter)
