"""Differentiable linear solve X = A^{-1} B over the C ABI.

``solve(A, B)`` factors A once with the device LU (``charmlu_dgetrf`` for
FP64, ``charmlu_sgetrf_mixed`` with an FP32 update for FP32) and solves
with ``charmlu_dgetrs`` / ``charmlu_sgetrs``.  The backward pass reuses the
stored factors and pivots, without refactoring:

    A^T Lambda = G          (the transposed solve from the same factors)
    dB = Lambda
    dA = -Lambda X^T        (one GEMM with alpha = -1, written directly)

The reference has no derivative; the forward solve is its factor + solve
(lu.ci:31-370, 387-458).  All arithmetic runs in libcharmlu_amd.so; torch
holds the device memory and supplies the stream.
"""
from __future__ import annotations

import torch

from . import _lib
from ._lib import check
from .lu import _ptr, _stream_ptr, getrf, getrf_mixed, getrs, getrs32


def pick_block(n: int) -> int:
    """The largest supported block size dividing n (n % block == 0 is the
    reference's requirement, benchmark.C:143-147)."""
    for b in (256, 192, 128, 96, 64, 48, 32, 16, 8, 4, 2, 1):
        if n % b == 0:
            return b
    return 1


def _factor(A, block):
    LU = A.detach().clone(memory_format=torch.contiguous_format)
    if LU.dtype == torch.float64:
        ipiv, info = getrf(LU, block)
    else:
        ipiv, info = getrf_mixed(LU, block, "fp32")
    if int(info.item()) != 0:
        raise torch.linalg.LinAlgError(
            f"singular pivot at column {int(info.item()) - 1} (|u_jj| <= 100 eps)")
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


class LUSolve(torch.autograd.Function):
    """X = A^{-1} B with A (n x n) and B (n or n x k), FP64 or FP32, on the
    GPU; differentiable in A and B."""

    @staticmethod
    def forward(ctx, A, B, block):
        if A.dim() != 2 or A.shape[0] != A.shape[1]:
            raise ValueError("A must be square")
        if B.shape[0] != A.shape[0] or B.dim() not in (1, 2):
            raise ValueError("B must be n or n x k")
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


def solve(A, B, block: int | None = None):
    """Differentiable X = A^{-1} B (see ``LUSolve``)."""
    return LUSolve.apply(A, B, pick_block(A.shape[0]) if block is None else block)
