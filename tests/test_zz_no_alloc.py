"""charmlu_gemm_fp8_dev_noalloc: the FP8 block op on the library's static
device scratch array (no allocation) gives the same bits as
charmlu_gemm_fp8_dev, which quantizes into the library's workspace, through
the C ABI -- two calls in a row, the second on a smaller problem (the array
keeps the first call's contents, so a stale byte would show), plus the
size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _run(lib, fn, M, N, K, alpha, beta, A, B, C):
    import torch
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(M, N, K, alpha, _p(A), A.stride(0), _p(B), B.stride(0), beta, _p(C), C.stride(0), s)
    torch.cuda.synchronize()
    return rc, lib.charmlu_last_error().decode()


def test_noalloc_fp8_gemm_matches_workspace_form(dev):
    import torch
    from charmlu_amd import _lib
    lib = _lib.load()
    g = torch.Generator(device="cpu").manual_seed(11)
    # (M, N, K): the first fills the scratch with its operands, the second
    # is smaller and ragged (K not a multiple of 64, partial 128-row tiles)
    for M, N, K, alpha, beta in [(3000, 2500, 1024, -1.0, 1.0), (700, 333, 200, 0.5, -2.0)]:
        A = (torch.rand(M, K, generator=g, dtype=torch.float32) - 0.5).cuda()
        B = (torch.rand(K, N, generator=g, dtype=torch.float32) - 0.5).cuda() * 3.0
        C0 = (torch.rand(M, N, generator=g, dtype=torch.float32) - 0.5).cuda()
        C1, C2 = C0.clone(), C0.clone()
        print(f"fp8 gemm M={M} N={N} K={K}: workspace form", flush=True)
        rc, msg = _run(lib, lib.charmlu_gemm_fp8_dev, M, N, K, alpha, beta, A, B, C1)
        assert rc == 0, msg
        print(f"fp8 gemm M={M} N={N} K={K}: no-allocation form", flush=True)
        rc, msg = _run(lib, lib.charmlu_gemm_fp8_dev_noalloc, M, N, K, alpha, beta, A, B, C2)
        assert rc == 0, msg
        assert torch.equal(C1, C2), (M, N, K)
        # and the result is the FP8 product: within the e4m3 rounding of the
        # operands (2^-4 relative per operand) of the FP64 product
        ref = alpha * (A.double() @ B.double()) + beta * C0.double()
        bound = 2 * 2.0 ** -4 * abs(alpha) * (A.double().abs() @ B.double().abs()) + 1e-5
        assert bool(((C2.double() - ref).abs() <= bound).all())


def test_noalloc_rejects_problems_beyond_the_scratch(dev):
    import torch
    from charmlu_amd import _lib
    lib = _lib.load()
    # (M + N) * Kp alone exceeds 32 MiB: rejected before any launch, so the
    # small stand-in buffers are never read
    A = torch.zeros(16, dtype=torch.float32, device="cuda")
    C = torch.full((16,), 7.0, dtype=torch.float32, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    M = N = 16384
    K = 1088
    rc = lib.charmlu_gemm_fp8_dev_noalloc(M, N, K, 1.0, _p(A), K, _p(A), N, 1.0, _p(C), N, s)
    torch.cuda.synchronize()
    assert rc == _lib.ERR_UNSUPPORTED
    assert "scratch" in lib.charmlu_last_error().decode()
    assert bool((C == 7.0).all())
