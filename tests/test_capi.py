"""CPU checks of the C ABI: the library loads, exports every function that
include/*.h declares, and rejects bad arguments before touching a device.
No compute call runs here (there is no GPU on this host)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(charmlu_\w+)\s*\(", txt, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    from charmlu_amd import _lib
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding types every one of them
    assert names <= set(_lib.SIGNATURES), names - set(_lib.SIGNATURES)


def test_nm_exports_are_c_symbols():
    import subprocess
    from charmlu_amd import _lib
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    for n in declared_functions():
        assert n in exported, n


def test_argument_errors_before_any_device_call():
    from charmlu_amd import _lib
    lib = _lib.load()
    # n < 0, block not dividing n (benchmark.C:143-147), bad lda
    assert lib.charmlu_dgetrf(-1, 8, None, 1, None, None, None) == -1
    assert lib.charmlu_dgetrf(30, 8, None, 30, None, None, None) == -2
    assert lib.charmlu_dgetrf(32, 0, None, 32, None, None, None) == -2
    assert lib.charmlu_dgetrf(32, 8, ctypes.c_void_p(16), 16, None, None, None) == -4
    assert lib.charmlu_dgetrs(7, 4, 1, None, 4, None, None, 1, None) == -1
    assert lib.charmlu_dgetrs(_lib.NO_TRANS, 4, -1, None, 4, None, None, 1, None) == -3
    assert lib.charmlu_cblas_dgemm(7, _lib.NO_TRANS, _lib.NO_TRANS, 1, 1, 1, 1.0, None, 1, None,
                                   1, 0.0, None, 1) == -1
    assert lib.charmlu_cblas_dgemm(_lib.ROW_MAJOR, 5, _lib.NO_TRANS, 1, 1, 1, 1.0, None, 1, None,
                                   1, 0.0, None, 1) == -2
    assert lib.charmlu_cblas_dtrsm(_lib.ROW_MAJOR, _lib.LEFT, 9, _lib.NO_TRANS, _lib.UNIT, 2, 2,
                                   1.0, None, 2, None, 2) == -3
    assert lib.charmlu_cblas_dger(_lib.ROW_MAJOR, 2, 2, 1.0, None, 0, None, 1, None, 2) == -6
    assert lib.charmlu_dgen_matrix(5, 1, 4, None, 4, None) == -1
    assert lib.charmlu_dresidual(-1, None, 1, None, None, None, None, None) == -1
    # mixed precision: FP8 and FP64 are not trailing-update precisions here
    assert lib.charmlu_sgetrf_mixed(32, 8, _lib.PREC_FP64, None, 32, None, None, None) == -3
    assert lib.charmlu_sgetrf_mixed(32, 7, _lib.PREC_BF16, None, 32, None, None, None) == -2
    assert lib.charmlu_sgetrs(_lib.NO_TRANS, 4, -1, None, 4, None, None, 1, None) == -3
    assert lib.charmlu_dgesv_ir(4, ctypes.c_void_p(16), 4, ctypes.c_void_p(16), 4,
                                ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.c_void_p(16),
                                0.0, 5, 10, None, None) == -9
    assert lib.charmlu_dgesv_ir(4, ctypes.c_void_p(16), 4, ctypes.c_void_p(16), 4,
                                ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.c_void_p(16),
                                1.0, 5, 0, None, None) == -11


def test_empty_problems_are_no_ops():
    from charmlu_amd import _lib
    lib = _lib.load()
    # M = N = 0: nothing to do, no device touched
    assert lib.charmlu_dgemm_dev(_lib.NO_TRANS, _lib.NO_TRANS, 0, 0, 0, 1.0, None, 1, None, 1,
                                 0.0, None, 1, None) == 0
    assert lib.charmlu_dgetrs(_lib.NO_TRANS, 0, 1, None, 1, None, None, 1, None) == 0


def test_product_has_no_oracle_dependency():
    # the product package must not import, load or call the oracle
    for path in glob.glob(os.path.join(ROOT, "charmlu_amd", "**", "*"), recursive=True):
        if os.path.isfile(path) and path.endswith((".py", ".hip", ".h", ".cpp")):
            txt = open(path, errors="replace").read()
            for needle in ("import oracle", "from oracle", "liboracle", "lu_oracle", "orc_"):
                assert needle not in txt, (path, needle)


# ---------------------------------------------------------------------------
# panel strip launch shapes (host arithmetic of charmlu_panel_plan)
# ---------------------------------------------------------------------------
def _plan(prec, m, w, cus, ncu=256, pivot=1):
    from charmlu_amd import _lib
    out = (ctypes.c_int * 6)()
    rc = _lib.load().charmlu_panel_plan(prec, m, w, cus, ncu, pivot, out)
    return rc, list(out)


def _poll_words(nt):
    return (nt // 64) * 64 * (4 if nt >= 512 else 5)


# Workgroups of a one-per-CU kernel resident at once on the top `cus` CUs
# of MI355X (256 CUs, 8 XCDs of 4 shader engines), as measured with a
# bounded-sleep probe for 16, 24, 32 and 48 (16, 24, 32 and 32: a 48-CU mask
# holds 1 or 2 CUs per engine, and workgroups are dealt round-robin over the
# engines); 0 is the unmasked layout's cap.
RESIDENT = {0: 32, 8: 8, 16: 16, 20: 16, 24: 24, 32: 32, 40: 32, 48: 32, 56: 32, 64: 64}


def test_resident_workgroups_follow_the_measured_layout():
    for cus, want in RESIDENT.items():
        assert _plan(1, 1024, 16, cus)[1][4] == want, cus


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("cus", [0, 8, 16, 24, 32, 48, 64])
def test_panel_plan_never_exceeds_resident_workgroups(prec, cus):
    """Every strip shape the library launches keeps its workgroups within
    what can be resident at once, covers the strip's rows, and lets the
    polling waves cover all records in one round; and whenever the tallest
    strip (the first, n rows) has a shape, every shorter one has too, so
    the factorizations' up-front check suffices."""
    from charmlu_amd import _lib
    vw = 2 if prec == 0 else 1
    resident = RESIDENT[cus]
    widest_prev = None
    for m in sorted(set(range(1, 4200, 37)) | set(range(4096, 300000, 4093))):
        rc, o = _plan(prec, m, 8, cus)
        widest = o[5]
        assert o[4] == resident
        if widest == 0:
            assert rc == _lib.ERR_UNSUPPORTED
        else:
            assert rc == 0
        # shorter strips never lose their shape
        if widest_prev == 0:
            assert widest == 0, (m, cus)
        widest_prev = widest
        for w in (8, 16, 32):
            rc, o = _plan(prec, m, w, cus)
            if w > widest or m < 1:
                assert rc == _lib.ERR_UNSUPPORTED
                continue
            assert rc == 0, (m, w, cus)
            sw, R, NT, G = o[:4]
            assert sw >= w and R * NT * G >= m and R * NT * (G - 1) < m
            assert G <= resident
            if G > 1:
                assert G * (sw * vw + vw + 1) + sw * vw <= _poll_words(NT)


def test_panel_plan_known_shapes():
    """The shapes of the measured default (32 panel CUs) and the fallbacks."""
    from charmlu_amd import _lib
    FP64, FP32 = _lib.PREC_FP64, _lib.PREC_FP32
    assert _plan(FP64, 32768, 16, 32)[1][:4] == [16, 4, 256, 32]   # C2's first strips
    assert _plan(FP64, 65536, 16, 32)[1][:4] == [16, 4, 512, 32]
    assert _plan(FP64, 8192, 32, 32)[1][:4] == [32, 2, 256, 16]    # 32-column FP64 strips
    assert _plan(FP64, 131072, 8, 32)[1][:4] == [8, 8, 512, 32]    # beyond 65536 rows
    assert _plan(FP64, 131072 + 256, 8, 32)[1][5] == 0             # too tall: rejected
    assert _plan(FP32, 65536, 32, 32)[1][:4] == [32, 4, 512, 32]   # C3's first strips
    assert _plan(FP32, 131072, 16, 32)[1][:4] == [16, 8, 512, 32]  # C5's first strips
    assert _plan(FP32, 262144, 8, 32)[1][:4] == [8, 16, 512, 32]
    # 16 panel CUs at n = 9216: 32-column FP64 strips would need 18
    # workgroups, so 16-column ones run
    rc, o = _plan(FP64, 9216, 32, 16)
    assert rc == _lib.ERR_UNSUPPORTED and o[5] == 16
    assert _plan(FP64, 9216, 16, 16)[1][:4] == [16, 4, 256, 9]
    # 8 panel CUs at C3 (FP32, 65536 rows): 8-column strips of 8192 rows each
    assert _plan(FP32, 65536, 8, 8)[1][:4] == [8, 16, 512, 8]
    # unpivoted strips are one workgroup: 256 threads up to 1024 rows, 512
    # threads up to 2048 (the diagonal block of block size 2048)
    assert _plan(FP32, 1024, 32, 32, pivot=0)[1][:4] == [32, 4, 256, 1]
    assert _plan(FP32, 1025, 32, 32, pivot=0)[1][:4] == [32, 4, 512, 1]
    assert _plan(FP32, 2048, 16, 32, pivot=0)[1][:4] == [16, 4, 512, 1]
    assert _plan(FP32, 2049, 16, 32, pivot=0)[0] == _lib.ERR_UNSUPPORTED
    assert _plan(FP64, 512, 32, 32, pivot=0)[0] == _lib.ERR_UNSUPPORTED  # FP64: 16 columns
    assert _plan(FP64, 512, 16, 32, pivot=0)[1][:4] == [16, 2, 256, 1]
    # argument errors
    assert _plan(7, 100, 8, 32)[0] == -1
    assert _plan(FP64, 0, 8, 32)[0] == -2
    assert _plan(FP64, 100, 0, 32)[0] == -3
    assert _plan(FP64, 100, 8, 65)[0] == -4


def test_dgesv_ir_ex_argument_errors():
    from charmlu_amd import _lib
    lib = _lib.load()
    p = ctypes.c_void_p(16)
    args = lambda **kw: [kw.get("trans", _lib.NO_TRANS), kw.get("n", 4), kw.get("nrhs", 2), p,
                         kw.get("lda", 4), p, 4, p, p, kw.get("ldb", 2), p, kw.get("ldx", 2),
                         kw.get("tol", 1.0), 5, kw.get("restart", 10), None, None]
    assert lib.charmlu_dgesv_ir_ex(*args(trans=9)) == -1
    assert lib.charmlu_dgesv_ir_ex(*args(n=-1)) == -2
    assert lib.charmlu_dgesv_ir_ex(*args(nrhs=-1)) == -3
    assert lib.charmlu_dgesv_ir_ex(*args(lda=3)) == -5
    assert lib.charmlu_dgesv_ir_ex(*args(ldb=1)) == -10
    assert lib.charmlu_dgesv_ir_ex(*args(ldx=1)) == -12
    assert lib.charmlu_dgesv_ir_ex(*args(tol=0.0)) == -13
    assert lib.charmlu_dgesv_ir_ex(*args(restart=0)) == -15
    # nothing to solve: no device touched
    assert lib.charmlu_dgesv_ir_ex(*args(n=0)) == 0
    assert lib.charmlu_dgesv_ir_ex(*args(nrhs=0)) == 0
