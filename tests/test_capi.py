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
