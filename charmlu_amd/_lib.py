"""ctypes binding of libcharmlu_amd.so (the C ABI of include/charmlu_amd.h).

The library must have been built (``python charmlu_amd/build.py`` or
``__graft_entry__.build()``); there is no fallback path.
"""
from __future__ import annotations

import atexit
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcharmlu_amd.so")

_c_int, _c_long, _c_double, _vp = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p
_c_u64 = ctypes.c_uint64

# name -> (restype, argtypes); every symbol declared in include/charmlu_amd.h
SIGNATURES = {
    "charmlu_version": (_c_int, []),
    "charmlu_last_error": (ctypes.c_char_p, []),
    "charmlu_init": (_c_int, [_c_int]),
    "charmlu_finalize": (_c_int, []),
    "charmlu_panel_probe": (_c_int, [_c_int, _c_int, _vp, _vp]),
    "charmlu_panel_plan": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                                    ctypes.POINTER(ctypes.c_int)]),
    "charmlu_profile": (_c_int, [_c_int]),
    "charmlu_profile_read": (_c_int, [ctypes.POINTER(ctypes.c_double), _c_int]),
    "charmlu_cblas_dgemm": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double,
                                     _vp, _c_int, _vp, _c_int, _c_double, _vp, _c_int]),
    "charmlu_cblas_dtrsm": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                                     _c_double, _vp, _c_int, _vp, _c_int]),
    "charmlu_cblas_dger": (_c_int, [_c_int, _c_int, _c_int, _c_double, _vp, _c_int, _vp, _c_int,
                                    _vp, _c_int]),
    "charmlu_cblas_dgemv": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_double, _vp, _c_int, _vp,
                                     _c_int, _c_double, _vp, _c_int]),
    "charmlu_dgemm_dev": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _vp, _c_long,
                                   _vp, _c_long, _c_double, _vp, _c_long, _vp]),
    "charmlu_dtrsm_dev": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _vp, _c_long, _vp,
                                   _c_long, _vp]),
    "charmlu_dger_dev": (_c_int, [_c_int, _c_int, _c_double, _vp, _c_long, _vp, _c_long, _vp,
                                  _c_long, _vp]),
    "charmlu_dgetrf": (_c_int, [_c_int, _c_int, _vp, _c_long, _vp, _vp, _vp]),
    "charmlu_dgetrs": (_c_int, [_c_int, _c_int, _c_int, _vp, _c_long, _vp, _vp, _c_long, _vp]),
    "charmlu_sgetrf_mixed": (_c_int, [_c_int, _c_int, _c_int, _vp, _c_long, _vp, _vp, _vp]),
    "charmlu_dlag2s": (_c_int, [_c_long, _c_long, _vp, _c_long, _vp, _c_long, _vp]),
    "charmlu_sgemm_dev": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, ctypes.c_float, _vp,
                                   _c_long, _vp, _c_long, ctypes.c_float, _vp, _c_long, _vp]),
    "charmlu_gemm_lp_dev": (_c_int, [_c_int, _c_int, _c_int, _c_int, ctypes.c_float, _vp, _c_long,
                                     _vp, _c_long, ctypes.c_float, _vp, _c_long, _vp]),
    "charmlu_gemm_fp8_dev": (_c_int, [_c_int, _c_int, _c_int, ctypes.c_float, _vp, _c_long, _vp,
                                      _c_long, ctypes.c_float, _vp, _c_long, _vp]),
    "charmlu_gemm_fp8_dev_noalloc": (_c_int, [_c_int, _c_int, _c_int, ctypes.c_float, _vp, _c_long,
                                             _vp, _c_long, ctypes.c_float, _vp, _c_long, _vp]),
    "charmlu_sgetrs": (_c_int, [_c_int, _c_int, _c_int, _vp, _c_long, _vp, _vp, _c_long, _vp]),
    "charmlu_dgesv_ir": (_c_int, [_c_int, _vp, _c_long, _vp, _c_long, _vp, _vp, _vp,
                                  ctypes.c_double, _c_int, _c_int, _vp, _vp]),
    "charmlu_dgesv_ir_ex": (_c_int, [_c_int, _c_int, _c_int, _vp, _c_long, _vp, _c_long, _vp, _vp,
                                     _c_long, _vp, _c_long, ctypes.c_double, _c_int, _c_int, _vp,
                                     _vp]),
    "charmlu_dgen_matrix": (_c_int, [_c_int, _c_u64, _c_long, _vp, _c_long, _vp]),
    "charmlu_dgen_rhs": (_c_int, [_c_int, _c_long, _c_long, _vp, _vp]),
    "charmlu_dresidual": (_c_int, [_c_int, _vp, _c_long, _vp, _vp, _vp, _vp, _vp]),
}

# CBLAS / ABI constants (include/charmlu_amd.h)
ROW_MAJOR, COL_MAJOR = 101, 102
NO_TRANS, TRANS = 111, 112
UPPER, LOWER = 121, 122
NON_UNIT, UNIT = 131, 132
LEFT, RIGHT = 141, 142
PREC_FP64, PREC_FP32, PREC_BF16, PREC_FP16, PREC_FP8 = 0, 1, 2, 3, 4
ERR_HIP, ERR_NOMEM, ERR_CONFIG, ERR_UNSUPPORTED = -100, -101, -102, -103
MIXED_NOPIV = 0x100

_lib = None


class CharmluError(RuntimeError):
    pass


def load():
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python charmlu_amd/build.py` "
                "(there is no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        # streams, events and workspaces go back before the HIP runtime
        # tears down at process exit
        atexit.register(lib.charmlu_finalize)
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().charmlu_last_error().decode(errors="replace")
        raise CharmluError(f"{what} returned {rc}: {msg}")
