"""charmlu_amd: MI355X-native dense LU with partial pivoting.

A device-resident re-design of UIUC-PPL/charmlu's block LU: hand-written
gfx950 HIP kernels behind the C ABI of ``include/charmlu_amd.h``
(``libcharmlu_amd.so``), driven from Python through ctypes.  Importing the
package loads the library and fails loudly if it has not been built.
"""
from . import _lib
from ._lib import CharmluError, load

load()

from .lu import (LUConfig, check_info, gen_matrix, gen_rhs, gesv_ir, gesv_ir_multi,  # noqa: E402
                 getrf, getrf_mixed, getrs, getrs32, hpl_gflops, lag2s, residual, run)

from .autograd import LUSolve, MixedLUSolve, solve  # noqa: E402

__all__ = ["LUSolve", "MixedLUSolve", "solve", "CharmluError", "LUConfig", "check_info",
           "gen_matrix", "gen_rhs", "gesv_ir", "gesv_ir_multi", "getrf", "getrf_mixed", "getrs",
           "getrs32", "hpl_gflops", "lag2s", "residual", "run", "load"]
