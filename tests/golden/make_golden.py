"""Writes tests/golden/*.npz from the CPU oracle (oracle/, a restatement of
the reference's block LU, pinned against LAPACK/MKL in tests/test_oracle.py).

The reference has no golden vectors of its own (its only test, test_lu.pl,
checks residual < 16); these fixtures freeze the oracle's outputs on the
reference's own test grid (test_lu.pl:9-52) at the sizes that are cheap,
so later changes to the oracle or the device path are caught.

Usage: python tests/golden/make_golden.py [--large]

--large writes large_fixtures.npz instead: the oracle's pivots, x and
near-tie log for the single-GPU FP64 configuration (n = 32768, block 256)
and for block 1024 on a pivoting matrix (n = 8192), at sizes the oracle
needs minutes for (the GPU tests compare against these files instead of
running the oracle on the GPU host).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

# (n, b, batch) from test_lu.pl:9-45 (mapping arguments are no-ops on one device)
GRID = [
    (48, 48, 0), (48, 24, 0), (48, 16, 0), (48, 12, 0),
    (32, 8, 9), (32, 8, 8), (32, 8, 7), (32, 8, 4), (32, 8, 3), (32, 8, 2), (32, 8, 1),
    (32, 4, 2), (32, 4, 1),
    (512, 128, 128), (512, 64, 128), (512, 256, 0), (512, 128, 0),
]


# (n, b) of the large fixtures: C2 and the FP64 block of the mixed-precision
# comparison (1024) on a uniform, pivoting matrix
LARGE = [(32768, 256), (8192, 1024)]
TIE_WINDOW = 1e-6  # columns whose two best candidates lie this close (relative) are kept


def large(nthreads: int):
    out = {}
    oracle.use_mkl()
    for n, b in LARGE:
        r = oracle.lu_solve_problem(n, b, 0, nthreads=nthreads, ties=True)
        t = r["ties"]
        key = f"n{n}_b{b}"
        out[key + "_ipiv"] = r["ipiv"].astype(np.int32)
        out[key + "_x"] = r["x"]
        out[key + "_res"] = np.array([r["residual"], r["residual_ns"]])
        out[key + "_info"] = np.array([r["info"]], np.int32)
        close = np.nonzero((t["row2"] >= 0) & (t["mag1"] - t["mag2"] <= TIE_WINDOW * t["mag1"]))[0]
        out[key + "_tie_cols"] = close.astype(np.int32)
        out[key + "_tie_row2"] = t["row2"][close].astype(np.int32)
        out[key + "_tie_mag1"] = t["mag1"][close]
        out[key + "_tie_mag2"] = t["mag2"][close]
        print(f"n={n} b={b}: {r['seconds']:.1f} s, info {r['info']}, residual "
              f"{r['residual']:.3f}, {len(close)} columns within {TIE_WINDOW:g}, "
              f"{len(oracle.near_ties(t))} within 4 ulp", flush=True)
    np.savez_compressed(os.path.join(HERE, "large_fixtures.npz"), **out)
    print("wrote", len(out), "arrays")


def main():
    if "--large" in sys.argv:
        large(os.cpu_count() or 1)
        return
    out = {}
    for n, b, batch in GRID:
        r = oracle.lu_solve_problem(n, b, batch)
        key = f"n{n}_b{b}_k{batch}"
        out[key + "_ipiv"] = r["ipiv"].astype(np.int32)
        out[key + "_x"] = r["x"]
        out[key + "_res"] = np.array([r["residual"], r["residual_ns"]])
        if n <= 64:
            out[key + "_lu"] = r["LU"]
    # block ops on seeded inputs (lu.C:36-41, 187, 247-252; benchmark.C:494-496)
    A = oracle.gen_uniform(11, 24, 20)
    B = oracle.gen_uniform(12, 20, 28)
    C = oracle.gen_uniform(13, 24, 28)
    out["gemm_A"], out["gemm_B"], out["gemm_C"] = A, B, C
    out["gemm_out"] = oracle.dgemm(-1.0, A, B, 1.0, C.copy())
    L = oracle.gen_uniform(14, 24, 24)
    X = oracle.gen_uniform(15, 24, 40)
    out["trsm_L"], out["trsm_B"] = L, X
    out["trsm_out"] = oracle.dtrsm_llnu(L, X.copy())
    x = oracle.gen_uniform(16, 1, 24)[0]
    y = oracle.gen_uniform(17, 1, 28)[0]
    out["ger_x"], out["ger_y"] = x, y
    out["ger_out"] = oracle.dger(-1.0, x, y, C.copy())
    v = oracle.gen_uniform(18, 1, 28)[0]
    out["gemv_x"] = v
    out["gemv_out"] = oracle.dgemv(1.0, C, v, 0.0, np.zeros(24))
    out["rhs_4096"] = oracle.gen_rhs(4096)
    out["gen_corner"] = oracle.gen_uniform(oracle.SEED_A, 8, 8)
    np.savez_compressed(os.path.join(HERE, "oracle_fixtures.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
