"""Write the LAPACK golden fixtures the oracle is pinned against.

The reference holds no fixtures or golden vectors for the sweep (its checks
are MATLAB scripts over data it does not ship, srcPerformanceTest/
checkSolution_60.m:8-69).  The arithmetic it delegates to is LAPACK
zgetrf/zgetrs (KrylovROMSubstructure.cpp:1428-1430), so the fixtures come from
an independent LAPACK: the OpenBLAS bundled with scipy (scipy.linalg.lapack):
zgetrf, zgetrs 'N'/'T'/'C', and their single-precision counterparts cgetrf
and cgetrs 'N'/'T'/'C' (the factor and solves of the refined mode).

    python tests/golden/gen_golden.py     -> tests/golden/lapack_small.npz
"""
import os

import numpy as np
from scipy.linalg import lapack

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20240601)
    out = {}
    for n in (1, 2, 3, 5, 8, 17, 33, 60):
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        B = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
        lu, piv, info = lapack.zgetrf(A)
        assert info == 0
        out[f"A_{n}"] = A
        out[f"B_{n}"] = B
        out[f"LU_{n}"] = lu
        out[f"ipiv_{n}"] = (piv + 1).astype(np.int32)  # LAPACK 1-based
        for t, code in (("N", 0), ("T", 1), ("C", 2)):
            x, info = lapack.zgetrs(lu, piv, B, trans=code)
            assert info == 0
            out[f"X{t}_{n}"] = x
        A32 = A.astype(np.complex64)
        lu32, piv32, info = lapack.cgetrf(A32)
        assert info == 0
        out[f"LU32_{n}"] = lu32
        out[f"ipiv32_{n}"] = (piv32 + 1).astype(np.int32)
        # cgetrs on the cgetrf factors, 'N' / 'T' / 'C'
        for t, code in (("N", 0), ("T", 1), ("C", 2)):
            x, info = lapack.cgetrs(lu32, piv32, B.astype(np.complex64), trans=code)
            assert info == 0
            out[f"X32{t}_{n}"] = x
    # an exactly zero second column: U(2,2) = 0 in any rounding, zgetrf info = 2
    S = np.array([[1.0, 0.0, 3.0], [2.0, 0.0, 1.0], [3.0, 0.0, 2.0]], dtype=np.complex128)
    lu, piv, info = lapack.zgetrf(S)
    out["S_sing"] = S
    out["LU_sing"] = lu
    out["ipiv_sing"] = (piv + 1).astype(np.int32)
    out["info_sing"] = np.array(info, dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, "lapack_small.npz"), **out)


if __name__ == "__main__":
    main()
