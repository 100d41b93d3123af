"""Many right-hand sides in one solve (the reference loops over vectors,
SpkSparseSolver.jl:266): each column of the solution equals the solve of that
column alone, bit for bit (the kernels read each factor entry once for a
block of right-hand sides and give every right-hand side the same operations
in the same order), for LU and SPD factors, double and float, small and
large fronts (more right-hand sides than the four waves of a chain and the
eight of a backward GEMV block). Float factors solve up to three right-hand
sides with the explicit inverses of the diagonal blocks and more by
substitution; each form is compared with itself (SPKHIP_SOLVE_INV selects
substitution for all), and a float solve of four or more right-hand sides
equals the substitution solve bit for bit."""
import os
import sys

import numpy as np
import pytest

from helpers import backward_error
from sparspak_hip import (SparseSolver, SparseSpdSolver, factor, findorder, inmatrix,
                          symbolicfactor, triangularsolve)
from sparspak_hip.gallery import convdiff3d, laplace2d, uniform

pytestmark = pytest.mark.gpu


def _prepared(A, cls, precision):
    s = cls(A)
    findorder(s)
    symbolicfactor(s)
    inmatrix(s)
    factor(s, precision)
    return s


@pytest.mark.parametrize("case", ["lu_fp64", "lu_fp32", "spd_fp64", "spd_fp32"])
def test_many_rhs_equal_single_rhs(case):
    kind, precision = case.split("_")
    print(f"case multi-rhs {case}", flush=True)
    sys.stdout.flush()
    A = convdiff3d(18) if kind == "lu" else laplace2d(90, shift=0.05)
    cls = SparseSolver if kind == "lu" else SparseSpdSolver
    s = _prepared(A, cls, precision)
    n = A.shape[0]
    nrhs = 11
    B = np.asfortranarray(np.stack([uniform(300 + q, n, -1.0, 1.0) for q in range(nrhs)], axis=1))
    old = os.environ.get("SPKHIP_SOLVE_INV")
    try:
        os.environ["SPKHIP_SOLVE_INV"] = "0"
        X = B.copy(order="F")
        triangularsolve(s, X, refine=False)
        for q in range(nrhs):
            x = B[:, q].copy()
            triangularsolve(s, x, refine=False)
            assert np.array_equal(x, X[:, q]), f"column {q} differs from its single solve"
        if precision == "fp32":
            # the default: four or more right-hand sides take the substitution
            # kernels; up to three the inverses, column by column alike
            os.environ.pop("SPKHIP_SOLVE_INV")
            Xd = B.copy(order="F")
            triangularsolve(s, Xd, refine=False)
            assert np.array_equal(Xd, X)
            X3 = np.array(B[:, :3], order="F")  # a copy: B stays the right-hand sides
            triangularsolve(s, X3, refine=False)
            for q in range(3):
                x = B[:, q].copy()
                triangularsolve(s, x, refine=False)
                assert np.array_equal(x, X3[:, q]), f"column {q} differs from its single solve"
    finally:
        if old is None:
            os.environ.pop("SPKHIP_SOLVE_INV", None)
        else:
            os.environ["SPKHIP_SOLVE_INV"] = old
    if precision == "fp64":
        assert max(backward_error(A, X[:, q], B[:, q]) for q in range(nrhs)) <= 1e-14
