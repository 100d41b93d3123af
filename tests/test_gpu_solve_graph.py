"""The solve sweeps replayed from a captured HIP graph give exactly the
results of direct launches, on every replay, with other device work and
synchronisations in between (refinement residuals, new right-hand sides)."""
import os

import numpy as np
import pytest

from helpers import backward_error
from sparspak_hip import (SparseSolver, factor, findorder, inmatrix, symbolicfactor,
                          triangularsolve)
from sparspak_hip.gallery import convdiff3d, uniform

pytestmark = pytest.mark.gpu


def _solve(s, b, refine):
    x = b.copy()
    triangularsolve(s, x, refine=refine, tol=1e-14)
    return x


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replays_match_direct_launches(precision):
    print(f"case solve graph convdiff3d 16^3 {precision}", flush=True)
    A = convdiff3d(16)
    n = A.shape[0]
    s = SparseSolver(A)
    findorder(s)
    symbolicfactor(s)
    inmatrix(s)
    factor(s, precision)
    rhs = [A @ uniform(70 + i, n, -1, 1) for i in range(4)]
    old = os.environ.get("SPKHIP_SOLVE_GRAPH")
    try:
        os.environ["SPKHIP_SOLVE_GRAPH"] = "0"
        direct = [_solve(s, b, precision != "fp64") for b in rhs]
        os.environ["SPKHIP_SOLVE_GRAPH"] = "1"
        for rep in range(3):
            replay = [_solve(s, b, precision != "fp64") for b in rhs]
            for b, x, y in zip(rhs, direct, replay):
                assert np.array_equal(x, y)
                assert backward_error(A, y, b) <= 1e-14
    finally:
        if old is None:
            os.environ.pop("SPKHIP_SOLVE_GRAPH", None)
        else:
            os.environ["SPKHIP_SOLVE_GRAPH"] = old
