"""The factorisation replayed from a captured HIP graph (both streams of the
lookahead, assembly of A's values included) gives exactly the factor of
direct launches, on every replay, for changing values of one pattern
(sparspaklu! refactorisation, SparseCSCInterface.jl:247-268), with other
device work and synchronisations in between."""
import os

import numpy as np
import pytest

from helpers import backward_error
from sparspak_hip import SparseSolver, findorder, symbolicfactor
from sparspak_hip import _lib
from sparspak_hip.gallery import convdiff3d, uniform

pytestmark = pytest.mark.gpu

PREC = {"fp64": 0, "fp32_bf16x3": 4}


def _factor(L, h, nz, prec):
    _lib.check(L.spkhip_inmatrix(h, _lib.pd(nz)), "inmatrix")
    _lib.check(L.spkhip_factor(h, PREC[prec]), "factor")


@pytest.mark.parametrize("prec", ["fp64", "fp32_bf16x3"])
def test_factor_graph_replays_match_direct_launches(prec):
    print(f"case factor graph convdiff3d 20^3 {prec}", flush=True)
    A = convdiff3d(20).tocsc()
    n = A.shape[0]
    s = SparseSolver(A)
    findorder(s)
    symbolicfactor(s)
    s.slvr._upload()
    L, h = s.slvr._dev.L, s.slvr._dev.h
    v1 = np.ascontiguousarray(A.data, dtype=np.float64)
    v2 = v1 * (1.0 + 0.25 * uniform(77, v1.size, -1.0, 1.0))  # same pattern, new values
    old = os.environ.get("SPKHIP_FACTOR_GRAPH")
    try:
        os.environ["SPKHIP_FACTOR_GRAPH"] = "0"
        _factor(L, h, v1, prec)
        ref1 = s.slvr.download()
        _factor(L, h, v2, prec)
        ref2 = s.slvr.download()
        os.environ["SPKHIP_FACTOR_GRAPH"] = "1"
        _factor(L, h, v1, prec)      # direct (first use of the key)
        _factor(L, h, v1, prec)      # captured, then launched
        for a, c in zip(ref1, s.slvr.download()):
            assert np.array_equal(a, c), "captured factor differs from direct launches"
        b = A @ uniform(78, n, -1.0, 1.0)
        for rep in range(3):
            for v, ref in ((v2, ref2), (v1, ref1)):
                _factor(L, h, v, prec)
                got = s.slvr.download()
                for a, c in zip(ref, got):
                    assert np.array_equal(a, c), f"replay {rep} differs from direct launches"
                # eager device work between replays: a refined solve
                x = b.copy()
                rc = L.spkhip_solve_refine(h, 1, _lib.pd(x), n, 1e-14, 10)
                _lib.check(rc, "solve_refine")
                if v is v1:
                    assert backward_error(A, x, b) <= 1e-14
    finally:
        if old is None:
            os.environ.pop("SPKHIP_FACTOR_GRAPH", None)
        else:
            os.environ["SPKHIP_FACTOR_GRAPH"] = old
