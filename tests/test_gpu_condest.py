"""Condition estimate of an SPD factor (the reference's commented
sparsespdcondest / sparsespdainverseonenorm, SpkSparseSpdSolver.jl:346-459),
computed on device solves, against (1) the same algorithm restated on the
CPU oracle's solves, (2) the exact cond_1 from a dense inverse (numpy), of
which Hager's estimate is a lower bound (exact on these matrices up to the
solves' rounding)."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle.oracle as orc
from helpers import csc1
from sparspak_hip import (SparseSolver, SparseSpdSolver, condest, factor, findorder, inmatrix,
                          symbolicfactor)
from sparspak_hip.gallery import laplace2d, uniform

pytestmark = pytest.mark.gpu


def _spd_random(n, seed):
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=0.08, random_state=rng, format="csc")
    M = (M + M.T) * 0.5
    d = np.asarray(abs(M).sum(axis=1)).ravel() + rng.uniform(0.01, 0.5, n)
    return sp.csc_matrix(M + sp.diags(d))


@pytest.mark.parametrize("name", ["laplace2d_20", "laplace2d_33_shift", "random_spd_150"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_condest_spd(name, precision):
    print(f"case condest {name} {precision}", flush=True)
    A = {"laplace2d_20": lambda: laplace2d(20),
         "laplace2d_33_shift": lambda: laplace2d(33, shift=0.01),
         "random_spd_150": lambda: _spd_random(150, 5)}[name]().tocsc()
    n = A.shape[0]
    s = SparseSpdSolver(A)
    findorder(s)
    symbolicfactor(s)
    inmatrix(s)
    factor(s, precision)
    cond, ainv = condest(s)
    Ad = A.toarray()
    exact_ainv = np.abs(np.linalg.inv(Ad)).sum(axis=0).max()
    exact = np.abs(Ad).sum(axis=0).max() * exact_ainv
    # the same algorithm on the oracle's LU solves of A (its SPD oracle)
    so = SparseSolver(A)
    findorder(so)
    symbolicfactor(so)
    sym = so.slvr.sym
    _, cp, rv, nz = csc1(A)
    lnz, unz, ipiv, rc, _ = orc.factor_csc(sym, cp, rv, nz)
    assert rc == 0
    est_o = orc.ainverse_onenorm(lambda b: orc.solve(sym, lnz, unz, ipiv, b.copy()), n)
    tol = 1e-10 if precision == "fp64" else 1e-4
    print(f"  gpu {ainv:.12e} oracle {est_o:.12e} exact {exact_ainv:.12e} cond {cond:.4e}")
    assert abs(ainv - est_o) <= tol * est_o
    assert ainv <= exact_ainv * (1 + tol)
    assert ainv >= exact_ainv / 3.0
    assert abs(cond - np.abs(Ad).sum(axis=0).max() * ainv) <= 1e-12 * cond


def test_condest_needs_spd_factor():
    A = laplace2d(8).tocsc()
    s = SparseSolver(A)
    findorder(s)
    symbolicfactor(s)
    inmatrix(s)
    factor(s, "fp64")
    with pytest.raises(ValueError):
        condest(s)
