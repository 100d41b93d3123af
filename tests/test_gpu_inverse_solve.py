"""Float solves with the explicit inverses of the diagonal blocks (the
default for float factors and up to three right-hand sides) against the same
solves by substitution (SPKHIP_SOLVE_INV=0, the reference's operation order,
SpkLUFactor.jl:269-377): fronts of one block and of several, with and
without rows below the blocks and update rows, LU and SPD, one and three
right-hand sides. The two differ only by float rounding (the inverses are
formed in float); the refined solutions meet the double backward-error bar.
Six right-hand sides take the substitution kernels in both settings."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from helpers import backward_error
from sparspak_hip import (SparseSolver, SparseSpdSolver, factor, findorder, inmatrix,
                          symbolicfactor, triangularsolve)
from sparspak_hip.gallery import convdiff3d, elasticity3d, uniform

pytestmark = pytest.mark.gpu


def _spd_random(n, seed):
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=0.08, random_state=rng, format="csc")
    M = (M + M.T) * 0.5
    d = np.asarray(abs(M).sum(axis=1)).ravel() + rng.uniform(0.01, 0.5, n)
    return sp.csc_matrix(M + sp.diags(d))


def _dense(n):
    rng = np.random.default_rng(3)
    return sp.csc_matrix(rng.uniform(-1, 1, (n, n)) + n * np.eye(n))


CASES = {
    "dense_200": (lambda: _dense(200), "lu"),         # one front, blocks of several pieces
    "spd_150": (lambda: _spd_random(150, 5), "spd"),  # one front, three one-piece blocks
    "convdiff3d_16": (lambda: convdiff3d(16), "lu"),
    "elasticity_6": (lambda: elasticity3d(6), "spd"),
}


def _solve(s, B, inv, refine):
    old = os.environ.get("SPKHIP_SOLVE_INV")
    os.environ["SPKHIP_SOLVE_INV"] = "1" if inv else "0"
    try:
        X = B.copy(order="F")
        triangularsolve(s, X, refine=refine)
        return X
    finally:
        if old is None:
            os.environ.pop("SPKHIP_SOLVE_INV", None)
        else:
            os.environ["SPKHIP_SOLVE_INV"] = old


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("nrhs", [1, 3, 6])
def test_inverse_solves_match_substitution(name, nrhs):
    gen, kind = CASES[name]
    print(f"case inverse solve {name} nrhs={nrhs}", flush=True)
    A = sp.csc_matrix(gen())
    cls = SparseSolver if kind == "lu" else SparseSpdSolver
    s = cls(A)
    findorder(s)
    symbolicfactor(s)
    inmatrix(s)
    factor(s, "fp32")
    n = A.shape[0]
    B = np.asfortranarray(np.stack([uniform(70 + q, n, -1.0, 1.0) for q in range(nrhs)], axis=1))
    Xs = _solve(s, B, inv=False, refine=False)
    Xi = _solve(s, B, inv=True, refine=False)
    if nrhs > 3:
        assert np.array_equal(Xi, Xs), "more than three right-hand sides: substitution"
    for q in range(nrhs):
        rel = np.max(np.abs(Xi[:, q] - Xs[:, q])) / np.max(np.abs(Xs[:, q]))
        assert rel <= 1e-4, f"column {q}: inverse and substitution solves differ by {rel:.2e}"
    # each column of a three-column inverse solve is that column's own solve,
    # bit for bit
    if nrhs == 3:
        x0 = _solve(s, np.array(B[:, :1], order="F"), inv=True, refine=False)
        assert np.array_equal(x0[:, 0], Xi[:, 0])
    X = _solve(s, B, inv=True, refine=True)
    assert max(backward_error(A, X[:, q], B[:, q]) for q in range(nrhs)) <= 1e-14
