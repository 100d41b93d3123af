"""The factor's plan variants against the oracle: the diagonal regions in
registers or in LDS (SPKHIP_DIAG_SLAB), each in-group update merged with the
next block's diagonal launch or launched after it (SPKHIP_MERGE_BORDER), for
LU with pivoting confined to the supernodes (SpkLUFactor.jl:230-240) and for
SPD (no interchanges, symmetric updates; merged, the update's first tile
column starts below the next block's region and mirrors the entries left of
it). Every variant gives the oracle's pivots and its entries within 1e-12
(double). The switches are read when the plan is built, so each variant
uploads its own symbolic factor."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle.oracle as orc
from helpers import backward_error, csc1, random_pattern, spd_meaningful
from sparspak_hip import (SparseSolver, SparseSpdSolver, factor, findorder, inmatrix,
                          symbolicfactor, triangularsolve)
from sparspak_hip.gallery import convdiff3d, elasticity3d, uniform

pytestmark = pytest.mark.gpu

VARIANTS = {"slab+merge": ("1", "1"), "slab": ("1", "0"), "lds": ("0", "0")}


def _factor_with(A, cls, slab, merge, maxblocksize=None, precision="fp64", tuning=None):
    old = {k: os.environ.get(k) for k in ("SPKHIP_DIAG_SLAB", "SPKHIP_MERGE_BORDER")}
    os.environ["SPKHIP_DIAG_SLAB"] = slab
    os.environ["SPKHIP_MERGE_BORDER"] = merge
    try:
        s = cls(A)
        if maxblocksize is not None:  # smaller pieces: more of them per block
            s.slvr.sym.maxblocksize = maxblocksize
        findorder(s)
        symbolicfactor(s)
        inmatrix(s)
        if tuning:
            s.slvr.set_tuning(**tuning)
        factor(s, precision)
        return s
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _weak_lu(n, seed):
    # an unsymmetric pattern with a weak diagonal: the pivot search matters
    return sp.csc_matrix(random_pattern(n, 0.05, seed, False, False) + sp.diags(np.full(n, 0.3)))


LU_CASES = {"convdiff3d_20": lambda: convdiff3d(20), "weakdiag_400": lambda: _weak_lu(400, 11)}


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("name", list(LU_CASES))
def test_lu_plan_variants_match_oracle(name, variant):
    print(f"case plan variant LU {name} {variant}", flush=True)
    A = sp.csc_matrix(LU_CASES[name]())
    s = _factor_with(A, SparseSolver, *VARIANTS[variant])
    lnz, unz, ipiv = s.slvr.download()
    n, cp, rv, nz = csc1(A)
    ol, ou, op, rc, _ = orc.factor_csc(s.slvr.sym, cp, rv, nz)
    assert rc == 0
    assert np.array_equal(ipiv, op), "pivots differ from the oracle"
    scale = max(np.max(np.abs(ol)), np.max(np.abs(ou)))
    assert np.max(np.abs(lnz - ol)) <= 1e-12 * scale
    assert np.max(np.abs(unz - ou)) <= 1e-12 * scale
    # the weak diagonal grows the factor: the oracle's own unrefined solve
    # reaches about 7e-12 there, so the bar is relative to it
    b = uniform(5, n, -1.0, 1.0)
    eta_o = backward_error(A, orc.solve(s.slvr.sym, ol, ou, op, b.copy()), b)
    x = b.copy()
    triangularsolve(s, x)
    assert backward_error(A, x, b) <= max(4.0 * eta_o, 1e-15)


def _dense_spd(n, seed):
    # one front of three pieces, 42, 43 and 65 columns at n = 150: the third
    # block is wider than 64 columns, so the step before it merges nothing
    rng = np.random.default_rng(seed)
    M = rng.uniform(-1, 1, (n, n))
    return sp.csc_matrix((M + M.T) / 2 + n * np.eye(n))


SPD_CASES = {"elasticity3d_8": lambda: elasticity3d(8), "dense_spd_150": lambda: _dense_spd(150, 3)}


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("name", list(SPD_CASES))
def test_spd_plan_variants_match_oracle(name, variant):
    print(f"case plan variant SPD {name} {variant}", flush=True)
    A = sp.csc_matrix(SPD_CASES[name]())
    s = _factor_with(A, SparseSpdSolver, *VARIANTS[variant])
    lnz = s.slvr.download()[0]
    n, cp, rv, nz = csc1(A)
    ol, _, _ = orc.inmatrix(s.slvr.sym, cp, rv, nz)
    orc.ldltfactor(s.slvr.sym, ol)
    m = spd_meaningful(s.slvr.sym)
    assert np.max(np.abs(lnz[m] - ol[m])) <= 1e-12 * np.max(np.abs(ol[m]))
    assert np.all(lnz[~m] == 0.0)
    b = uniform(6, n, -1.0, 1.0)
    x = b.copy()
    triangularsolve(s, x)
    assert backward_error(A, x, b) < 1e-14


# pieces of at most 6 or 8 columns: factor blocks of up to eight pieces (the
# descriptors' piece tables full) and solve blocks of up to eight pieces
# (blocks of 96 or 128 columns take the LDS diagonal kernel and the general
# off-diagonal path, whatever the variant)
MANY = {"convdiff3d_16_lu_mb8": (lambda: convdiff3d(16), SparseSolver, 8, None),
        "convdiff3d_16_lu_mb6": (lambda: convdiff3d(16), SparseSolver, 6, None),
        "convdiff3d_14_lu_mb4": (lambda: convdiff3d(14), SparseSolver, 4, None),
        "convdiff3d_16_lu_mb12": (lambda: convdiff3d(16), SparseSolver, 12, None),
        "elasticity3d_6_spd_mb6": (lambda: elasticity3d(6), SparseSpdSolver, 6, None),
        "convdiff3d_16_lu_mb8_block128": (lambda: convdiff3d(16), SparseSolver, 8,
                                          dict(block_cols=128)),
        "elasticity3d_6_spd_mb6_block96": (lambda: elasticity3d(6), SparseSpdSolver, 6,
                                           dict(block_cols=96))}


@pytest.mark.parametrize("variant", ["slab+merge", "lds"])
@pytest.mark.parametrize("name", list(MANY))
def test_many_pieces_per_block_match_oracle(name, variant):
    print(f"case plan variant many pieces {name} {variant}", flush=True)
    gen, cls, mb, tuning = MANY[name]
    A = sp.csc_matrix(gen())
    s = _factor_with(A, cls, *VARIANTS[variant], maxblocksize=mb, tuning=tuning)
    sym = s.slvr.sym
    assert max(np.diff(np.asarray(sym.xsuper))) <= 1.5 * mb
    lnz, unz, ipiv = s.slvr.download()
    n, cp, rv, nz = csc1(A)
    if cls is SparseSolver:
        ol, ou, op, rc, _ = orc.factor_csc(sym, cp, rv, nz)
        assert rc == 0
        assert np.array_equal(ipiv, op), "pivots differ from the oracle"
        scale = max(np.max(np.abs(ol)), np.max(np.abs(ou)))
        assert np.max(np.abs(lnz - ol)) <= 1e-12 * scale
        assert np.max(np.abs(unz - ou)) <= 1e-12 * scale
    else:
        ol, _, _ = orc.inmatrix(sym, cp, rv, nz)
        orc.ldltfactor(sym, ol)
        m = spd_meaningful(sym)
        assert np.max(np.abs(lnz[m] - ol[m])) <= 1e-12 * np.max(np.abs(ol[m]))
    # float factor: explicit-inverse solves (blocks of up to eight pieces)
    # against substitution, and the refined solution
    s32 = _factor_with(A, cls, *VARIANTS[variant], maxblocksize=mb, precision="fp32",
                       tuning=tuning)
    b = uniform(9, n, -1.0, 1.0)
    xs = []
    for inv in ("1", "0"):
        old = os.environ.get("SPKHIP_SOLVE_INV")
        os.environ["SPKHIP_SOLVE_INV"] = inv
        try:
            x = b.copy()
            triangularsolve(s32, x, refine=False)
            xs.append(x)
        finally:
            if old is None:
                os.environ.pop("SPKHIP_SOLVE_INV", None)
            else:
                os.environ["SPKHIP_SOLVE_INV"] = old
    assert np.max(np.abs(xs[0] - xs[1])) <= 1e-4 * np.max(np.abs(xs[1]))
    x = b.copy()
    triangularsolve(s32, x)
    assert backward_error(A, x, b) <= 1e-14
