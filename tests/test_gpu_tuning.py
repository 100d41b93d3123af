"""Plan knobs (small-versus-MFMA front split, block width, update depth,
small-front wave threshold) and the autotuner: every setting gives the
factor of the default plan to rounding, with the same interchanges, and the
oracle's on the same input; the autotuner returns a setting from its search
and leaves the solver factored with it."""
import numpy as np
import pytest

import oracle.oracle as orc
from helpers import backward_error, csc1
from sparspak_hip import (SparseSolver, SparseSpdSolver, autotune, factor, findorder,
                          inmatrix, symbolicfactor, triangularsolve)
from sparspak_hip.gallery import convdiff3d, laplace2d, uniform

pytestmark = pytest.mark.gpu

SETTINGS = [
    dict(small_lmax=32),
    dict(small_lmax=0),
    dict(block_cols=64),
    dict(block_cols=96, upd_group=3),
    dict(upd_group=1),
    dict(upd_group=16),
    dict(small_wave_min=1),
    dict(small_lmax=48, block_cols=32, upd_group=2, small_wave_min=64),
]


def _prepared(A, cls=SparseSolver):
    s = cls(A)
    findorder(s)
    symbolicfactor(s)
    inmatrix(s)
    return s


@pytest.fixture(scope="module")
def cd16():
    A = convdiff3d(16)
    s = _prepared(A)
    factor(s, "fp64")
    return A, s.slvr.download()


@pytest.mark.parametrize("knobs", SETTINGS, ids=lambda k: ",".join(f"{a}={b}" for a, b in k.items()))
def test_plan_knobs_keep_the_factor(cd16, knobs):
    print(f"case tuning convdiff3d 16^3 fp64 {knobs}", flush=True)
    A, (l0, u0, p0) = cd16
    s = _prepared(A)
    s.slvr.set_tuning(**knobs)
    got = s.slvr.get_tuning()
    assert all(got[k] == v for k, v in knobs.items())
    factor(s, "fp64")
    lnz, unz, ipiv = s.slvr.download()
    assert np.array_equal(ipiv, p0)
    assert np.max(np.abs(lnz - l0)) <= 1e-12 * np.max(np.abs(l0))
    assert np.max(np.abs(unz - u0)) <= 1e-12 * np.max(np.abs(u0))
    b = A @ uniform(7, A.shape[0], -1, 1)
    x = b.copy()
    triangularsolve(s, x, refine=False)
    assert backward_error(A, x, b) <= 1e-14


def test_small_split_against_oracle():
    # all fronts through the blocked (matrix-core) path, checked against the
    # oracle's factor of the same input
    print("case tuning convdiff3d 10^3 small_lmax=0 vs oracle", flush=True)
    A = convdiff3d(10)
    s = _prepared(A)
    s.slvr.set_tuning(small_lmax=0, block_cols=64)
    factor(s, "fp64")
    lnz, unz, ipiv = s.slvr.download()
    n_, cp, rv, nz = csc1(A)
    ol, ou, op, rc, _ = orc.factor_csc(s.slvr.sym, cp, rv, nz)
    assert np.array_equal(ipiv, op)
    assert np.max(np.abs(lnz - ol)) <= 1e-12 * np.max(np.abs(ol))
    assert np.max(np.abs(unz - ou)) <= 1e-12 * np.max(np.abs(ou))


def test_spd_knobs():
    print("case tuning laplace2d 60 SPD", flush=True)
    A = laplace2d(60)
    b = A @ uniform(9, A.shape[0], -1, 1)
    s0 = _prepared(A, SparseSpdSolver)
    factor(s0, "fp64")
    l0 = s0.slvr.download()[0]
    s = _prepared(A, SparseSpdSolver)
    s.slvr.set_tuning(small_lmax=16, upd_group=2, small_wave_min=1)
    factor(s, "fp64")
    l1 = s.slvr.download()[0]
    assert np.max(np.abs(l1 - l0)) <= 1e-12 * np.max(np.abs(l0))
    x = b.copy()
    triangularsolve(s, x, refine=False)
    assert backward_error(A, x, b) <= 1e-14


def test_knobs_out_of_range_are_rejected():
    print("case tuning range checks", flush=True)
    s = _prepared(laplace2d(10))
    for bad in (dict(small_lmax=65), dict(small_lmax=-1), dict(block_cols=16),
                dict(block_cols=256), dict(upd_group=0), dict(small_wave_min=0)):
        with pytest.raises(RuntimeError):
            s.slvr.set_tuning(**bad)
    with pytest.raises(ValueError):
        s.slvr.set_tuning(tile=3)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autotune_picks_a_searched_setting(precision):
    print(f"case autotune convdiff3d 20^3 {precision}", flush=True)
    A = convdiff3d(20)
    s = _prepared(A)
    best = autotune(s, precision, reps=1)
    assert best["small_lmax"] in (64, 48, 32)
    assert best["block_cols"] in (128, 96, 64, 48, 32)
    assert best["upd_group"] in (8, 4, 16, 2)
    assert best["small_wave_min"] in (16384, 4096, 65536)
    assert best["factor_ms"] > 0
    assert s.slvr.get_tuning() == best
    b = A @ uniform(11, A.shape[0], -1, 1)
    x = b.copy()
    triangularsolve(s, x, refine=True, tol=1e-14)   # factored by autotune
    assert backward_error(A, x, b) <= 1e-14
