"""The benchmark configurations as GPU parity cases (reduced sizes where the
full one is a benchmark, not a test): 2D Laplacian (SPD and LU), 3D
convection-diffusion (FP64 and FP32 + refinement), Q1 elasticity, batched
uneven SPD systems, many value sets on one pattern."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle.oracle as orc
from helpers import backward_error, csc1
from sparspak_hip import (SparseSolver, SparseSpdSolver, factor, findorder, inmatrix,
                          symbolicfactor, triangularsolve)
from sparspak_hip.batch import BatchedSolver
from sparspak_hip.gallery import (convdiff3d, convdiff3d_varcoef, elasticity3d, laplace2d,
                                  normal, splitmix64, uniform)

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _solver(A, cls=SparseSolver, precision="fp64"):
    s = cls(A)
    findorder(s)
    symbolicfactor(s)
    inmatrix(s)
    factor(s, precision)
    return s


def test_config1_laplace2d_100_spd_and_lu():
    # 5-point Laplacian 100 x 100, b = A x_true, x_true ~ U[-1,1] (seed 1)
    _say("case config1 laplace2d 100")
    A = laplace2d(100)
    n = A.shape[0]
    xt = uniform(1, n, -1, 1)
    b = A @ xt
    for cls in (SparseSpdSolver, SparseSolver):
        s = _solver(A, cls)
        x = b.copy()
        triangularsolve(s, x)
        eta = backward_error(A, x, b)
        _say(f"  {cls.__name__}: eta {eta:.2e}")
        assert eta < 1e-15
        assert np.max(np.abs(x - xt)) < 1e-11
    # SPD factor against the corrected LDL^T oracle (the reference's own
    # LDL^T is defective for supernodes wider than one column)
    s = _solver(A, SparseSpdSolver)
    lnz = s.slvr.download()[0]
    n_, cp, rv, nz = csc1(A)
    ol, _, _ = orc.inmatrix(s.slvr.sym, cp, rv, nz)
    orc.ldltfactor(s.slvr.sym, ol)
    d = np.array([lnz[s.slvr.sym.xlnz[j] - 1 + (j - (s.slvr.sym.xsuper[s.slvr.sym.snode[j] - 1] - 1))]
                  for j in range(n)])
    do = np.array([ol[s.slvr.sym.xlnz[j] - 1 + (j - (s.slvr.sym.xsuper[s.slvr.sym.snode[j] - 1] - 1))]
                   for j in range(n)])
    assert np.max(np.abs(d - do)) < 1e-12 * np.max(np.abs(do))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_config2_convdiff3d_reduced(precision):
    _say(f"case config2 convdiff3d 24^3 {precision}")
    A = convdiff3d(24)
    n = A.shape[0]
    xt = uniform(2, n, -1, 1)
    b = A @ xt
    s = _solver(A, precision=precision)
    x = b.copy()
    triangularsolve(s, x, refine=True, tol=1e-14)
    eta = backward_error(A, x, b)
    _say(f"  eta {eta:.2e} iters {s.slvr.stats().refine_iters}")
    assert eta <= 1e-14
    assert np.max(np.abs(x - xt)) < 1e-10


def test_convdiff3d_32_accuracy_matches_oracle():
    # the device factor is at least as accurate as the reference's own
    # algorithm (the oracle) on this operator, not merely under the 1e-14 bar
    _say("case convdiff3d 32^3 fp64 accuracy")
    A = convdiff3d(32)
    n = A.shape[0]
    xt = uniform(2, n, -1, 1)
    b = A @ xt
    s = _solver(A)
    x = b.copy()
    triangularsolve(s, x, refine=False)
    eta = backward_error(A, x, b)
    n_, cp, rv, nz = csc1(A)
    ol, ou, op, rc, _ = orc.factor_csc(s.slvr.sym, cp, rv, nz)
    xo = orc.solve(s.slvr.sym, ol, ou, op, b.copy())
    eta_o = backward_error(A, xo, b)
    _say(f"  eta {eta:.2e} oracle {eta_o:.2e}")
    assert eta <= max(eta_o, 1e-16)


def test_config3_elasticity_reduced():
    _say("case config3 elasticity 5^3")
    A = elasticity3d(5)
    n = A.shape[0]
    f = normal(3, n)
    for cls, prec in ((SparseSpdSolver, "fp64"), (SparseSolver, "fp64"), (SparseSolver, "fp32")):
        s = _solver(A, cls, prec)
        x = f.copy()
        triangularsolve(s, x, refine=True, tol=1e-14)
        eta = backward_error(A, x, f)
        _say(f"  {cls.__name__} {prec}: eta {eta:.2e}")
        assert eta <= 1e-14
    # LU factor against the oracle
    s = _solver(A)
    lnz, unz, ipiv = s.slvr.download()
    n_, cp, rv, nz = csc1(A)
    ol, ou, op, rc, _ = orc.factor_csc(s.slvr.sym, cp, rv, nz)
    assert np.array_equal(ipiv, op)
    assert np.max(np.abs(lnz - ol)) < 1e-12 * np.max(np.abs(ol))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_config4_batched_uneven(precision):
    # uneven 2D Laplacians + a per-system diagonal shift U[0.1, 1], 4 rhs each
    _say(f"case config4 batched {precision}")
    nsys = 40
    sides = 8 + (splitmix64(4, nsys) % np.uint64(17)).astype(int)  # 8..24
    shifts = uniform(40, nsys, 0.1, 1.0)
    mats = [laplace2d(int(sd), shift=float(sh)) for sd, sh in zip(sides, shifts)]
    bs = BatchedSolver(mats, kind="spd")
    bs.set_values()
    bs.factor(precision)
    rhs = [np.stack([uniform(400 + 4 * i + k, A.shape[0], -1, 1) for k in range(4)], axis=1)
           for i, A in enumerate(mats)]
    xs = bs.solve(rhs, refine=True, tol=1e-14)
    worst = 0.0
    for A, b, x in zip(mats, rhs, xs):
        worst = max(worst, float(np.max(backward_error(A, x, b))))
    _say(f"  worst eta {worst:.2e}")
    assert worst <= 1e-14
    # one system against its own factorisation (same kernels, no batching)
    s = _solver(mats[3], SparseSpdSolver, precision)
    x3 = rhs[3].copy(order="F")
    triangularsolve(s, x3, refine=True, tol=1e-14)
    assert np.max(np.abs(x3 - xs[3])) < 1e-12 * np.max(np.abs(x3))


def test_config5_value_sets_one_pattern():
    # 16 value sets on a 10^3 pattern, log-uniform diffusivity, 16 rhs each
    _say("case config5 value sets")
    g, nsets = 10, 16
    mats = []
    for k in range(nsets):
        kappa = 10.0 ** uniform(5000 + k, g ** 3, -2.0, 2.0)
        mats.append(convdiff3d_varcoef(g, kappa))
    bs = BatchedSolver(mats)
    assert all(s[4] is bs.systems[0][4] for s in bs.systems)  # one symbolic factor
    bs.set_values()
    bs.factor("fp32")
    rhs = [np.stack([uniform(900 + 16 * i + k, g ** 3, -1, 1) for k in range(16)], axis=1)
           for i in range(nsets)]
    xs = bs.solve(rhs, refine=True, tol=1e-14)
    worst = max(float(np.max(backward_error(A, x, b))) for A, b, x in zip(mats, rhs, xs))
    _say(f"  worst eta {worst:.2e} iters {bs.stats().refine_iters}")
    assert worst <= 1e-14


def test_fp16_trailing_updates_refine_to_1e14():
    # as the bf16 case with diagonally scaled fp16 operands; at most as many
    # refinement steps as with bf16 (three more bits per operand)
    _say("case convdiff3d 24^3 fp32_fp16")
    A = convdiff3d(24)
    n = A.shape[0]
    xt = uniform(2, n, -1, 1)
    b = A @ xt
    its = {}
    for prec in ("fp32_bf16", "fp32_fp16"):
        s = _solver(A, precision=prec)
        x = b.copy()
        triangularsolve(s, x, refine=True, tol=1e-14, max_iter=30)
        eta = backward_error(A, x, b)
        its[prec] = s.slvr.stats().refine_iters
        _say(f"  {prec}: eta {eta:.2e} iters {its[prec]}")
        assert eta <= 1e-14
    assert its["fp32_fp16"] <= its["fp32_bf16"]


def test_split_bf16_trailing_updates_refine_to_1e14():
    # split-bf16 operands (hi + lo, three products): the refinement needs at
    # most as many steps as plain float updates on this operator need plus
    # one, and no more than plain bf16
    _say("case convdiff3d 24^3 fp32_bf16x3")
    A = convdiff3d(24)
    n = A.shape[0]
    xt = uniform(2, n, -1, 1)
    b = A @ xt
    its = {}
    for prec in ("fp32", "fp32_bf16", "fp32_bf16x3"):
        s = _solver(A, precision=prec)
        x = b.copy()
        triangularsolve(s, x, refine=True, tol=1e-14, max_iter=30)
        eta = backward_error(A, x, b)
        its[prec] = s.slvr.stats().refine_iters
        _say(f"  {prec}: eta {eta:.2e} iters {its[prec]}")
        assert eta <= 1e-14
    assert its["fp32_bf16x3"] <= min(its["fp32_bf16"], its["fp32"] + 1)


def test_bf16_trailing_updates_refine_to_1e14():
    # float factor with bf16 operands in the trailing updates of fronts wider
    # than 256 columns (the 24^3 top separator is), FP64 refinement
    _say("case convdiff3d 24^3 fp32_bf16")
    A = convdiff3d(24)
    n = A.shape[0]
    xt = uniform(2, n, -1, 1)
    b = A @ xt
    s = _solver(A, precision="fp32_bf16")
    x = b.copy()
    triangularsolve(s, x, refine=True, tol=1e-14, max_iter=30)
    eta = backward_error(A, x, b)
    _say(f"  eta {eta:.2e} iters {s.slvr.stats().refine_iters}")
    assert eta <= 1e-14
