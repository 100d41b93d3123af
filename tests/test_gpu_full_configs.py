"""The remaining benchmark configurations at their full sizes, checked through
properties that do not depend on size: every system's (every right-hand
side's) normwise backward error after refinement is <= 1e-14, the float
factor of a well-conditioned system refines in few iterations, and one
system of each batch equals its own unbatched factorisation.

  Q1 elasticity 40^3 (3 dof per node, clamped dofs kept as identity rows,
  n = 206 763), FP32 factor with bf16 trailing updates in fronts wider than
  256 columns, FP64 refinement, body load N(0, 1);
  2 000 uneven 2D Laplacians (side 8..96, diagonal shift U[0.1, 1]), SPD,
  4 right-hand sides each, FP64 and FP32 + refinement;
  256 value sets on one 32^3 convection-diffusion pattern (log-uniform
  diffusivity in [1e-2, 1e2]), FP32 + refinement, 16 right-hand sides each.
"""
import time

import numpy as np
import pytest

from helpers import backward_error
from sparspak_hip import (SparseSolver, SparseSpdSolver, factor, findorder, inmatrix,
                          symbolicfactor, triangularsolve)
from sparspak_hip.batch import BatchedSolver
from sparspak_hip.gallery import (convdiff3d_varcoef, elasticity3d, laplace2d, normal,
                                  splitmix64, uniform)

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def elasticity40():
    A = elasticity3d(40)
    s = SparseSolver(A)
    findorder(s)
    symbolicfactor(s)
    inmatrix(s)
    return A, s


@pytest.mark.parametrize("precision", ["fp32_bf16", "fp32_bf16x3", "fp64"])
def test_config3_elasticity40_refines_to_1e14(elasticity40, precision):
    _say(f"case full config3 elasticity 40^3 {precision}")
    A, s = elasticity40
    n = A.shape[0]
    assert n == 206763
    f = normal(3, n)
    factor(s, precision)
    s.slvr._factor(precision)  # timed again after the arena's first touch
    x = f.copy()
    triangularsolve(s, x, refine=True, tol=1e-14, max_iter=30)
    st = s.slvr.stats()
    eta = backward_error(A, x, f)
    _say(f"  n {n} factor {st.t_factor_ms:.1f} ms ({st.factor_flops / st.t_factor_ms / 1e9:.1f}"
         f" TF/s) refine {st.t_refine_ms:.1f} ms iters {st.refine_iters} eta {eta:.2e}")
    assert eta <= 1e-14
    assert st.refine_iters <= 10


def _batched_laplacians():
    nsys = 2000
    sides = 8 + (splitmix64(4, nsys) % np.uint64(89)).astype(int)  # 8..96
    shifts = uniform(40, nsys, 0.1, 1.0)
    mats = [laplace2d(int(sd), shift=float(sh)) for sd, sh in zip(sides, shifts)]
    rhs = [np.stack([uniform(400 + 4 * i + k, A.shape[0], -1, 1) for k in range(4)], axis=1)
           for i, A in enumerate(mats)]
    return sides, mats, rhs


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_config4_batched2000_full(precision):
    _say(f"case full config4 batched 2000 {precision}")
    sides, mats, rhs = _batched_laplacians()
    assert sides.min() >= 8 and sides.max() <= 96
    bs = BatchedSolver(mats, kind="spd")
    bs.set_values()
    bs.factor(precision)
    xs = bs.solve(rhs, refine=True, tol=1e-14)
    st = bs.stats()
    worst = max(float(np.max(backward_error(A, x, b))) for A, b, x in zip(mats, rhs, xs))
    _say(f"  n {bs.n} factor {st.t_factor_ms:.1f} ms refine {st.t_refine_ms:.1f} ms"
         f" iters {st.refine_iters} worst eta {worst:.2e}")
    assert worst <= 1e-14
    # the largest system against its own factorisation (same kernels)
    i = int(np.argmax(sides))
    s = SparseSpdSolver(mats[i])
    findorder(s)
    symbolicfactor(s)
    inmatrix(s)
    factor(s, precision)
    xi = np.asfortranarray(rhs[i].copy())
    triangularsolve(s, xi, refine=True, tol=1e-14)
    assert np.max(np.abs(xi - xs[i])) <= 1e-12 * np.max(np.abs(xi))


def test_config5_valuesets256_full():
    _say("case full config5 256 value sets 32^3 fp32")
    g, nsets = 32, 256
    t0 = time.time()
    mats = [convdiff3d_varcoef(g, 10.0 ** uniform(5000 + k, g ** 3, -2.0, 2.0))
            for k in range(nsets)]
    bs = BatchedSolver(mats)
    assert all(s[4] is bs.systems[0][4] for s in bs.systems)  # one symbolic factor
    bs.set_values()
    bs.factor("fp32")
    rhs = np.stack([uniform(900 + k, bs.n, -1, 1) for k in range(16)], axis=1)
    x = bs.solve(rhs, refine=True, tol=1e-14)
    st = bs.stats()
    worst = 0.0
    for i, (A, off) in enumerate(zip(mats, bs.offsets)):
        n = A.shape[0]
        worst = max(worst, float(np.max(backward_error(A, x[off:off + n], rhs[off:off + n]))))
    _say(f"  n {bs.n} factor {st.t_factor_ms:.1f} ms ({bs.flops() / st.t_factor_ms / 1e9:.1f}"
         f" TF/s) refine {st.t_refine_ms:.1f} ms iters {st.refine_iters} worst eta {worst:.2e}"
         f" (host {time.time() - t0:.0f} s)")
    assert worst <= 1e-14
