"""Value sets: several matrices of one pattern factored and solved in the
same launches (spkhip_set_value_sets, the batched form of the reference's
refactorisation with new values, sparspaklu!, SparseCSCInterface.jl:247-268).
Each set's factor (lnz, unz, pivots) and unrefined solutions equal, bit for
bit, those of a handle that holds that set alone: no arithmetic crosses
sets. LU and SPD, double and float factors (float solves of one and three
right-hand sides take the explicit-inverse steps, eight the substitution
kernels), fronts that take the blocked large-front kernels, the captured
factor and solve graphs (second use of a key), the allocation-free solve,
and the refined solve's per-set backward error."""
import ctypes

import numpy as np
import pytest

from helpers import backward_error
from sparspak_hip import _lib
from sparspak_hip.batch import BatchedSolver
from sparspak_hip.gallery import convdiff3d_varcoef, elasticity3d, uniform

pytestmark = pytest.mark.gpu

PREC = {"fp64": 0, "fp32": 1, "fp32_bf16x3": 4}


def _say(msg):
    print(msg, flush=True)


def _lu_sets(g, nsets):
    return [convdiff3d_varcoef(g, 10.0 ** uniform(7000 + k, g ** 3, -1.5, 1.5))
            for k in range(nsets)]


def _spd_sets(m, nsets):
    A = elasticity3d(m).tocsc()
    A.sort_indices()
    out = []
    for k in range(nsets):
        B = A.copy()
        B.data = B.data * (1.0 + 0.1 * k)  # a scaled operator plus a shift per set
        B.setdiag(B.diagonal() + 0.05 * (k + 1))  # stored diagonal: same pattern
        out.append(B)
    return out


def _single(A, kind, prec):
    b = BatchedSolver([A], kind=kind, value_sets=False)
    b.set_values()
    b.factor(prec)
    return b


def _download(b, nsets):
    L, h = b._dev.L, b._dev.h
    st = b.stats()
    nl, nu, n = st.nnz_lnz, st.nnz_unz, st.n
    if b.precision == 0:
        lnz, unz = np.zeros(nl * nsets), np.zeros(max(nu, 1) * nsets)
        pl, pu = _lib.pd(lnz), _lib.pd(unz)
        fn = L.spkhip_download
    else:
        lnz = np.zeros(nl * nsets, dtype=np.float32)
        unz = np.zeros(max(nu, 1) * nsets, dtype=np.float32)
        pl = lnz.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        pu = unz.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        fn = L.spkhip_download_f32
    ipvt = np.zeros(n * nsets, dtype=np.int64)
    _lib.check(fn(h, pl, pu, _lib.p64(ipvt)), "download")
    return lnz, unz[:nu * nsets], ipvt


CASES = [
    ("lu", "convdiff3d_varcoef 16^3 x4", lambda: _lu_sets(16, 4)),
    ("spd", "elasticity3d 6 x3", lambda: _spd_sets(6, 3)),
]


@pytest.mark.parametrize("kind,name,gen", CASES, ids=[c[1] for c in CASES])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp32_bf16x3"])
def test_value_sets_equal_single_set_handles(kind, name, gen, prec):
    _say(f"case value sets {name} {prec}")
    mats = gen()
    nsets = len(mats)
    n = mats[0].shape[0]
    vs = BatchedSolver(mats, kind=kind)
    assert vs.value_sets
    vs.set_values()
    vs.factor(prec)
    vs.factor(prec)  # second use of the key: the captured factor graph
    lnz, unz, ipvt = _download(vs, nsets)
    st = vs.stats()
    nl, nu = st.nnz_lnz, st.nnz_unz
    assert st.n == n
    singles = [_single(A, kind, prec) for A in mats]
    for v, s1 in enumerate(singles):
        l1, u1, p1 = _download(s1, 1)
        assert np.array_equal(lnz[v * nl:(v + 1) * nl], l1), f"set {v}: lnz differs"
        assert np.array_equal(unz[v * nu:(v + 1) * nu], u1), f"set {v}: unz differs"
        assert np.array_equal(ipvt[v * n:(v + 1) * n], p1), f"set {v}: pivots differ"
    for nrhs in (1, 3, 8):
        rhs = [np.stack([uniform(800 + 10 * v + q, n, -1, 1) for q in range(nrhs)], axis=1)
               for v in range(nsets)]
        for rep in range(2):  # the second solve replays the captured solve graph
            xs = vs.solve(rhs, refine=False)
            for v, s1 in enumerate(singles):
                x1 = s1.solve([rhs[v]], refine=False)[0]
                assert np.array_equal(xs[v], x1), f"set {v} nrhs {nrhs} rep {rep}: solve differs"
    rhs = [np.stack([uniform(900 + 10 * v + q, n, -1, 1) for q in range(2)], axis=1)
           for v in range(nsets)]
    xs = vs.solve(rhs, refine=True, tol=1e-14)
    worst = max(float(np.max(backward_error(A, x, b))) for A, b, x in zip(mats, rhs, xs))
    _say(f"  refined: worst eta {worst:.2e} iters {vs.stats().refine_iters}")
    assert worst <= 1e-14


def test_value_sets_noalloc_and_device_solves_agree():
    import torch
    _say("case value sets noalloc convdiff3d_varcoef 12^3 x3 fp32")
    mats = _lu_sets(12, 3)
    n, nsets, nrhs = mats[0].shape[0], 3, 2
    vs = BatchedSolver(mats)
    vs.set_values()
    vs.factor("fp32")
    L, h = vs._dev.L, vs._dev.h
    B = np.asfortranarray(np.stack([uniform(50 + q, n * nsets, -1, 1) for q in range(nrhs)],
                                   axis=1))
    d1 = torch.tensor(B.T.copy(), device="cuda")   # rows of d1 = columns of B
    d2 = d1.clone()
    torch.cuda.synchronize()
    _lib.check(L.spkhip_solve_device(h, nrhs, ctypes.c_void_p(d1.data_ptr()), n * nsets),
               "solve_device")
    _lib.check(L.spkhip_solve_device_noalloc(h, nrhs, ctypes.c_void_p(d2.data_ptr()),
                                             n * nsets), "solve_device_noalloc")
    torch.cuda.synchronize()
    assert torch.equal(d1, d2)
    X = vs.solve(B, refine=False)
    assert np.array_equal(X, d1.cpu().numpy().T)
    # a leading dimension short of every set's rows is rejected
    assert L.spkhip_solve_device(h, nrhs, ctypes.c_void_p(d1.data_ptr()), n) == -4


def test_value_sets_api_errors():
    _say("case value sets api: condest, new pattern")
    mats = _spd_sets(4, 2)
    vs = BatchedSolver(mats, kind="spd")
    vs.set_values()
    vs.factor("fp64")
    L, h = vs._dev.L, vs._dev.h
    c1, c2 = ctypes.c_double(), ctypes.c_double()
    assert L.spkhip_condest(h, ctypes.byref(c1), ctypes.byref(c2)) == -4
    assert L.spkhip_set_value_sets(h, 0) == -4
    assert L.spkhip_set_value_sets(h, 3) == 0
    assert L.spkhip_factor(h, 0) == -3  # values of two sets dropped with the change
