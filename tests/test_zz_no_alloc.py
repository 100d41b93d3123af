"""spkhip_solve_device_noalloc: the solve for callers that cannot allocate
device memory, through the C ABI, against spkhip_solve_device on the same
factor and right-hand sides (bit-identical: same kernels, same order).

* Two calls in a row on one GPU, the second with a smaller input (another
  handle, fewer unknowns and right-hand sides, float factor): the shared
  scratch keeps the first call's contents, which must not leak into the
  second.
* An SPD factor (the other front-width instantiation of the sweeps).
* A first solve after a factorisation allocates no device memory (the free
  memory of the device is unchanged while spkhip_solve_device would need
  megabytes for its buffers).
* More right-hand sides than the scratch holds: -2, the right-hand sides
  untouched.
"""
import ctypes
import sys

import numpy as np
import pytest

from helpers import backward_error
from sparspak_hip import SparseSolver, SparseSpdSolver, _lib, findorder, symbolicfactor
from sparspak_hip.gallery import convdiff3d, laplace2d, uniform

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _factored(A, prec, kind="lu"):
    s = SparseSolver(A) if kind == "lu" else SparseSpdSolver(A)
    findorder(s)
    symbolicfactor(s)
    s.slvr._upload()
    L, h = s.slvr._dev.L, s.slvr._dev.h
    nz = np.ascontiguousarray(A.data, dtype=np.float64)
    _lib.check(L.spkhip_inmatrix(h, _lib.pd(nz)), "inmatrix")
    _lib.check(L.spkhip_factor(h, prec), "factor")
    return s, L, h


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _rhs(n, nrhs, seed):
    return np.stack([uniform(seed + q, n, -1.0, 1.0) for q in range(nrhs)], axis=1)


def _solve_both(A, L, h, nrhs, seed, check_alloc=False):
    """noalloc solve first (before the handle's own solve buffers exist),
    then spkhip_solve_device on the same rhs; returns both solutions."""
    import torch
    n = A.shape[0]
    B = _rhs(n, nrhs, seed)
    Bf = np.asfortranarray(B)
    d_na = torch.from_numpy(Bf.T.copy()).cuda()   # column-major n x nrhs
    d_ref = d_na.clone()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    rc = L.spkhip_solve_device_noalloc(h, nrhs, _vp(d_na), n)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert rc == 0, f"noalloc rc {rc}"
    _lib.check(L.spkhip_solve_device(h, nrhs, _vp(d_ref), n), "solve_device")
    torch.cuda.synchronize()
    free2 = torch.cuda.mem_get_info()[0]
    x_na = d_na.cpu().numpy().T
    x_ref = d_ref.cpu().numpy().T
    _say(f"  n {n} nrhs {nrhs}: free memory drop noalloc {free0 - free1} B, "
         f"then solve_device {free1 - free2} B")
    if check_alloc:
        # solve_device allocates its rhs and vector buffers; the noalloc
        # call must not have
        assert free1 - free2 >= 4 << 20, "solve_device was expected to allocate"
        assert free0 - free1 < 1 << 20, "the noalloc solve allocated device memory"
    return B, x_na, x_ref


def test_noalloc_two_calls_second_smaller():
    _say("case noalloc call 1: convdiff3d 24^3 LU fp64, 8 rhs")
    A = convdiff3d(24)
    s1, L, h1 = _factored(A, 0)
    B, x_na, x_ref = _solve_both(A, L, h1, 8, 11, check_alloc=True)
    assert np.array_equal(x_na, x_ref), "noalloc differs from solve_device (call 1)"
    for q in range(B.shape[1]):
        assert backward_error(A, x_na[:, q], B[:, q]) <= 1e-14

    _say("case noalloc call 2: convdiff3d 12^3 LU fp32, 1 rhs")
    A2 = convdiff3d(12)
    s2, L, h2 = _factored(A2, 1)
    B2, x2_na, x2_ref = _solve_both(A2, L, h2, 1, 21)
    assert np.array_equal(x2_na, x2_ref), "noalloc differs from solve_device (call 2)"
    assert backward_error(A2, x2_na[:, 0], B2[:, 0]) <= 1e-5

    _say("case noalloc call 3: the first handle again, 2 rhs")
    B3, x3_na, x3_ref = _solve_both(A, L, h1, 2, 31)
    assert np.array_equal(x3_na, x3_ref), "noalloc differs from solve_device (call 3)"


def test_noalloc_spd():
    _say("case noalloc SPD: laplace2d 60 + shift, fp64, 3 rhs")
    A = laplace2d(60, shift=0.1).tocsc()
    s, L, h = _factored(A, 0, kind="spd")
    B, x_na, x_ref = _solve_both(A, L, h, 3, 41)
    assert np.array_equal(x_na, x_ref), "noalloc differs from solve_device (SPD)"
    for q in range(3):
        assert backward_error(A, x_na[:, q], B[:, q]) <= 1e-14


def test_noalloc_rejects_too_many_rhs():
    import torch
    _say("case noalloc too many rhs: convdiff3d 20^3 fp64")
    A = convdiff3d(20)
    n = A.shape[0]
    s, L, h = _factored(A, 0)
    nrhs = SCRATCH_BYTES // (n * 8) + 1   # the right-hand sides alone overflow
    d_b = torch.ones((nrhs, n), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = L.spkhip_solve_device_noalloc(h, nrhs, _vp(d_b), n)
    torch.cuda.synchronize()
    assert rc == -2
    assert bool(torch.all(d_b == 1.0)), "rejected call modified the right-hand sides"
