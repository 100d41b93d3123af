"""The device-resident path the benchmark times, through the C ABI:
spkhip_inmatrix_device -> spkhip_factor -> spkhip_solve_refine_device
(sparspaklu! refactorisation with values and right-hand sides already in
HBM, SparseCSCInterface.jl:247-268).

* 16^3: values from the device give a factor bit-identical to values from
  the host, which matches the oracle (1e-12, identical pivots); the device
  refined solve is bit-identical to the host refined solve.
* Stream order: values produced by a kernel on the caller's stream just
  before the call, with no host synchronisation, are the ones factored.
* Full 64^3 convection-diffusion with the plan the benchmark tunes to
  (small_lmax 64, block_cols 64, upd_group 4, small_wave_min 16384), in
  fp64 and fp32 with split-bf16 updates: backward error <= 1e-14 within
  three refinement steps.
* A NaN in the values is never certified: the refined solve returns -6 and
  a non-finite backward error.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle.oracle as orc
from helpers import backward_error, csc1
from sparspak_hip import SparseSolver, _lib, findorder, symbolicfactor
from sparspak_hip.gallery import convdiff3d, uniform

pytestmark = pytest.mark.gpu

BENCH_PLAN = (64, 64, 4, 16384)
PREC = {"fp64": 0, "fp32": 1, "fp32_bf16x3": 4}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _handle(A):
    s = SparseSolver(A)
    findorder(s)
    symbolicfactor(s)
    s.slvr._upload()
    return s, s.slvr._dev.L, s.slvr._dev.h


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _download(s):
    lnz, unz, ipiv = s.slvr.download()
    return lnz, unz, ipiv


def test_device_values_match_host_values_and_oracle():
    import torch
    _say("case device values 16^3 fp64 and fp32")
    A = convdiff3d(16)
    n = A.shape[0]
    s, L, h = _handle(A)
    nz = np.ascontiguousarray(A.data, dtype=np.float64)
    d_nz = torch.from_numpy(nz).cuda()
    b = A @ uniform(7, n, -1.0, 1.0)
    for prec in ("fp64", "fp32"):
        _lib.check(L.spkhip_inmatrix(h, _lib.pd(nz)), "inmatrix")
        _lib.check(L.spkhip_factor(h, PREC[prec]), "factor")
        host = _download(s)
        xh = b.copy()
        rc_h = L.spkhip_solve_refine(h, 1, _lib.pd(xh), n, 1e-14, 10)
        torch.cuda.synchronize()
        _lib.check(L.spkhip_inmatrix_device(h, _vp(d_nz)), "inmatrix_device")
        _lib.check(L.spkhip_factor(h, PREC[prec]), "factor")
        dev = _download(s)
        for a, c in zip(host, dev):
            assert np.array_equal(a, c), f"{prec}: device-value factor differs from host-value factor"
        d_x = torch.from_numpy(b.copy()).cuda()
        rc_d = L.spkhip_solve_refine_device(h, 1, _vp(d_x), n, 1e-14, 10)
        assert rc_h == 0 and rc_d == 0
        xd = d_x.cpu().numpy()
        assert np.array_equal(xh, xd), f"{prec}: device refined solve differs from host"
        assert backward_error(A, xd, b) <= 1e-14
    # the fp64 factor against the oracle
    _lib.check(L.spkhip_inmatrix_device(h, _vp(d_nz)), "inmatrix_device")
    _lib.check(L.spkhip_factor(h, 0), "factor")
    lnz, unz, ipiv = _download(s)
    _, cp, rv, nzv = csc1(A)
    ol, ou, op, rc, _ = orc.factor_csc(s.slvr.sym, cp, rv, nzv)
    assert rc == 0
    assert np.array_equal(ipiv, op)
    assert np.max(np.abs(lnz - ol)) <= 1e-12 * np.max(np.abs(ol))
    assert np.max(np.abs(unz - ou)) <= 1e-12 * np.max(np.abs(ou))


def test_device_values_follow_caller_stream():
    """Values and right-hand side written by kernels on torch's stream right
    before the calls, no torch.cuda.synchronize() in between."""
    import torch
    _say("case caller stream order 12^3")
    A = convdiff3d(12)
    n = A.shape[0]
    s, L, h = _handle(A)
    nz = np.ascontiguousarray(A.data, dtype=np.float64)
    b = A @ uniform(8, n, -1.0, 1.0)
    _lib.check(L.spkhip_inmatrix(h, _lib.pd(nz)), "inmatrix")
    _lib.check(L.spkhip_factor(h, 0), "factor")
    ref = _download(s)
    xr = b.copy()
    _lib.check(L.spkhip_solve_refine(h, 1, _lib.pd(xr), n, 1e-14, 10), "solve_refine")
    _lib.check(L.spkhip_set_caller_stream(
        h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "set_caller_stream")
    src = torch.from_numpy(nz).cuda()
    bsrc = torch.from_numpy(b).cuda()
    torch.cuda.synchronize()
    for rep in range(3):
        # a long queue of kernels ahead of the producers keeps them pending
        # when the library is called
        big = torch.ones(1 << 24, dtype=torch.float64, device="cuda")
        for _ in range(20):
            big.mul_(1.0)
        d_nz = torch.zeros_like(src)
        d_nz.add_(src)
        _lib.check(L.spkhip_inmatrix_device(h, _vp(d_nz)), "inmatrix_device")
        _lib.check(L.spkhip_factor(h, 0), "factor")
        got = _download(s)
        for a, c in zip(ref, got):
            assert np.array_equal(a, c), f"rep {rep}: factor of stale values"
        for _ in range(20):
            big.mul_(1.0)
        d_x = torch.zeros_like(bsrc)
        d_x.add_(bsrc)
        _lib.check(L.spkhip_solve_refine_device(h, 1, _vp(d_x), n, 1e-14, 10), "refine")
        assert np.array_equal(d_x.cpu().numpy(), xr), f"rep {rep}: solve of a stale rhs"
    _lib.check(L.spkhip_set_caller_stream(h, None), "set_caller_stream")


@pytest.mark.parametrize("prec", ["fp64", "fp32_bf16x3"])
def test_bench_path_full_64cubed(prec):
    import torch
    _say(f"case bench path convdiff3d 64^3 {prec} plan {BENCH_PLAN}")
    A = convdiff3d(64)
    n = A.shape[0]
    s, L, h = _handle(A)
    tu = _lib.Tuning(*BENCH_PLAN, 0.0)
    _lib.check(L.spkhip_set_tuning(h, ctypes.byref(tu)), "set_tuning")
    xtrue = uniform(2, n, -1.0, 1.0)
    b = A @ xtrue
    d_nz = torch.from_numpy(np.ascontiguousarray(A.data, dtype=np.float64)).cuda()
    d_x = torch.from_numpy(b.copy()).cuda()
    torch.cuda.synchronize()
    _lib.check(L.spkhip_inmatrix_device(h, _vp(d_nz)), "inmatrix_device")
    _lib.check(L.spkhip_factor(h, PREC[prec]), "factor")
    rc = L.spkhip_solve_refine_device(h, 1, _vp(d_x), n, 1e-14, 10)
    st = _lib.Stats()
    L.spkhip_stats(h, ctypes.byref(st))
    x = d_x.cpu().numpy()
    eta = backward_error(A, x, b)
    _say(f"  rc {rc} iters {st.refine_iters} eta {eta:.3e} (library {st.backward_error:.3e}) "
         f"factor {st.t_factor_ms:.1f} ms refine {st.t_refine_ms:.1f} ms")
    assert rc == 0
    assert eta <= 1e-14
    assert st.backward_error <= 1e-14
    assert st.refine_iters <= 3
    assert np.max(np.abs(x - xtrue)) <= 1e-10


def test_nan_values_are_not_certified():
    import torch
    _say("case NaN value 10^3 fp32 refine")
    A = convdiff3d(10).tocsc()
    n = A.shape[0]
    s, L, h = _handle(A)
    nz = np.ascontiguousarray(A.data, dtype=np.float64).copy()
    nz[len(nz) // 2] = np.nan
    d_nz = torch.from_numpy(nz).cuda()
    _lib.check(L.spkhip_inmatrix_device(h, _vp(d_nz)), "inmatrix_device")
    L.spkhip_factor(h, 1)   # a NaN pivot need not be zero: rc 0 or -1
    b = uniform(9, n, -1.0, 1.0)
    d_x = torch.from_numpy(b.copy()).cuda()
    rc = L.spkhip_solve_refine_device(h, 1, _vp(d_x), n, 1e-14, 3)
    st = _lib.Stats()
    L.spkhip_stats(h, ctypes.byref(st))
    _say(f"  rc {rc} backward_error {st.backward_error}")
    assert rc == -6
    assert not np.isfinite(st.backward_error)
