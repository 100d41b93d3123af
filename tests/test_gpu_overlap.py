"""The wide trailing updates that run on the second stream (beside the next
block group's panel work) give exactly the factor of the serial launch order
(the per-launch profile runs everything on one stream): a missing wait
between the two streams would show as a difference."""
import numpy as np
import pytest

from sparspak_hip import SparseSolver, SparseSpdSolver, factor, findorder, inmatrix, symbolicfactor
from sparspak_hip.gallery import convdiff3d, laplace2d

pytestmark = pytest.mark.gpu


def _factors(s, precision, profile):
    L, h = s.slvr._dev.L, s.slvr._dev.h
    L.spkhip_profile_enable(h, 1 if profile else 0)
    s._factordone = False
    factor(s, precision)
    L.spkhip_profile_enable(h, 0)
    if precision == "fp64":
        return s.slvr.download()
    import ctypes
    from sparspak_hip import _lib
    sym = s.slvr.sym
    lnz = np.zeros(sym.nlnz, dtype=np.float32)
    unz = np.zeros(max(sym.nunz, 1), dtype=np.float32)
    ipiv = np.zeros(s.slvr.n, dtype=np.int64)
    _lib.check(L.spkhip_download_f32(h, _lib.pf(lnz), _lib.pf(unz), _lib.p64(ipiv)), "download")
    return lnz, unz, ipiv


@pytest.mark.parametrize("precision,group", [("fp64", 8), ("fp64", 1), ("fp32", 2),
                                             ("fp32_bf16x3", 3)])
def test_two_streams_match_serial_order(precision, group):
    print(f"case overlap convdiff3d 22^3 {precision} upd_group={group}", flush=True)
    A = convdiff3d(22)
    s = SparseSolver(A)
    findorder(s)
    symbolicfactor(s)
    inmatrix(s)
    s.slvr.set_tuning(upd_group=group, block_cols=64)
    ser = _factors(s, precision, True)
    for _ in range(2):
        par = _factors(s, precision, False)
        for a, b in zip(ser, par):
            assert np.array_equal(a, b)


def test_two_streams_spd():
    print("case overlap laplace2d 120 SPD", flush=True)
    s = SparseSpdSolver(laplace2d(120))
    findorder(s)
    symbolicfactor(s)
    inmatrix(s)
    s.slvr.set_tuning(upd_group=1)
    ser = _factors(s, "fp64", True)
    par = _factors(s, "fp64", False)
    for a, b in zip(ser, par):
        assert np.array_equal(a, b)
