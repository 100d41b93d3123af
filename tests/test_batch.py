"""Batched systems: the concatenated (block-diagonal) symbolic factor is a
valid symbolic factor of the block-diagonal matrix, and the CPU oracle
factors it into exactly the per-system factors (CPU only)."""
import numpy as np
import scipy.sparse as sp

import oracle.oracle as orc
from helpers import csc1
from sparspak_hip.batch import BatchedSolver
from sparspak_hip.gallery import convdiff3d, laplace2d, uniform
from sparspak_hip.symbolic import SymbolicFactor


class _Sym:
    """Duck-typed symbolic factor for the oracle from concatenated arrays."""

    def __init__(self, b):
        self.n = b.n
        self.nsuper = b.nsuper
        self.kind = b.kind
        self.xsuper, self.snode, self.xlindx, self.lindx = b.xsuper, b.snode, b.xlindx, b.lindx
        self.xlnz, self.xunz = b.xlnz, b.xunz
        self.order = type("O", (), {"rperm": b.perm, "rinvp": b.invp})()
        self.nlnz = int(b.xlnz[-1] - 1)
        self.nunz = int(b.xunz[-1] - 1)


def _systems():
    out = []
    for k, side in enumerate([5, 9, 7, 9, 6]):
        out.append(laplace2d(side) + sp.diags(np.full(side * side, 0.1 * (k + 1))))
    out.append(convdiff3d(4))
    return out


def test_concatenated_symbolic_matches_per_system():
    mats = _systems()
    b = BatchedSolver(mats)
    assert b.n == sum(A.shape[0] for A in mats)
    for (n, cp, rv, nz, s), off in zip(b.systems, b.offsets):
        assert np.array_equal(b.perm[off:off + n], s.order.rperm + off)
    # pattern reuse for identical systems
    assert b.systems[1][4] is b.systems[3][4] or b.systems[1][4].nsuper == b.systems[3][4].nsuper


def test_oracle_on_concatenation_equals_per_system():
    mats = _systems()
    b = BatchedSolver(mats)
    sym = _Sym(b)
    lnz, unz, ipiv, rc, _ = orc.factor_csc(sym, b.colptr, b.rowval, b.nzval)
    assert rc == 0
    lo = uo = 0
    for n, cp, rv, nz, s in b.systems:
        l1, u1, p1, rc1, _ = orc.factor_csc(s, cp, rv, nz)
        assert np.array_equal(lnz[lo:lo + l1.size], l1)
        assert np.array_equal(unz[uo:uo + u1.size], u1)
        lo += l1.size
        uo += u1.size
    B = b.block_diagonal()
    rhs = uniform(5, b.n, -1, 1)
    x = orc.solve(sym, lnz, unz, ipiv, rhs.copy())
    r = rhs - B @ x
    assert np.max(np.abs(r)) < 1e-13
