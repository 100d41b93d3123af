"""The A -> slot map of the upload (spk_amap, the host phase of _inmatrix!,
SpkSparseBase.jl:302-372, SparseCSCInterface.jl:112-163, SPD
SpkSparseSpdBase.jl:274-282): integer-identical for every host thread count,
and the values scattered through it equal the oracle's _inmatrix! output
exactly (LU and SPD); an entry outside the structure gives the same map and
return code for every thread count."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle.oracle as orc
from helpers import csc1, random_pattern, symbolic
from sparspak_hip import _lib
from sparspak_hip.gallery import convdiff3d, laplace2d


def _amap(sym, cp, rv, nthreads):
    L = _lib.lib()
    P = _lib.p64
    kind = 0 if sym.kind == "lu" else 1
    xunz = sym.xunz if sym.xunz is not None else np.zeros(sym.n + 1, np.int64)
    out = np.zeros(int(cp[-1] - 1), dtype=np.int64)
    rc = L.spk_amap(sym.n, kind, sym.nsuper, P(sym.xsuper), P(sym.snode), P(sym.xlindx),
                    P(sym.lindx), P(sym.xlnz), P(_lib.as_i64(xunz)), P(sym.order.rinvp),
                    P(_lib.as_i64(cp)), P(_lib.as_i64(rv)), nthreads, P(out))
    return rc, out


@pytest.mark.parametrize("case", ["random_lu", "convdiff_lu", "laplace_spd", "random_spd"])
def test_amap_threads_and_oracle(case):
    if case == "random_lu":
        A, kind = random_pattern(300, 0.02, 7, symmetric=False), "lu"
    elif case == "convdiff_lu":
        A, kind = convdiff3d(10), "lu"
    elif case == "laplace_spd":
        A, kind = laplace2d(30, shift=0.5), "spd"
    else:
        A, kind = random_pattern(200, 0.03, 8, symmetric=True), "spd"
    sym = symbolic(A, kind=kind)
    n, cp, rv, nz = csc1(A)
    rc1, m1 = _amap(sym, cp, rv, 1)
    assert rc1 == 0
    for t in (2, 3, 8, 16):
        rct, mt = _amap(sym, cp, rv, t)
        assert rct == rc1 and np.array_equal(mt, m1), f"{t} threads"
    lnz = np.zeros(int(sym.xlnz[-1] - 1))
    unz = np.zeros(int(sym.xunz[-1] - 1)) if sym.kind == "lu" else None
    for k, slot in enumerate(m1):
        if slot >= 0:
            lnz[slot] += nz[k]
        elif slot <= -2:
            unz[-(slot + 2)] += nz[k]
    ol, ou, _ = orc.inmatrix(sym, cp, rv, nz)
    assert np.array_equal(lnz, ol)
    if unz is not None:
        assert np.array_equal(unz, ou)


def test_amap_entry_outside_structure_same_for_all_threads():
    A = laplace2d(12)
    sym = symbolic(A, kind="lu")
    B = sp.lil_matrix(A)
    B[0, 140] = 1.0      # a far fill-in entry the symbolic factor lacks
    B[143, 3] = 1.0
    n, cp, rv, nz = csc1(sp.csc_matrix(B))
    rc1, m1 = _amap(sym, cp, rv, 1)
    assert rc1 in (0, -1)
    for t in (2, 5, 16):
        rct, mt = _amap(sym, cp, rv, t)
        assert rct == rc1 and np.array_equal(mt, m1)
