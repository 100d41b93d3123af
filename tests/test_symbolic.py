"""Host symbolic phase: golden vectors from the reference's tests, brute-force
symbolic elimination, and the supernode split rule (CPU only)."""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from helpers import brute_symbolic, csc1, random_pattern, symbolic
from sparspak_hip.gallery import convdiff3d, elasticity3d, laplace2d
from sparspak_hip.problem import Problem
from sparspak_hip.symbolic import Graph, SymbolicFactor


def _fig313_problem(fx, key):
    d = fx[key]
    p = Problem(6, 6)
    A = sp.csc_matrix((np.ones(len(d["I"])), (np.array(d["I"]) - 1, np.array(d["J"]) - 1)),
                      shape=(6, 6))
    p.insparse(A)
    return p, d


def test_graph_from_problem_fig313(fixtures):
    # test_graph.jl mgrap001
    p, d = _fig313_problem(fixtures, "fig313")
    n, cp, rv, _ = p.csc()
    g = Graph.from_csc(n, cp, rv)
    assert g.xadj.tolist() == d["xadj"]
    assert g.adj.tolist() == d["adj"]
    assert g.isstructuresymmetric()  # mgrap002


def test_graph_from_csc_fig313(fixtures):
    # test_cscinterface.jl csc_mgrap001 (same Figure 3.1.3 adjacency)
    d = fixtures["fig313"]
    A = sp.csc_matrix((np.ones(len(d["I"])), (np.array(d["I"]) - 1, np.array(d["J"]) - 1)),
                      shape=(6, 6))
    n, cp, rv, _ = csc1(A)
    g = Graph.from_csc(n, cp, rv)
    assert g.xadj.tolist() == d["xadj"]
    assert g.adj.tolist() == d["adj"]


def test_makestructuresymmetric_fig313(fixtures):
    # test_graph.jl mgrap003/mgrap004
    p, d = _fig313_problem(fixtures, "fig313_unsym")
    n, cp, rv, _ = p.csc()
    g = Graph.from_csc(n, cp, rv)
    assert not g.isstructuresymmetric()
    g.makestructuresymmetric()
    assert g.isstructuresymmetric()
    assert g.xadj.tolist() == d["xadj"]
    assert g.adj.tolist() == d["adj"]


def test_makestructuresymmetric_is_sorted_union():
    # symmetrised graph == sorted pattern of A + A^T without the diagonal
    for seed in range(20):
        A = random_pattern(25, 0.12, seed, symmetric=False)
        n, cp, rv, _ = csc1(A)
        g = Graph.from_csc(n, cp, rv)
        g.makestructuresymmetric()
        S = ((abs(A) + abs(A).T) != 0).tocsc().astype(float)
        S.setdiag(0)
        S.eliminate_zeros()
        S.sort_indices()
        assert g.xadj.tolist() == (S.indptr + 1).tolist()
        assert g.adj.tolist() == (S.indices + 1).tolist()


def _tridiag11(fx):
    n = fx["tridiag11"]["n"]
    p = Problem(n, n)
    # maketridiagproblem (test_sparse_method.jl:67-80), insertion order kept
    for i in range(1, n):
        p.inaij(i + 1, i, -1.0)
        p.inaij(i, i, 4.0)
        p.inaij(i, i + 1, -1.0)
        p.inbi(i, 2.0 * i)
    p.inaij(n, n, 4.0)
    p.inbi(n, 3.0 * n + 1.0)
    return p


def test_tridiag11_ordering(fixtures):
    # msprs003
    d = fixtures["tridiag11"]
    p = _tridiag11(fixtures)
    n, cp, rv, _ = p.csc()
    s = SymbolicFactor(n, Graph.from_csc(n, cp, rv))
    s.findorder()
    assert s.order.rperm.tolist() == d["rperm"]
    assert s.order.rinvp.tolist() == d["rinvp"]
    assert s.order.cperm.tolist() == d["rperm"]
    assert s.order.cinvp.tolist() == d["rinvp"]


def test_tridiag11_symbolic(fixtures):
    # msprs004
    d = fixtures["tridiag11"]
    p = _tridiag11(fixtures)
    n, cp, rv, _ = p.csc()
    s = SymbolicFactor(n, Graph.from_csc(n, cp, rv))
    s.findorder()
    s.symbolicfactor()
    assert s.xlnz.tolist() == d["xlnz"]
    assert s.xunz.tolist() == d["xunz"]
    assert s.xlindx.tolist() == d["xlindx"]
    assert s.lindx.tolist() == d["lindx"]


def _check_against_brute(A, s):
    n = A.shape[0]
    perm = s.order.rperm
    assert sorted(perm.tolist()) == list(range(1, n + 1))
    assert np.array_equal(s.order.rinvp[perm - 1], np.arange(1, n + 1))
    cols, parent = brute_symbolic(A, perm)
    # elimination tree and column counts
    assert [p + 1 if p >= 0 else 0 for p in parent] == s.parent.tolist()
    assert [len(c) for c in cols] == s.colcnt.tolist()
    assert s.nnzl == sum(len(c) for c in cols)
    # postorder: every subtree is a contiguous range ending at its root
    for j in range(n):
        if parent[j] >= 0:
            assert parent[j] > j
    # supernodes: index list = structure of the first column
    for J in range(s.nsuper):
        f = s.xsuper[J] - 1
        w = s.xsuper[J + 1] - s.xsuper[J]
        I = s.lindx[s.xlindx[J] - 1: s.xlindx[J + 1] - 1] - 1
        assert I.tolist() == sorted(cols[f])
        assert I[:w].tolist() == list(range(f, f + w))
        for c in range(f, f + w):  # nested column structures
            assert sorted(cols[c]) == I[c - f:].tolist()
        assert w <= 1.5 * s.maxblocksize
    # panel pointers are prefix sums of the panel sizes
    jl = np.diff(s.xlindx)
    w = np.diff(s.xsuper)
    if s.kind == "lu":
        assert s.xlnz[-1] - 1 == int(np.sum(jl * w))
        assert s.xunz[-1] - 1 == int(np.sum((jl - w) * w))
    # supernodes are maximal up to the split: neighbouring supernodes that
    # could merge (fundamental rule) were split by the width limit only
    for J in range(s.nsuper - 1):
        l = s.xsuper[J + 1] - 2
        merge = parent[l] == l + 1 and len(cols[l]) == len(cols[l + 1]) + 1
        if merge:
            assert s.xsuper[J + 1] - s.xsuper[J] > 0


@pytest.mark.parametrize("seed", range(12))
def test_symbolic_vs_bruteforce(seed):
    sym = seed % 2 == 0
    A = random_pattern(30 + seed, 0.08, seed, symmetric=sym)
    s = symbolic(A)
    _check_against_brute(A, s)


def test_symbolic_grid_vs_bruteforce():
    from sparspak_hip.gallery import convdiff3d, laplace2d
    for A in (laplace2d(9, 7), convdiff3d(5, 4, 4)):
        _check_against_brute(A, symbolic(A))
        _check_against_brute(A, symbolic(A, kind="spd"))


def _split_closed_form(f, l, maxsize):
    """First columns of the pieces of a fundamental supernode [f, l]
    (1-based) per SpkSymFct.jl:434-445."""
    size = l - f + 1
    starts = [f]
    if size > maxsize:
        delta = size / (1.0 + size / maxsize)
        limit = math.floor((l + 1) - delta / 2.0)
        k = 1
        idx = math.floor(f + delta)
        while idx < limit:
            starts.append(idx)
            k += 1
            idx = math.floor(f + k * delta)
    return starts


@pytest.mark.parametrize("m,maxsize", [(31, 30), (45, 30), (100, 30), (61, 60), (150, 60),
                                       (97, 30)])
def test_supernode_split(m, maxsize):
    # a dense m x m matrix is one fundamental supernode of width m
    A = sp.csc_matrix(np.ones((m, m)) + m * np.eye(m))
    s = symbolic(A, kind="lu" if maxsize == 30 else "spd")
    assert s.xsuper.tolist() == _split_closed_form(1, m, maxsize) + [m + 1]
    # the rule does not cap the last piece at the limit: 100 columns at
    # limit 30 end with a 31-column piece; the widest pieces it can make are
    # 45 (limit 30) and 90 (limit 60)
    w = np.diff(s.xsuper)
    assert w[:-1].max(initial=0) <= maxsize
    assert w.max() <= 1.5 * maxsize


def test_dense_block_example():
    # 100 columns, limit 30: delta = 100 / (1 + 100/30) = 23.08; limit
    # floor(101 - 11.54) = 89; pieces start at 1, 24, 47, 70 (93 >= 89 stops)
    A = sp.csc_matrix(np.ones((100, 100)) + 100 * np.eye(100))
    s = symbolic(A)
    assert s.xsuper.tolist() == [1, 24, 47, 70, 101]


def test_mmd_fill_sanity():
    # MMD on a 2D grid keeps fill far below the natural band ordering
    from sparspak_hip.gallery import laplace2d
    A = laplace2d(30)
    s = symbolic(A)
    band_fill = 30 * 30 * 31  # ~ n * bandwidth for the lexicographic order
    assert s.nnzl < 0.5 * band_fill


def test_findorderperm_identity():
    from sparspak_hip.gallery import laplace2d
    A = laplace2d(6)
    n, cp, rv, _ = csc1(A)
    s = SymbolicFactor(n, Graph.from_csc(n, cp, rv))
    s.findorderperm(np.arange(1, n + 1))
    s.symbolicfactor()
    _check_against_brute(A, s)


def _fronts(s):
    """Chains of consecutive supernodes whose index lists nest (piece k+1's
    list is piece k's without its own columns): the device's fronts."""
    xs, xl, li = np.asarray(s.xsuper), np.asarray(s.xlindx), np.asarray(s.lindx)
    nj = lambda J: xs[J + 1] - xs[J]
    jl = lambda J: xl[J + 1] - xl[J]
    fronts = []
    for J in range(s.nsuper):
        if (J > 0 and jl(J - 1) > nj(J - 1) and jl(J - 1) - nj(J - 1) == jl(J)
                and li[xl[J - 1] - 1 + nj(J - 1)] == xs[J]):
            fronts[-1].append(J)
        else:
            fronts.append([J])
    return fronts


@pytest.mark.parametrize("kind,gen", [
    ("lu", lambda: random_pattern(300, 0.03, 4, False, False)),
    ("lu", lambda: convdiff3d(14)),
    ("spd", lambda: elasticity3d(5)),
    ("spd", lambda: laplace2d(30)),
])
def test_front_piece_column_stride(kind, gen):
    # the solve kernels address a front's factor columns from each piece's
    # first column: inside a front with index list length L, every column of
    # a piece starting at front column cq holds L - cq entries of lnz, and the
    # columns are consecutive in lnz
    s = symbolic(gen(), kind=kind)
    xs, xl, xlnz = np.asarray(s.xsuper), np.asarray(s.xlindx), np.asarray(s.xlnz)
    for fr in _fronts(s):
        L = xl[fr[0] + 1] - xl[fr[0]]
        fc0 = xs[fr[0]] - 1
        for p in fr:
            cq = xs[p] - 1 - fc0
            cols = np.arange(xs[p] - 1, xs[p + 1] - 1)
            assert np.all(xlnz[cols + 1] - xlnz[cols] == L - cq)
