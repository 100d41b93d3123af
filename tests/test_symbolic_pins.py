"""Pins of the host integer work the device layout depends on.

* MMD (SpkMMD.jl:78-633): hand-traced orderings of five tiny graphs with
  ties (tests/golden/mmd_traces.json, every tie decision with its reference
  line); the literal Python restatement (oracle/symbolic_ref.py) must
  reproduce each trace event for event, and the product's C++ restatement
  (spk_mmd) each ordering; then C++ and Python agree on random graphs
  (disconnected, isolated nodes, grids, stars, dense blocks).
* Postorders (SpkETree.jl:106-274): a hand-traced weighted postorder with a
  tie and a lighter child appended last; C++ and Python agree on random
  elimination trees, plain and weighted.
* Relative-index maps (SpkSpdMMOps.jl:70-102): every map the device receives
  (spk_front_maps, the schedule spkhip_symbolic_upload builds) equals the
  closed form of _ldindx!/_igathr! for every (updating supernode, target
  supernode) pair the reference's left-looking factor forms.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

from helpers import symbolic
from oracle import symbolic_ref as R
from sparspak_hip import _lib
from sparspak_hip._lib import p64
from sparspak_hip.gallery import convdiff3d, laplace2d

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def traces():
    with open(os.path.join(HERE, "golden", "mmd_traces.json")) as f:
        return json.load(f)


def _graph(n, edges):
    nb = {i: set() for i in range(1, n + 1)}
    for a, b in edges:
        nb[a].add(b)
        nb[b].add(a)
    xadj, adj = [1], []
    for i in range(1, n + 1):
        adj += sorted(nb[i])
        xadj.append(xadj[-1] + len(nb[i]))
    return xadj, adj


def _cmmd(n, xadj, adj):
    L = _lib.lib()
    xa = np.array(xadj, dtype=np.int64)
    ad = np.array(adj if len(adj) else [0], dtype=np.int64)
    perm = np.zeros(n, np.int64)
    invp = np.zeros(n, np.int64)
    L.spk_mmd(n, p64(xa), p64(ad), p64(perm), p64(invp))
    return [int(v) for v in perm], [int(v) for v in invp]


def _cpostorder(n, parent, perm, invp, weight=None):
    L = _lib.lib()
    pa = np.array(parent, np.int64)
    pe = np.array(perm, np.int64)
    iv = np.array(invp, np.int64)
    w = None if weight is None else np.array(weight, np.int64)
    L.spk_postorder(n, p64(pa), p64(pe), p64(iv), None if w is None else p64(w))
    return ([int(v) for v in pa], [int(v) for v in pe], [int(v) for v in iv],
            None if w is None else [int(v) for v in w])


def _ev(trace):
    return [list(t[:-1]) + [list(t[-1])] if isinstance(t[-1], list) else list(t) for t in trace]


def test_mmd_hand_traces(traces):
    for case in traces["mmd"]:
        n = case["n"]
        xadj, adj = _graph(n, case["edges"])
        assert xadj == case["xadj"] and adj == case["adj"], case["name"]
        tr = []
        perm, invp = R.mmd(n, xadj, adj, tr)
        assert _ev(tr) == case["events"], f"{case['name']}: oracle trace differs from the hand trace"
        assert perm == case["perm"] and invp == case["invp"], case["name"]
        cperm, cinvp = _cmmd(n, xadj, adj)
        assert cperm == case["perm"] and cinvp == case["invp"], \
            f"{case['name']}: product MMD differs from the hand trace"


def _random_graphs(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        kind = t % 5
        if kind == 0:       # Erdos-Renyi with isolated nodes
            n = int(rng.integers(1, 45))
            M = np.triu(rng.random((n, n)) < rng.uniform(0.02, 0.4), 1)
            M = M | M.T
            iso = rng.random(n) < 0.2
            M[iso, :] = False
            M[:, iso] = False
        elif kind == 1:     # disjoint union of paths and cycles
            n = 0
            edges = []
            for _ in range(int(rng.integers(1, 5))):
                m = int(rng.integers(2, 9))
                edges += [(n + i, n + i + 1) for i in range(m - 1)]
                if rng.random() < 0.5 and m > 2:
                    edges.append((n, n + m - 1))
                n += m
            M = np.zeros((n, n), bool)
            for a, b in edges:
                M[a, b] = M[b, a] = True
        elif kind == 2:     # grid with holes
            a, b = int(rng.integers(2, 7)), int(rng.integers(2, 7))
            n = a * b
            M = np.zeros((n, n), bool)
            for y in range(b):
                for x in range(a):
                    v = x + a * y
                    if x + 1 < a and rng.random() < 0.9:
                        M[v, v + 1] = M[v + 1, v] = True
                    if y + 1 < b and rng.random() < 0.9:
                        M[v, v + a] = M[v + a, v] = True
        elif kind == 3:     # stars sharing leaves
            n = int(rng.integers(3, 30))
            M = np.zeros((n, n), bool)
            for c in rng.choice(n, size=max(1, n // 6), replace=False):
                for leaf in rng.choice(n, size=int(rng.integers(1, n)), replace=False):
                    if leaf != c:
                        M[c, leaf] = M[leaf, c] = True
        else:               # dense blocks joined by a few edges
            n = 0
            blocks = []
            for _ in range(int(rng.integers(1, 4))):
                m = int(rng.integers(1, 8))
                blocks.append(range(n, n + m))
                n += m
            M = np.zeros((n, n), bool)
            for bl in blocks:
                for i in bl:
                    for j in bl:
                        if i != j:
                            M[i, j] = True
            for _ in range(len(blocks)):
                i, j = rng.integers(0, n, 2)
                if i != j:
                    M[i, j] = M[j, i] = True
        xadj, adj = [1], []
        for i in range(n):
            nb = list(np.nonzero(M[i])[0] + 1)
            adj += [int(v) for v in nb]
            xadj.append(xadj[-1] + len(nb))
        out.append((n, xadj, adj))
    return out


def test_mmd_product_equals_restatement_random():
    for n, xadj, adj in _random_graphs(400, 11):
        if n == 0:
            continue
        assert list(R.mmd(n, xadj, adj)) == list(_cmmd(n, xadj, adj)), (n, xadj, adj)


def test_weighted_postorder_hand_trace(traces):
    for case in traces["weighted_postorder"]:
        n = case["n"]
        tr = []
        fson, brother = R.weightedbinarytree(n, case["parent"], case["weight"], tr)
        assert [list(t) for t in tr] == case["events"]
        invpos, _ = R.postordertree(n, fson, brother, [0] + case["parent"])
        assert invpos[1:] == case["invpos"]
        ident = list(range(1, n + 1))
        par, rp, ri, w = R.getpostorder(n, case["parent"], ident, ident, case["weight"])
        assert par == case["new_parent"] and w == case["new_weight"]
        assert rp == case["rperm_from_identity"] and ri == case["rinvp_from_identity"]
        cpar, crp, cri, cw = _cpostorder(n, case["parent"], ident, ident, case["weight"])
        assert (cpar, crp, cri, cw) == (par, rp, ri, w), "product weighted postorder differs"


def test_postorders_product_equals_restatement_random():
    rng = np.random.default_rng(12)
    for n, xadj, adj in _random_graphs(300, 13):
        if n < 2:
            continue
        perm, invp = R.mmd(n, xadj, adj)
        parent = R.findetree(n, xadj, adj, perm, invp)
        # plain postorder, then the weighted one with random weights (ties
        # frequent) as _symbolicfactor! does with the column counts
        p1, rp1, ri1, _ = R.getpostorder(n, parent, perm, invp)
        c1 = _cpostorder(n, parent, perm, invp)
        assert (c1[0], c1[1], c1[2]) == (p1, rp1, ri1)
        w = [int(v) for v in rng.integers(1, 4, n)]
        p2, rp2, ri2, w2 = R.getpostorder(n, p1, rp1, ri1, w)
        c2 = _cpostorder(n, p1, rp1, ri1, w)
        assert c2 == (p2, rp2, ri2, w2)


def test_etree_product_equals_restatement():
    L = _lib.lib()
    for n, xadj, adj in _random_graphs(150, 14):
        if n == 0:
            continue
        perm, invp = R.mmd(n, xadj, adj)
        par = np.zeros(n, np.int64)
        L.spk_etree(n, p64(np.array(xadj, np.int64)), p64(np.array(adj or [0], np.int64)),
                    p64(np.array(perm, np.int64)), p64(np.array(invp, np.int64)), p64(par))
        assert [int(v) for v in par] == R.findetree(n, xadj, adj, perm, invp)


def _front_maps(s):
    L = _lib.lib()
    ns = s.nsuper
    nsub = int(s.xlindx[-1] - 1)
    fsp0 = np.zeros(ns, np.int64)
    fnp = np.zeros(ns, np.int64)
    fpar = np.zeros(ns, np.int64)
    frel = np.zeros(ns + 1, np.int64)
    rel = np.zeros(max(nsub, 1), np.int64)
    nf = L.spk_front_maps(s.n, ns, p64(s.xsuper), p64(s.snode), p64(s.xlindx), p64(s.lindx),
                          p64(fsp0), p64(fnp), p64(fpar), p64(frel), p64(rel))
    assert nf > 0
    return int(nf), fsp0[:nf], fnp[:nf], fpar[:nf], frel[:nf + 1], rel


def _ldindx(jlen, lindx_j, mp):
    """_ldindx! (SpkSpdMMOps.jl:70-79): map[I_J[t]] = jlen - t (t 1-based)."""
    for t in range(1, jlen + 1):
        mp[lindx_j[t - 1]] = jlen - t


def _igathr(klen, lindx_k, mp):
    """_igathr! (SpkSpdMMOps.jl:94-102): relind[i] = map[lindx_k[i]]."""
    return [mp[lindx_k[i]] for i in range(klen)]


MATRICES = {
    "laplace2d_12": lambda: laplace2d(12),
    "convdiff3d_7": lambda: convdiff3d(7),
    "rand_unsym": lambda: sp.random(90, 90, density=0.04, random_state=3, format="csc")
    + sp.eye(90),
    "band": lambda: sp.diags([1.0] * 9, list(range(-4, 5)), shape=(150, 150), format="csc"),
}


@pytest.mark.parametrize("name", list(MATRICES))
@pytest.mark.parametrize("kind", ["lu", "spd"])
def test_relative_index_maps_closed_form(name, kind):
    A = sp.csc_matrix(MATRICES[name]())
    if kind == "spd":
        A = A + A.T
    s = symbolic(A, kind=kind)
    nf, fsp0, fnp, fpar, frel, rel = _front_maps(s)
    xs, xl, li = s.xsuper, s.xlindx, s.lindx
    ns = s.nsuper

    def lst(J):  # index list of supernode J (0-based J), 1-based values
        return [int(v) for v in li[xl[J] - 1: xl[J + 1] - 1]]

    front_of = np.zeros(ns, np.int64)
    for f in range(nf):
        front_of[fsp0[f] - 1: fsp0[f] - 1 + fnp[f]] = f
    # the pieces of a front nest: each piece's list = the previous one's
    # without its own columns (what makes the front one dense matrix)
    for f in range(nf):
        for J in range(fsp0[f], fsp0[f] - 1 + fnp[f]):
            prev = lst(J - 1)
            w = int(xs[J] - xs[J - 1])
            assert lst(J) == prev[w:]
    checked = 0
    mp = np.zeros(s.n + 1, np.int64)
    for f in range(nf):
        p = int(fpar[f]) - 1
        if p < 0:
            continue
        K = int(fsp0[f] - 1 + fnp[f] - 1)           # last piece: the updating supernode
        Ik = lst(K)
        wk = int(xs[K + 1] - xs[K])
        upd = Ik[wk:]                               # the front's update rows
        Lp = len(lst(int(fsp0[p]) - 1))
        got = [int(v) for v in rel[frel[f] - 1: frel[f] - 1 + len(upd)]]
        assert len(upd) == int(frel[f + 1] - frel[f])
        # the device map is taken against the parent front's whole index
        # list (its first piece's list): _ldindx! on it, _igathr! on the
        # child's update rows, position = L_p - 1 - relind
        P0 = int(fsp0[p]) - 1
        IP = lst(P0)
        _ldindx(len(IP), IP, mp)
        relind_front = _igathr(len(upd), upd, mp)
        assert got == [Lp - 1 - r for r in relind_front], f"front {f}"
        checked += len(upd)
        # for every target supernode J inside the parent front, the
        # reference's own pair map (K -> J, over K's rows from J's first
        # column on) gives the same distances from the bottom, because the
        # pieces' lists nest as suffixes of the front's list
        i = 0
        while i < len(upd):
            J = int(s.snode[upd[i] - 1]) - 1
            if front_of[J] != p:
                break           # rows of later ancestors travel in the parent's update matrix
            IJ = lst(J)
            _ldindx(len(IJ), IJ, mp)
            relind = _igathr(len(upd) - i, upd[i:], mp)
            assert relind == relind_front[i:], (f, J)
            last_col = int(xs[J + 1] - 1)
            while i < len(upd) and upd[i] <= last_col:
                i += 1
    assert checked > 0
