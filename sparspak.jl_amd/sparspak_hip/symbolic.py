"""Host symbolic phase: graph, ordering and symbolic factorisation.

Mirrors the reference's storage object and its symbolic phase bodies
(SpkSparseBase.jl:99-251 for LU, SpkSparseSpdBase.jl:84-232 for SPD). The
integer work runs in the library's C++ restatements (spk_* in
include/sparspak_hip.h); this module sequences them exactly as
``_findorder!`` and ``_symbolicfactor!`` do and keeps the resulting arrays
(1-based values, int64) for the numeric phase.
"""
import numpy as np

from . import _lib
from ._lib import as_i64, p64

LU_MAXBLOCKSIZE = 30   # SpkSparseBase.jl:133, SparseCSCInterface.jl:60
SPD_MAXBLOCKSIZE = 60  # SpkSparseSpdBase.jl:111


class Graph:
    """Adjacency structure (xadj, adj) with 1-based values (SpkGraph.jl:31-40)."""

    def __init__(self, nv, xadj, adj):
        self.nv = int(nv)
        self.xadj = as_i64(xadj)
        self.adj = as_i64(adj)

    @property
    def nedges(self):
        return int(self.xadj[-1] - 1)

    @classmethod
    def from_csc(cls, n, colptr, rowval, diagonal=False):
        """Graph(::SparseMatrixCSC) (SparseCSCInterface.jl:12-56): column i's
        row indices in stored order, the diagonal dropped unless asked for."""
        colptr = as_i64(colptr)
        rowval = as_i64(rowval)
        cols = np.repeat(np.arange(1, n + 1, dtype=np.int64), np.diff(colptr))
        keep = np.ones(rowval.shape[0], dtype=bool) if diagonal else (rowval != cols)
        adj = rowval[keep]
        counts = np.bincount(cols[keep] - 1, minlength=n)
        xadj = np.empty(n + 1, dtype=np.int64)
        xadj[0] = 1
        np.cumsum(counts, out=xadj[1:])
        xadj[1:] += 1
        return cls(n, xadj, adj)

    def isstructuresymmetric(self):
        L = _lib.lib()
        return bool(L.spk_isstructuresymmetric(self.nv, p64(self.xadj), p64(self.adj)))

    def makestructuresymmetric(self):
        """SpkGraph.jl:94-268 (+ sortgraph! when changed). In place."""
        L = _lib.lib()
        ne = int(L.spk_symmetrized_nedges(self.nv, p64(self.xadj), p64(self.adj)))
        hx = np.zeros(self.nv + 1, dtype=np.int64)
        ha = np.zeros(max(ne, 1), dtype=np.int64)
        L.spk_makestructuresymmetric(self.nv, p64(self.xadj), p64(self.adj),
                                     p64(hx), p64(ha))
        self.xadj = hx
        self.adj = ha[:ne].copy()
        return True


class Ordering:
    """rperm/rinvp (= cperm/cinvp) of SpkOrdering.jl:66-77; identity at start."""

    def __init__(self, n):
        self.rperm = np.arange(1, n + 1, dtype=np.int64)
        self.rinvp = np.arange(1, n + 1, dtype=np.int64)

    @property
    def cperm(self):
        return self.rperm

    @property
    def cinvp(self):
        return self.rinvp


def mmd(g, order):
    """mmd!(g, order) (SpkMMD.jl:38-42)."""
    L = _lib.lib()
    rc = L.spk_mmd(g.nv, p64(g.xadj), p64(g.adj), p64(order.rperm), p64(order.rinvp))
    if rc != 0:
        raise RuntimeError(f"mmd failed ({rc})")


class SymbolicFactor:
    """The integer state of _SparseBase / _SparseSpdBase after the symbolic
    phase: ordering, etree, colcnt, supernodes and the panel pointers."""

    def __init__(self, n, g, kind="lu", maxblocksize=None):
        self.n = int(n)
        self.g = g
        self.kind = kind
        self.maxblocksize = int(maxblocksize if maxblocksize is not None else
                                (LU_MAXBLOCKSIZE if kind == "lu" else SPD_MAXBLOCKSIZE))
        self.order = Ordering(self.n)
        self.parent = None
        self.colcnt = None
        self.nnzl = 0
        self.nsuper = 0
        self.nsub = 0

    # SpkSparseBase.jl:173-185 (_findorder!)
    def findorder(self, orderfunction=None):
        if self.n == 0:
            raise RuntimeError("An empty problem, no ordering found.")
        self.g.makestructuresymmetric()
        if orderfunction is None:
            mmd(self.g, self.order)
        else:
            orderfunction(self.g, self.order)
        return True

    def findorderperm(self, perm):
        """Order with a caller's permutation (new -> old, 1-based). The
        reference's findorderperm! passes a Vector where a function is
        expected (SpkSparseSolver.jl:143-151 -> SpkSparseBase.jl:179); here
        the permutation is applied as intended."""
        if self.n == 0:
            raise RuntimeError("An empty problem, no ordering found.")
        perm = as_i64(perm)
        if sorted(perm.tolist()) != list(range(1, self.n + 1)):
            raise ValueError("perm is not a permutation of 1..n")
        self.g.makestructuresymmetric()
        self.order.rperm = perm.copy()
        self.order.rinvp = np.empty(self.n, dtype=np.int64)
        self.order.rinvp[perm - 1] = np.arange(1, self.n + 1, dtype=np.int64)
        return True

    # SpkSparseBase.jl:193-251 (_symbolicfactor!); SPD: SpkSparseSpdBase.jl:178-232
    def symbolicfactor(self):
        if self.n == 0:
            raise RuntimeError("An empty problem. No symbolic factorization done.")
        L = _lib.lib()
        n = self.n
        g = self.g
        o = self.order
        parent = np.zeros(n, dtype=np.int64)
        L.spk_etree(n, p64(g.xadj), p64(g.adj), p64(o.rperm), p64(o.rinvp), p64(parent))
        L.spk_postorder(n, p64(parent), p64(o.rperm), p64(o.rinvp), None)
        colcnt = np.zeros(n, dtype=np.int64)
        self.nnzl = int(L.spk_colcounts(n, p64(g.xadj), p64(g.adj), p64(o.rperm),
                                        p64(o.rinvp), p64(parent), p64(colcnt)))
        L.spk_postorder(n, p64(parent), p64(o.rperm), p64(o.rinvp), p64(colcnt))
        xsuper = np.zeros(n + 1, dtype=np.int64)
        snode = np.zeros(n, dtype=np.int64)
        nsub = np.zeros(1, dtype=np.int64)
        nsuper = int(L.spk_supernodes(n, p64(parent), p64(colcnt), self.maxblocksize,
                                      p64(xsuper), p64(snode), p64(nsub)))
        self.nsuper = nsuper
        self.nsub = int(nsub[0])
        self.xsuper = xsuper[: nsuper + 1].copy()
        self.snode = snode
        self.parent = parent
        self.colcnt = colcnt
        self.xlnz = np.zeros(n + 1, dtype=np.int64)
        if self.kind == "lu":
            self.xunz = np.zeros(n + 1, dtype=np.int64)
            L.spk_nonzero_indexs(n, p64(colcnt), nsuper, p64(self.xsuper),
                                 p64(self.xlnz), p64(self.xunz))
        else:
            self.xunz = None
            L.spk_nonzero_indexs_spd(n, p64(colcnt), nsuper, p64(self.xsuper),
                                     p64(self.xlnz))
        self.xlindx = np.zeros(nsuper + 1, dtype=np.int64)
        self.lindx = np.zeros(max(self.nsub, 1), dtype=np.int64)
        rc = L.spk_symbfact(n, p64(g.xadj), p64(g.adj), p64(o.rperm), p64(o.rinvp),
                            p64(colcnt), nsuper, p64(self.xsuper), p64(snode),
                            self.nsub, p64(self.xlindx), p64(self.lindx))
        if rc != 0:
            raise RuntimeError("Inconsistency in data structure.")
        self.lindx = self.lindx[: self.nsub]
        return True

    # sizes of the reference layout
    @property
    def nlnz(self):
        return int(self.xlnz[-1] - 1)

    @property
    def nunz(self):
        return int(self.xunz[-1] - 1) if self.xunz is not None else 0

    def supernode_parent(self):
        """_findsupernodetree! (SpkSymFct.jl:372-380); 0 marks a root."""
        L = _lib.lib()
        sp = np.zeros(self.nsuper, dtype=np.int64)
        L.spk_supernode_tree(self.nsuper, p64(self.xsuper), p64(self.snode),
                             p64(self.xlindx), p64(self.lindx), p64(sp))
        return sp

    def flops(self):
        """Algorithmic factor flops from the column counts: F_LU = sum_j
        (m_j + 2 m_j^2) with m_j = colcnt_j - 1, F_Chol = sum_j colcnt_j^2."""
        c = self.colcnt.astype(np.float64)
        m = c - 1.0
        if self.kind == "lu":
            return float(np.sum(m + 2.0 * m * m))
        return float(np.sum(c * c))
