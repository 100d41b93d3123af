"""The reference's `Problem` input format (SpkProblem.jl:105-139).

A column-oriented store whose entries are ADDED on repeated input
(`inaij!`, SpkProblem.jl:177-251), with a right-hand side (`inbi!`,
`infullrhs!`) and a solution vector `x`. Each column keeps its row indices in
ascending order, as the reference's sorted linked lists do, so the graph and
the assembled values are the same as from the equivalent CSC matrix.
"""
import numpy as np
import scipy.sparse as sp


class Problem:
    def __init__(self, nrows=0, ncols=0, nnz=2500, z=0.0, info=""):
        self.info = info
        self.nrows = int(nrows)
        self.ncols = int(ncols)
        self._cols = {}      # column -> {row: value}
        self.nnz = 0         # number of stored entries (values)
        self.dnz = 0
        self.nedges = 0
        self.dedges = 0
        self.rhs = np.zeros(self.nrows)
        self.x = np.zeros(self.ncols)

    def _grow(self):
        if self.rhs.shape[0] < self.nrows:
            self.rhs = np.concatenate([self.rhs, np.zeros(self.nrows - self.rhs.shape[0])])
        if self.x.shape[0] < self.ncols:
            self.x = np.concatenate([self.x, np.zeros(self.ncols - self.x.shape[0])])

    def inaij(self, rnum, cnum, aij=0.0):
        """inaij! (SpkProblem.jl:177-251): add aij to entry (rnum, cnum);
        invalid subscripts are ignored with a warning (returns False)."""
        if rnum < 1 or cnum < 1:
            import warnings
            warnings.warn(f"invalid matrix subscripts {rnum}, {cnum}: input ignored")
            return False
        self.nrows = max(self.nrows, int(rnum))
        self.ncols = max(self.ncols, int(cnum))
        self._grow()
        col = self._cols.setdefault(int(cnum), {})
        if int(rnum) in col:
            col[int(rnum)] += aij
        else:
            col[int(rnum)] = aij
            self.nnz += 1
            self.nedges += 1
            if rnum == cnum:
                self.dnz += 1
                self.dedges += 1
        return True

    def inbi(self, rnum, bi):
        """inbi! (SpkProblem.jl:258-271)."""
        if rnum < 1:
            raise ValueError(f"Invalid rhs subscript {rnum}.")
        self.nrows = max(self.nrows, int(rnum))
        self._grow()
        self.rhs[int(rnum) - 1] += bi
        return True

    def insparse(self, spm_or_I, J=None, V=None):
        """insparse! (SpkProblem.jl:280-298) from a scipy matrix or COO
        triplets (1-based)."""
        if J is None:
            m = sp.coo_matrix(spm_or_I)
            I, J, V = m.row + 1, m.col + 1, m.data
        else:
            I = spm_or_I
        for i, j, v in zip(I, J, V):
            if not self.inaij(int(i), int(j), float(v)):
                return False
        return True

    def infullrhs(self, rhs):
        """infullrhs! (SpkProblem.jl:512-517)."""
        for i in range(self.nrows, 0, -1):
            self.inbi(i, float(rhs[i - 1]))
        return self

    def csc(self):
        """(n, colptr, rowval, nzval): 1-based CSC with ascending rows,
        explicit zeros kept (they are part of the structure)."""
        n = self.ncols
        colptr = np.ones(n + 1, dtype=np.int64)
        rows, vals = [], []
        for j in range(1, n + 1):
            col = self._cols.get(j, {})
            rs = sorted(col)
            rows.extend(rs)
            vals.extend(col[r] for r in rs)
            colptr[j] = colptr[j - 1] + len(rs)
        return (n, colptr, np.asarray(rows, dtype=np.int64),
                np.asarray(vals, dtype=np.float64))

    def outsparse(self):
        """outsparse (SpkProblem.jl:305-326) as a scipy CSC matrix."""
        n, colptr, rowval, nzval = self.csc()
        return sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=(self.nrows, self.ncols))
