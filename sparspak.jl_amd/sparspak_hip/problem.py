"""The reference's `Problem` input format (SpkProblem.jl:105-139).

The same storage as the reference: per column a linked list of its entries,
sorted by row (`head[c]` the first entry, `link[e]` the next, `rowsubs[e]`
the row, `values[e]` the value, entries in input order up to `lastused`),
entries ADDED on repeated input (`inaij!`, SpkProblem.jl:177-251), with a
right-hand side (`inbi!`, `infullrhs!`) and a solution vector `x`.
`linked_to_csc` walks the lists into the CSC arrays that the device upload
takes (the graph and values are those of the equivalent CSC matrix).
"""
import numpy as np
import scipy.sparse as sp

# the reference's marker of a slot without a value (SpkUtilities.jl:7)
BIGGY = float(np.finfo(np.float64).max)


def linked_to_csc(ncols, head, link, rowsubs, values):
    """CSC arrays (1-based colptr, rowval, nzval) of a Problem's linked lists:
    column c's entries in list order from head[c] (ascending rows, as
    inaij! keeps them). 0-based arrays of the reference's 1-based pointer
    values: head[c - 1], link[e - 1], rowsubs[e - 1], values[e - 1]."""
    colptr = np.ones(ncols + 1, dtype=np.int64)
    rows, vals = [], []
    for c in range(1, ncols + 1):
        e = int(head[c - 1]) if c - 1 < len(head) else 0
        cnt = 0
        while e > 0:
            rows.append(int(rowsubs[e - 1]))
            vals.append(float(values[e - 1]))
            cnt += 1
            e = int(link[e - 1])
        colptr[c] = colptr[c - 1] + cnt
    return colptr, np.asarray(rows, dtype=np.int64), np.asarray(vals, dtype=np.float64)


class Problem:
    def __init__(self, nrows=0, ncols=0, nnz=2500, z=0.0, info=""):
        self.info = info
        self.nrows = int(nrows)
        self.ncols = int(ncols)
        # linked lists (1-based pointer values, 0 ends a list), as Problem's
        # head / link / rowsubs / values / lastused (SpkProblem.jl:105-139)
        self.head = np.zeros(max(self.ncols, 1), dtype=np.int64)
        self.link = np.zeros(max(int(nnz), 1), dtype=np.int64)
        self.rowsubs = np.zeros(max(int(nnz), 1), dtype=np.int64)
        self.values = np.full(max(int(nnz), 1), BIGGY)
        self.lastused = 0
        self.nnz = 0         # number of stored values
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
        if self.head.shape[0] < self.ncols:
            self.head = np.concatenate(
                [self.head, np.zeros(2 * self.ncols - self.head.shape[0], dtype=np.int64)])
        if self.lastused >= self.link.shape[0]:
            grow = max(2 * self.link.shape[0], 3 * self.ncols) - self.link.shape[0]
            self.link = np.concatenate([self.link, np.zeros(grow, dtype=np.int64)])
            self.rowsubs = np.concatenate([self.rowsubs, np.zeros(grow, dtype=np.int64)])
            self.values = np.concatenate([self.values, np.full(grow, BIGGY)])

    def inaij(self, rnum, cnum, aij=0.0):
        """inaij! (SpkProblem.jl:177-251): add aij to entry (rnum, cnum) of
        column cnum's row-sorted list (a new entry is linked in at its row's
        place); invalid subscripts are ignored with a warning (returns
        False)."""
        if rnum < 1 or cnum < 1:
            import warnings
            warnings.warn(f"invalid matrix subscripts {rnum}, {cnum}: input ignored")
            return False
        rnum, cnum = int(rnum), int(cnum)
        self.nrows = max(self.nrows, rnum)
        self.ncols = max(self.ncols, cnum)
        self._grow()
        ptr = int(self.head[cnum - 1])
        lastptr = 0
        while ptr > 0:
            r = int(self.rowsubs[ptr - 1])
            if r > rnum:
                break
            if r == rnum:
                if self.values[ptr - 1] == BIGGY:
                    self.nnz += 1
                    self.values[ptr - 1] = aij
                    if rnum == cnum:
                        self.dnz += 1
                else:
                    self.values[ptr - 1] += aij
                return True
            lastptr = ptr
            ptr = int(self.link[ptr - 1])
        self.lastused += 1
        e = self.lastused
        self.nedges += 1
        self.rowsubs[e - 1] = rnum
        if rnum == cnum:
            self.dedges += 1
        if lastptr == 0:          # first entry of the list
            self.link[e - 1] = self.head[cnum - 1]
            self.head[cnum - 1] = e
        elif ptr == 0:            # at the end of the list
            self.link[lastptr - 1] = e
            self.link[e - 1] = 0
        else:                     # between lastptr and ptr
            self.link[lastptr - 1] = e
            self.link[e - 1] = ptr
        self.values[e - 1] = aij
        self.nnz += 1
        if rnum == cnum:
            self.dnz += 1
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
        """(n, colptr, rowval, nzval): 1-based CSC with ascending rows from
        the linked lists (linked_to_csc); explicit zeros kept (they are part
        of the structure)."""
        n = self.ncols
        colptr, rowval, nzval = linked_to_csc(n, self.head, self.link, self.rowsubs, self.values)
        return n, colptr, rowval, nzval

    def outsparse(self):
        """outsparse (SpkProblem.jl:305-326) as a scipy CSC matrix."""
        n, colptr, rowval, nzval = self.csc()
        return sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=(self.nrows, self.ncols))
