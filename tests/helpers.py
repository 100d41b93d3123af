"""Shared helpers for the tests: symbolic set-up from scipy matrices,
brute-force symbolic elimination, a dense textbook LU with pivoting confined
to diagonal blocks, and error measures."""
import numpy as np
import scipy.sparse as sp

from sparspak_hip.symbolic import Graph, SymbolicFactor


def csc1(A):
    """1-based canonical CSC arrays of a scipy matrix."""
    A = sp.csc_matrix(A)
    A.sum_duplicates()
    A.sort_indices()
    return (A.shape[0], A.indptr.astype(np.int64) + 1, A.indices.astype(np.int64) + 1,
            A.data.astype(np.float64))


def symbolic(A, kind="lu", maxblocksize=None):
    n, cp, rv, _ = csc1(A)
    g = Graph.from_csc(n, cp, rv)
    s = SymbolicFactor(n, g, kind=kind, maxblocksize=maxblocksize)
    s.findorder()
    s.symbolicfactor()
    return s


def random_pattern(n, density, seed, symmetric=True, dominant=True):
    rng = np.random.default_rng(seed)
    M = (rng.random((n, n)) < density).astype(float)
    if symmetric:
        M = np.maximum(M, M.T)
    vals = rng.uniform(-1, 1, (n, n)) * M
    np.fill_diagonal(vals, 0.0)
    d = np.abs(vals).sum(axis=1) + (1.0 if dominant else 0.0)
    vals[np.arange(n), np.arange(n)] = d * rng.choice([-1, 1], n) if dominant else rng.uniform(-1, 1, n)
    return sp.csc_matrix(vals)


def brute_symbolic(A, perm):
    """Boolean elimination of P (A + A^T) P^T: list of row sets of L's
    columns (0-based new numbering), etree parent (0-based, -1 root)."""
    n = A.shape[0]
    B = (abs(A) + abs(A).T).toarray() != 0
    p = perm - 1
    B = B[np.ix_(p, p)]
    np.fill_diagonal(B, True)
    S = B.copy()
    cols = []
    for j in range(n):
        rows = set(np.nonzero(S[j:, j])[0] + j)
        cols.append(rows)
        below = sorted(r for r in rows if r > j)
        for a in below:
            for b in below:
                S[a, b] = True
    parent = [min([r for r in cols[j] if r > j], default=-1) for j in range(n)]
    return cols, parent


def blocked_lu_reference(Ap, xsuper):
    """Dense right-looking LU of the permuted matrix Ap with row pivoting
    confined to each supernode's diagonal block: a row interchange inside
    block [f, l] spans columns f..n-1 only. Returns the packed LU and the
    local pivots (1-based within the block)."""
    A = np.array(Ap, dtype=np.float64, copy=True)
    n = A.shape[0]
    piv = np.zeros(n, dtype=np.int64)
    for J in range(len(xsuper) - 1):
        f, l = xsuper[J] - 1, xsuper[J + 1] - 1
        for k in range(f, l):
            kp = k + int(np.argmax(np.abs(A[k:l, k])))  # first max
            piv[k] = kp - f + 1
            if kp != k and A[kp, k] != 0:
                A[[k, kp], f:] = A[[kp, k], f:]
            if A[k, k] != 0:
                A[k + 1:, k] *= 1.0 / A[k, k]
            A[k + 1:, k + 1:] -= np.outer(A[k + 1:, k], A[k, k + 1:])
    return A, piv


def panels_to_dense(s, lnz, unz):
    """Scatter the reference layout (lnz/unz panels) into a dense packed LU
    (new numbering); entries outside the structure are 0."""
    n = s.n
    D = np.zeros((n, n))
    for J in range(s.nsuper):
        f = s.xsuper[J] - 1
        w = s.xsuper[J + 1] - s.xsuper[J]
        I = s.lindx[s.xlindx[J] - 1: s.xlindx[J + 1] - 1] - 1
        jl = len(I)
        P = lnz[s.xlnz[f] - 1: s.xlnz[f] - 1 + jl * w].reshape(w, jl).T
        D[np.ix_(I, np.arange(f, f + w))] = P
        if unz is not None and jl > w:
            U = unz[s.xunz[f] - 1: s.xunz[f] - 1 + (jl - w) * w].reshape(w, jl - w)
            D[np.ix_(np.arange(f, f + w), I[w:])] = U
    return D


def backward_error(A, x, b):
    r = b - A @ x
    an = abs(A).sum(axis=1).max()
    return np.max(np.abs(r), axis=0) / (an * np.max(np.abs(x), axis=0) + np.max(np.abs(b), axis=0))


def spd_meaningful(sym):
    """Mask of the LDL^T panel entries the factor defines: the diagonal
    block's diagonal and lower triangle and the rows below it. The strict
    upper triangle of a diagonal block is never read by _ldltsolve!; the
    reference leaves partial update sums there, this library zeros."""
    m = np.ones(sym.nlnz, dtype=bool)
    for J in range(sym.nsuper):
        f = sym.xsuper[J] - 1
        w = sym.xsuper[J + 1] - sym.xsuper[J]
        jl = sym.xlindx[J + 1] - sym.xlindx[J]
        base = sym.xlnz[f] - 1
        for c in range(w):
            m[base + c * jl: base + c * jl + c] = False
    return m
