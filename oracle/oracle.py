"""ORACLE — test infrastructure only (tests/, the smoke check and bench.py's
cpu_baseline leg). ctypes access to liboracle.so, the CPU restatement of
Sparspak.jl's numeric phase (see spk_oracle.cpp for the function-by-function
citations). Never used by the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

i64 = ctypes.c_int64
i64p = ctypes.POINTER(ctypes.c_int64)
dp = ctypes.POINTER(ctypes.c_double)
fp = ctypes.POINTER(ctypes.c_float)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        for suf, P in (("f64", dp), ("f32", fp)):
            getattr(L, f"orc_lufactor_{suf}").argtypes = [i64, i64, i64p, i64p, i64p, i64p, i64p,
                                                          P, i64p, P, i64p, i64p]
            getattr(L, f"orc_lulsolve_{suf}").argtypes = [i64, i64p, i64p, i64p, i64p, P, i64p, P]
            getattr(L, f"orc_luusolve_{suf}").argtypes = [i64, i64, i64p, i64p, i64p, i64p, P,
                                                          i64p, P, P]
            getattr(L, f"orc_inmatrix_csc_{suf}").argtypes = [i64, i64p, i64p, dp, i64p, i64p,
                                                              i64p, i64p, i64p, i64p, P, i64p, P,
                                                              i64p]
            getattr(L, f"orc_inmatrix_spd_csc_{suf}").argtypes = [i64, i64p, i64p, dp, i64p, i64p,
                                                                  i64p, i64p, i64p, i64p, P]
            getattr(L, f"orc_ldltfactor_{suf}").argtypes = [i64, i64, i64p, i64p, i64p, i64p,
                                                            i64p, P]
            getattr(L, f"orc_ldltfactor_ref_{suf}").argtypes = [i64, i64, i64p, i64p, i64p, i64p,
                                                                i64p, P]
            getattr(L, f"orc_ldltsolve_{suf}").argtypes = [i64, i64p, i64p, i64p, i64p, P, P]
            for name in ("lufactor", "lulsolve", "luusolve", "inmatrix_csc", "inmatrix_spd_csc",
                         "ldltfactor", "ldltfactor_ref", "ldltsolve"):
                getattr(L, f"orc_{name}_{suf}").restype = i64
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(i64p)


def _f(a):
    return a.ctypes.data_as(dp if a.dtype == np.float64 else fp)


def _suf(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def inmatrix(sym, colptr, rowval, nzval, dtype=np.float64):
    """_inmatrix! on CSC input -> (lnz, unz, ipiv)."""
    L = lib()
    n = sym.n
    lnz = np.zeros(sym.nlnz, dtype=dtype)
    nzval = np.ascontiguousarray(nzval, dtype=np.float64)
    colptr = np.ascontiguousarray(colptr, dtype=np.int64)
    rowval = np.ascontiguousarray(rowval, dtype=np.int64)
    if sym.kind == "lu":
        unz = np.zeros(max(sym.nunz, 1), dtype=dtype)
        ipiv = np.zeros(n, dtype=np.int64)
        rc = getattr(L, f"orc_inmatrix_csc_{_suf(dtype)}")(
            n, _p(colptr), _p(rowval), nzval.ctypes.data_as(dp), _p(sym.order.rinvp),
            _p(sym.snode), _p(sym.xsuper), _p(sym.xlindx), _p(sym.lindx), _p(sym.xlnz),
            _f(lnz), _p(sym.xunz), _f(unz), _p(ipiv))
        if rc != 0:
            raise RuntimeError("No space for matrix element.")
        return lnz, unz[: sym.nunz], ipiv
    rc = getattr(L, f"orc_inmatrix_spd_csc_{_suf(dtype)}")(
        n, _p(colptr), _p(rowval), nzval.ctypes.data_as(dp), _p(sym.order.rinvp),
        _p(sym.snode), _p(sym.xsuper), _p(sym.xlindx), _p(sym.lindx), _p(sym.xlnz), _f(lnz))
    if rc != 0:
        raise RuntimeError("No space for matrix element.")
    return lnz, None, None


def lufactor(sym, lnz, unz, ipiv):
    """_lufactor! in place -> (iflag, first failing supernode)."""
    L = lib()
    ff = np.zeros(1, dtype=np.int64)
    u = unz if unz.size else np.zeros(1, dtype=lnz.dtype)
    rc = getattr(L, f"orc_lufactor_{_suf(lnz.dtype)}")(
        sym.n, sym.nsuper, _p(sym.xsuper), _p(sym.snode), _p(sym.xlindx), _p(sym.lindx),
        _p(sym.xlnz), _f(lnz), _p(sym.xunz), _f(u), _p(ipiv), _p(ff))
    if unz.size:
        unz[:] = u[: unz.size]
    return int(rc), int(ff[0])


def ldltfactor(sym, lnz, faithful=False):
    L = lib()
    name = "ldltfactor_ref" if faithful else "ldltfactor"
    return int(getattr(L, f"orc_{name}_{_suf(lnz.dtype)}")(
        sym.n, sym.nsuper, _p(sym.xsuper), _p(sym.snode), _p(sym.xlindx), _p(sym.lindx),
        _p(sym.xlnz), _f(lnz)))


def lusolve_permuted(sym, lnz, unz, ipiv, rhs):
    """_lulsolve! then _luusolve! on an already permuted rhs (in place)."""
    L = lib()
    s = _suf(lnz.dtype)
    u = unz if unz.size else np.zeros(1, dtype=lnz.dtype)
    getattr(L, f"orc_lulsolve_{s}")(sym.nsuper, _p(sym.xsuper), _p(sym.xlindx), _p(sym.lindx),
                                     _p(sym.xlnz), _f(lnz), _p(ipiv), _f(rhs))
    getattr(L, f"orc_luusolve_{s}")(sym.n, sym.nsuper, _p(sym.xsuper), _p(sym.xlindx),
                                     _p(sym.lindx), _p(sym.xlnz), _f(lnz), _p(sym.xunz), _f(u),
                                     _f(rhs))
    return rhs


def lulsolve(sym, lnz, ipiv, rhs):
    L = lib()
    getattr(L, f"orc_lulsolve_{_suf(lnz.dtype)}")(sym.nsuper, _p(sym.xsuper), _p(sym.xlindx),
                                                   _p(sym.lindx), _p(sym.xlnz), _f(lnz),
                                                   _p(ipiv), _f(rhs))
    return rhs


def luusolve(sym, lnz, unz, rhs):
    L = lib()
    u = unz if unz.size else np.zeros(1, dtype=lnz.dtype)
    getattr(L, f"orc_luusolve_{_suf(lnz.dtype)}")(sym.n, sym.nsuper, _p(sym.xsuper),
                                                   _p(sym.xlindx), _p(sym.lindx), _p(sym.xlnz),
                                                   _f(lnz), _p(sym.xunz), _f(u), _f(rhs))
    return rhs


def ldltsolve_permuted(sym, lnz, rhs):
    L = lib()
    getattr(L, f"orc_ldltsolve_{_suf(lnz.dtype)}")(sym.nsuper, _p(sym.xsuper), _p(sym.xlindx),
                                                    _p(sym.lindx), _p(sym.xlnz), _f(lnz),
                                                    _f(rhs))
    return rhs


def solve(sym, lnz, unz, ipiv, b):
    """_triangularsolve!: rhs = b[rperm]; solves; x = rhs[rinvp]."""
    o = sym.order
    rhs = np.ascontiguousarray(b[o.rperm - 1], dtype=lnz.dtype)
    if sym.kind == "lu":
        lusolve_permuted(sym, lnz, unz, ipiv, rhs)
    else:
        ldltsolve_permuted(sym, lnz, rhs)
    return rhs[o.rinvp - 1]


def factor_csc(sym, colptr, rowval, nzval, dtype=np.float64):
    """inmatrix + factor on the CPU -> (lnz, unz, ipiv, iflag, first_fail)."""
    lnz, unz, ipiv = inmatrix(sym, colptr, rowval, nzval, dtype)
    if sym.kind == "lu":
        rc, ff = lufactor(sym, lnz, unz, ipiv)
        return lnz, unz, ipiv, rc, ff
    rc = ldltfactor(sym, lnz)
    return lnz, None, None, rc, 0


def ainverse_onenorm(solve, n, itmax=5):
    """Hager's 1-norm estimate of A^-1 for a symmetric A, restated from the
    reference's commented sparsespdainverseonenorm
    (SpkSparseSpdSolver.jl:346-431): solve(x) returns A^-1 x (which is also
    A^-T x). Plain numpy; test infrastructure."""
    def sgn(v):
        return np.where(v < 0.0, -1.0, 1.0)
    x = solve(np.full(n, 1.0 / n))
    if n == 1:
        return abs(x[0])
    est = float(np.sum(np.abs(x)))
    x = sgn(x)
    isgn = x.astype(np.int64)
    x = solve(x)
    j = int(np.argmax(np.abs(x)))
    it = 1
    while True:
        it += 1
        x = np.zeros(n)
        x[j] = 1.0
        x = solve(x)
        estold, est = est, float(np.sum(np.abs(x)))
        if np.all(sgn(x).astype(np.int64) == isgn):
            break
        if est <= estold:
            break
        x = sgn(x)
        isgn = x.astype(np.int64)
        x = solve(x)
        jlast, j = j, int(np.argmax(np.abs(x)))
        if x[jlast] == abs(x[j]) or it >= itmax:
            break
    i = np.arange(n, dtype=np.float64)
    x = np.where(np.arange(n) % 2 == 0, 1.0, -1.0) * (1.0 + i / (n - 1))
    x = solve(x)
    temp = 2.0 * float(np.sum(np.abs(x))) / (3.0 * n)
    return max(temp, est)
