"""SparseSolver / SparseSpdSolver with the reference's staged API.

Mirrors SpkSparseSolver.jl:17-275 (LU), SpkSparseSpdSolver.jl:21-265 (SPD)
and SparseCSCInterface.jl:171-300 (CSC input, `sparspaklu`, `sparspaklu!`,
`ldiv!`, `\\`): the same phase functions, the same phase flags and
invalidation rules, the same sequence errors. Julia's `f!` names appear
here without the bang (`findorder`, `symbolicfactor`, `inmatrix`, `factor`,
`triangularsolve`, `solve`); `sparspaklu!` is `sparspaklu_refactor`.

The symbolic phase runs on the host (library's C++ restatement), the
numeric phase (`inmatrix`, `factor`, `triangularsolve`) on the GPU through
libsparspak_hip.so. There is no CPU fallback.
"""
import ctypes

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import as_i64, check, p64, pd, pf
from .problem import Problem
from .symbolic import Graph, SymbolicFactor

FP64 = 0
FP32 = 1
FP32_BF16 = 2  # float factor, bf16 operands in the trailing updates of large fronts
FP32_FP16 = 3  # the same with diagonally scaled fp16 operands
FP32_BF16X3 = 4  # the same with split-bf16 operands (three bf16 products per pair)
_PREC = {"fp64": FP64, "float64": FP64, FP64: FP64,
         "fp32": FP32, "float32": FP32, FP32: FP32,
         "fp32_bf16": FP32_BF16, FP32_BF16: FP32_BF16,
         "fp32_fp16": FP32_FP16, FP32_FP16: FP32_FP16,
         "fp32_bf16x3": FP32_BF16X3, FP32_BF16X3: FP32_BF16X3}


def _csc_arrays(m):
    """Canonical 1-based CSC arrays (sorted rows, duplicates summed, explicit
    zeros kept) of a scipy sparse matrix."""
    m = sp.csc_matrix(m)
    m.sum_duplicates()
    m.sort_indices()
    n = m.shape[1]
    return (n, m.indptr.astype(np.int64) + 1, m.indices.astype(np.int64) + 1,
            m.data.astype(np.float64))


class _Device:
    """A device handle (spkhip_t) owned by one solver."""

    def __init__(self, device=0):
        L = _lib.lib()
        h = ctypes.c_void_p()
        check(L.spkhip_create(ctypes.byref(h), int(device)), "spkhip_create")
        self.h = h
        self.L = L

    def close(self):
        if self.h:
            self.L.spkhip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _SparseBase:
    """Storage + phase bodies (SpkSparseBase.jl:99-416 / SpkSparseSpdBase.jl)."""

    def __init__(self, n, colptr, rowval, kind="lu", device=0):
        self.n = int(n)
        self.kind = kind
        self.colptr = as_i64(colptr)
        self.rowval = as_i64(rowval)
        self.nnz = int(self.colptr[-1] - 1) if self.n else 0
        g = Graph.from_csc(self.n, self.colptr, self.rowval) if self.n else Graph(0, [1], [])
        self.sym = SymbolicFactor(self.n, g, kind=kind)
        self.device = device
        self._dev = None
        self._uploaded = False
        self.errflag = 0
        self.precision = FP64

    @property
    def order(self):
        return self.sym.order

    def _findorder(self, orderfunction=None):
        if self.n == 0:
            raise RuntimeError("An empty problem, no ordering found.")
        self._uploaded = False
        return self.sym.findorder(orderfunction)

    def _findorderperm(self, perm):
        self._uploaded = False
        return self.sym.findorderperm(perm)

    def _symbolicfactor(self):
        if self.n == 0:
            raise RuntimeError("An empty problem. No symbolic factorization done.")
        self._uploaded = False
        return self.sym.symbolicfactor()

    def _upload(self):
        if self._uploaded:
            return
        if self._dev is None:
            self._dev = _Device(self.device)
        s = self.sym
        kind = 0 if self.kind == "lu" else 1
        xunz = s.xunz if s.xunz is not None else np.zeros(self.n + 1, dtype=np.int64)
        rc = self._dev.L.spkhip_symbolic_upload(
            self._dev.h, kind, self.n, s.nsuper, p64(s.xsuper), p64(s.snode),
            p64(s.xlindx), p64(s.lindx), p64(s.xlnz), p64(xunz),
            p64(s.order.rperm), p64(s.order.rinvp), p64(self.colptr), p64(self.rowval))
        if rc == -1:
            raise RuntimeError("No space for matrix element.")
        check(rc, "spkhip_symbolic_upload")
        self._uploaded = True

    def _inmatrix(self, nzval):
        """_inmatrix! (SpkSparseBase.jl:302-372, SparseCSCInterface.jl:101-169)."""
        if self.n == 0:
            raise RuntimeError("An empty problem. No matrix.")
        self._upload()
        nz = np.ascontiguousarray(nzval, dtype=np.float64)
        if nz.shape[0] != self.nnz:
            raise ValueError("value count does not match the pattern")
        rc = self._dev.L.spkhip_inmatrix(self._dev.h, pd(nz))
        if rc == -1:
            raise RuntimeError("No space for matrix element.")
        check(rc, "spkhip_inmatrix")
        return True

    def _factor(self, precision=FP64):
        """_factor! (SpkSparseBase.jl:378-391). A zero pivot in a diagonal
        block raises, as the reference does through its error flag."""
        if self.n == 0:
            raise RuntimeError("An empty problem. No matrix.")
        self.precision = _PREC[precision]
        rc = int(self._dev.L.spkhip_factor(self._dev.h, self.precision))
        self.errflag = rc
        if rc == -1:
            st = self.stats()
            raise RuntimeError(
                "Numerical factorization: zero pivot in the diagonal block of "
                f"supernode {st.first_fail_supernode}.")
        check(rc, "spkhip_factor")
        return True

    def _triangularsolve(self, rhs, refine=False, tol=1e-14, max_iter=10):
        """_triangularsolve! (SpkSparseBase.jl:400-416) on a vector or an
        n x k array (columns solved together); in place."""
        if self.n == 0:
            raise RuntimeError("An empty problem. No solution.")
        b = np.asarray(rhs)
        if b.dtype != np.float64:
            raise TypeError("right-hand side must be float64")
        vec = b.ndim == 1
        B = np.asfortranarray(b.reshape(self.n, -1))
        nrhs = B.shape[1]
        if refine:
            rc = self._dev.L.spkhip_solve_refine(self._dev.h, nrhs, pd(B), self.n,
                                                 float(tol), int(max_iter))
            self.refine_rc = int(rc)
            _lib.check_refine(rc, "spkhip_solve_refine",
                              self.stats().backward_error if rc == -5 else 0.0)
        else:
            check(self._dev.L.spkhip_solve(self._dev.h, nrhs, pd(B), self.n), "spkhip_solve")
        if vec:
            rhs[:] = B[:, 0]
        else:
            rhs[...] = B
        return True

    def get_tuning(self):
        self._upload()
        t = _lib.Tuning()
        check(self._dev.L.spkhip_get_tuning(self._dev.h, ctypes.byref(t)), "spkhip_get_tuning")
        return {f: getattr(t, f) for f in _lib.TUNING_KNOBS + ("factor_ms",)}

    def set_tuning(self, **knobs):
        """Plan knobs (small_lmax, block_cols, upd_group, small_wave_min);
        the factor must be recomputed afterwards."""
        cur = self.get_tuning()
        bad = set(knobs) - set(_lib.TUNING_KNOBS)
        if bad:
            raise ValueError(f"unknown tuning knobs {sorted(bad)}")
        cur.update(knobs)
        t = _lib.Tuning(*(int(cur[f]) for f in _lib.TUNING_KNOBS), 0.0)
        check(self._dev.L.spkhip_set_tuning(self._dev.h, ctypes.byref(t)), "spkhip_set_tuning")

    def _autotune(self, precision=FP32, reps=2):
        t = _lib.Tuning()
        rc = int(self._dev.L.spkhip_autotune(self._dev.h, _PREC[precision], int(reps),
                                             ctypes.byref(t)))
        self.errflag = rc
        if rc == -1:
            raise RuntimeError("Numerical factorization: zero pivot during autotuning.")
        check(rc, "spkhip_autotune")
        self.precision = _PREC[precision]
        return {f: getattr(t, f) for f in _lib.TUNING_KNOBS + ("factor_ms",)}

    def _condest(self):
        c, a = ctypes.c_double(0.0), ctypes.c_double(0.0)
        check(self._dev.L.spkhip_condest(self._dev.h, ctypes.byref(c), ctypes.byref(a)),
              "spkhip_condest")
        return c.value, a.value

    def stats(self):
        st = _lib.Stats()
        check(self._dev.L.spkhip_stats(self._dev.h, ctypes.byref(st)), "spkhip_stats")
        return st

    def download(self):
        """(lnz, unz, ipiv) of the current factor in the reference layout."""
        s = self.sym
        lnz = np.zeros(s.nlnz)
        unz = np.zeros(max(s.nunz, 0))
        ipiv = np.zeros(self.n, dtype=np.int64)
        check(self._dev.L.spkhip_download(self._dev.h, pd(lnz), pd(unz) if unz.size else None,
                                          p64(ipiv)), "spkhip_download")
        return lnz, unz, ipiv

    @property
    def lnz(self):
        return self.download()[0]

    @property
    def unz(self):
        return self.download()[1]

    @property
    def ipiv(self):
        return self.download()[2]

    def __getattr__(self, name):
        # symbolic arrays (xlnz, xunz, xlindx, lindx, xsuper, snode, colcnt,
        # nsuper, ...) are read from the symbolic factor
        if name in ("sym",):
            raise AttributeError(name)
        return getattr(self.sym, name)


class _SolverBase:
    kind = "lu"

    def __init__(self, p, device=0):
        if isinstance(p, Problem):
            n, colptr, rowval, nzval = p.csc()
            self.ma, self.na = p.nrows, p.ncols
        else:
            n, colptr, rowval, nzval = _csc_arrays(p)
            self.ma, self.na = p.shape[1], p.shape[0]
        self.p = p
        self.n = n
        self.mc = 0
        self.nc = 0
        self._nzval = nzval
        self._colptr = colptr
        self._rowval = rowval
        self.slvr = _SparseBase(n, colptr, rowval, kind=self.kind, device=device)
        self._orderdone = False
        self._symbolicdone = False
        self._inmatrixdone = False
        self._factordone = False
        self._trisolvedone = False
        self._refinedone = False
        self._condestdone = False
        self.precision = FP64


class SparseSolver(_SolverBase):
    """SparseSolver (SpkSparseSolver.jl:17-63, SparseCSCInterface.jl:171-186)."""
    kind = "lu"


class SparseSpdSolver(_SolverBase):
    """SparseSpdSolver (SpkSparseSpdSolver.jl:21-265): LDL^T layout, lower
    triangle of the input used, supernodes split at 60 columns."""
    kind = "spd"


# -- phase functions (SpkSparseSolver.jl:80-275) ----------------------------

def solve(s, rhs=None, **kw):
    """solve! (SpkSparseSolver.jl:80-87; with rhs: SparseCSCInterface.jl:194-201)."""
    findorder(s) or _err("Finding Order.")
    symbolicfactor(s) or _err("Symbolic Factorization.")
    inmatrix(s) or _err("Matrix input.")
    factor(s, **{k: v for k, v in kw.items() if k == "precision"}) or _err("Numerical Factorization.")
    kw.pop("precision", None)
    if rhs is None:
        triangularsolve(s, **kw) or _err("Triangular Solve.")
    else:
        triangularsolve(s, rhs, **kw) or _err("Triangular Solve.")
    return True


def _err(msg):
    raise RuntimeError(msg)


def findorder(s, orderfunction=None):
    """findorder! (SpkSparseSolver.jl:103-131)."""
    if s._orderdone:
        return True
    s.slvr._findorder(orderfunction)
    s._orderdone = True
    s._symbolicdone = False
    return True


def findorderperm(s, perm):
    """findorderperm! (SpkSparseSolver.jl:143-151), with the permutation
    actually applied (the reference passes it where a function is expected)."""
    if s._orderdone:
        return True
    s.slvr._findorderperm(perm)
    s._orderdone = True
    s._symbolicdone = False
    return True


def symbolicfactor(s):
    """symbolicfactor! (SpkSparseSolver.jl:163-175)."""
    if s._symbolicdone:
        return True
    if not s._orderdone:
        raise RuntimeError("Sequence error. Ordering not done yet.")
    s.slvr._symbolicfactor()
    s._symbolicdone = True
    s._inmatrixdone = False
    return True


def inmatrix(s):
    """inmatrix! (SpkSparseSolver.jl:188-200)."""
    if s._inmatrixdone:
        return True
    if not s._symbolicdone:
        raise RuntimeError("Sequence error. Symbolic factor not done yet.")
    ok = s.slvr._inmatrix(s._nzval)
    s._inmatrixdone = True
    s._factordone = False
    return ok


def factor(s, precision=FP64):
    """factor! (SpkSparseSolver.jl:209-227). precision: "fp64" (the
    reference's Float64 path) or "fp32" (float factor, refine in double)."""
    if s._factordone and _PREC[precision] == s.precision:
        return True
    if not s._inmatrixdone:
        raise RuntimeError("Sequence error. Matrix input not done yet.")
    s._trisolvedone = False
    s.precision = _PREC[precision]
    s.slvr._factor(precision)
    if s.slvr.errflag == 0:
        s._factordone = True
        return True
    return False


def autotune(s, precision=FP32, reps=2):
    """Choose the plan knobs (the small-versus-MFMA front split, the block
    width, the update depth, the small-front wave threshold) by timing the
    factorisation of the current values at `precision`; leaves `s`
    factored with the fastest setting and returns it (with its factor_ms).
    No counterpart in the reference."""
    if not s._inmatrixdone:
        raise RuntimeError("Sequence error. Matrix input not done yet.")
    s._trisolvedone = False
    best = s.slvr._autotune(precision, reps)
    s.precision = _PREC[precision]
    s._factordone = True
    return best


def condest(s):
    """sparsespdcondest / sparsespdainverseonenorm (SpkSparseSpdSolver.jl:346-459,
    commented Fortran in the reference): for a factored SparseSpdSolver, the
    1-norm condition number estimate ||A||_1 est(||A^-1||_1) with est from
    Hager's method (at most five iterations) on device solves. Returns
    (estimate of cond_1(A), estimate of ||A^-1||_1, a lower bound)."""
    if s.kind != "spd":
        raise ValueError("condest needs a SparseSpdSolver (its solves with A^T are "
                         "the solves with A)")
    if not s._factordone:
        raise RuntimeError("Sequence error. Factorization not done yet.")
    cond, ainv = s.slvr._condest()
    s._condestdone = True
    return cond, ainv


def triangularsolve(s, rhs=None, refine=None, tol=1e-14, max_iter=10):
    """triangularsolve! (SpkSparseSolver.jl:237-275). Without rhs the
    Problem's rhs is solved into p.x; with rhs (vector or n x k array) the
    solve is in place and always runs. refine defaults to True for a float
    factor (double-precision iterative refinement)."""
    if refine is None:
        refine = s.precision != FP64
    if rhs is None:
        if s._trisolvedone:
            return True
        if not s._factordone:
            raise RuntimeError("Sequence error. Factorization not done yet.")
        if not isinstance(s.p, Problem):
            raise RuntimeError("no right-hand side stored in the problem")
        temp = np.array(s.p.rhs[: s.p.nrows], dtype=np.float64)
        s.slvr._triangularsolve(temp, refine=refine, tol=tol, max_iter=max_iter)
        s.p.x[: temp.shape[0]] = temp
        s._trisolvedone = True
        s._refinedone = bool(refine)
        return True
    if not s._factordone:
        raise RuntimeError("Sequence error. Factorization not done yet.")
    s.slvr._triangularsolve(rhs, refine=refine, tol=tol, max_iter=max_iter)
    s._trisolvedone = True
    s._refinedone = bool(refine)
    return True


# -- CSC interface (SparseCSCInterface.jl:218-300) ---------------------------

def sparspaklu(m, factorize=True, precision=FP64, device=0):
    """sparspaklu (SparseCSCInterface.jl:218-227)."""
    lu = SparseSolver(m, device=device)
    if factorize:
        findorder(lu) or _err("Finding Order.")
        symbolicfactor(lu) or _err("Symbolic Factorization.")
        inmatrix(lu) or _err("Matrix input.")
        factor(lu, precision) or _err("Numerical Factorization.")
    return lu


def sparspaklu_refactor(lu, m, allow_pattern_change=True, precision=FP64):
    """sparspaklu! (SparseCSCInterface.jl:247-268): new values, ordering and
    symbolic factor reused when the pattern is unchanged."""
    n, colptr, rowval, nzval = _csc_arrays(m)
    pattern_changed = (n != lu.n or not np.array_equal(colptr, lu._colptr)
                       or not np.array_equal(rowval, lu._rowval))
    if pattern_changed:
        if allow_pattern_change or not lu._symbolicdone:
            fresh = SparseSolver(m, device=lu.slvr.device)
            lu.__dict__.update(fresh.__dict__)
        else:
            raise RuntimeError("'allow_pattern_change=false', but sparsity pattern of matrix "
                               "'m' does not match that used to create 'lu'")
    lu.p = m
    lu._nzval = nzval
    lu._orderdone or findorder(lu) or _err("Finding Order.")
    lu._symbolicdone or symbolicfactor(lu) or _err("Symbolic Factorization.")
    lu._inmatrixdone = False
    lu._factordone = False
    lu._trisolvedone = False
    inmatrix(lu) or _err("Matrix input.")
    factor(lu, precision) or _err("Numerical Factorization.")
    return lu


def ldiv(lu, v, u=None, **kw):
    """ldiv!(lu, v) / ldiv!(u, lu, v) (SparseCSCInterface.jl:277-293)."""
    if u is not None:
        u[...] = v
        v = u
    solve(lu, v, **kw) or _err("Triangular Solve.")
    lu._trisolvedone = False
    return v


def backslash(lu, v, **kw):
    """lu \\ v (SparseCSCInterface.jl:300)."""
    return ldiv(lu, np.array(v, dtype=np.float64, copy=True), **kw)
