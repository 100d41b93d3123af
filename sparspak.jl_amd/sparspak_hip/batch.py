"""Many independent systems in one device factorisation.

Each system keeps its own host symbolic phase (the reference's ordering and
symbolic factor, exactly as `SparseSolver(A_i)` would compute them), and the
per-system symbolic factors are concatenated into one block-diagonal symbolic
factor: offsets are added to every index and pointer array. The device then
sees one forest with one root per system and factors every system's fronts
of the same elimination-tree level in the same launches — thousands of
uneven systems without a launch per system. Consecutive systems with an
identical pattern share one host symbolic factor.

When every system has the same pattern (many value sets of one matrix), the
device holds that one symbolic factor and its plan once, and every launch
covers all value sets (spkhip_set_value_sets): the schedule, index maps and
solve plan are not replicated per system. The refinement certifies the
backward error per system. Results per system are identical to factoring
each system alone with the same kernels (no arithmetic crosses systems).
"""
import ctypes

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import as_i64, check, p64, pd
from .solver import FP32, FP64, _PREC, _csc_arrays, _Device
from .symbolic import Graph, SymbolicFactor


class BatchedSolver:
    """Factor and solve a list of independent sparse systems together.

    matrices: list of scipy sparse matrices (square). Systems whose pattern
    is identical to the previous one reuse its symbolic factor.
    kind: "lu" (SparseSolver) or "spd" (SparseSpdSolver).
    value_sets: None (default) takes the value-set mode when there are
    several systems and all share one pattern; False always concatenates the
    symbolic factors; True requires one shared pattern.
    """

    def __init__(self, matrices, kind="lu", device=0, value_sets=None):
        self.kind = kind
        self.device = device
        self.systems = []
        prev = None
        for A in matrices:
            n, cp, rv, nz = _csc_arrays(A)
            if prev is not None and prev[0] == n and np.array_equal(prev[1], cp) \
                    and np.array_equal(prev[2], rv):
                sym = prev[3]
            else:
                sym = SymbolicFactor(n, Graph.from_csc(n, cp, rv), kind=kind)
                sym.findorder()
                sym.symbolicfactor()
            self.systems.append((n, cp, rv, nz, sym))
            prev = (n, cp, rv, sym)
        one = all(s[4] is self.systems[0][4] for s in self.systems)
        if value_sets and not one:
            raise ValueError("value sets need one pattern shared by every system")
        self.value_sets = len(self.systems) > 1 and one if value_sets is None else bool(value_sets)
        self._concat()
        self._dev = None
        self._uploaded = False
        self.precision = FP64

    def _concat(self):
        parts = {k: [] for k in ("perm", "invp", "xsuper", "snode", "xlindx", "lindx", "xlnz",
                                 "xunz", "colptr", "rowval", "nzval")}
        col = sup = sub = lz = uz = nnz = 0
        self.offsets = []
        for n, cp, rv, nz, s in self.systems:
            o = s.order
            self.offsets.append(col)
            parts["perm"].append(o.rperm + col)
            parts["invp"].append(o.rinvp + col)
            parts["xsuper"].append(s.xsuper[:-1] + col)
            parts["snode"].append(s.snode + sup)
            parts["xlindx"].append(s.xlindx[:-1] + sub)
            parts["lindx"].append(s.lindx + col)
            parts["xlnz"].append(s.xlnz[:-1] + lz)
            if s.xunz is not None:
                parts["xunz"].append(s.xunz[:-1] + uz)
            parts["colptr"].append(cp[:-1] + nnz)
            parts["rowval"].append(rv + col)
            parts["nzval"].append(nz)
            col += n
            sup += s.nsuper
            sub += int(s.xlindx[-1] - 1)
            lz += int(s.xlnz[-1] - 1)
            uz += int(s.xunz[-1] - 1) if s.xunz is not None else 0
            nnz += int(cp[-1] - 1)
        cat = lambda k, last=None: as_i64(np.concatenate(parts[k] + ([np.array([last])]
                                                                    if last is not None else [])))
        self.n = col
        self.nsuper = sup
        self.perm = cat("perm")
        self.invp = cat("invp")
        self.xsuper = cat("xsuper", col + 1)
        self.snode = cat("snode")
        self.xlindx = cat("xlindx", sub + 1)
        self.lindx = cat("lindx")
        self.xlnz = cat("xlnz", lz + 1)
        self.xunz = cat("xunz", uz + 1) if self.kind == "lu" else np.zeros(col + 1, np.int64)
        self.colptr = cat("colptr", nnz + 1)
        self.rowval = cat("rowval")
        self.nzval = np.concatenate(parts["nzval"])
        self.nnz = nnz

    def flops(self):
        return float(sum(s.flops() for *_, s in self.systems))

    def upload(self):
        if self._uploaded:
            return
        if self._dev is None:
            self._dev = _Device(self.device)
        L = self._dev.L
        kind = 0 if self.kind == "lu" else 1
        if self.value_sets:
            n, cp, rv, _, s = self.systems[0]
            xunz = s.xunz if s.xunz is not None else np.zeros(n + 1, dtype=np.int64)
            rc = L.spkhip_symbolic_upload(self._dev.h, kind, n, s.nsuper, p64(s.xsuper),
                                          p64(s.snode), p64(s.xlindx), p64(s.lindx),
                                          p64(s.xlnz), p64(xunz), p64(s.order.rperm),
                                          p64(s.order.rinvp), p64(cp), p64(rv))
            if rc == -1:
                raise RuntimeError("No space for matrix element.")
            check(rc, "spkhip_symbolic_upload")
            check(L.spkhip_set_value_sets(self._dev.h, len(self.systems)),
                  "spkhip_set_value_sets")
            self._uploaded = True
            return
        rc = L.spkhip_symbolic_upload(self._dev.h, kind, self.n, self.nsuper, p64(self.xsuper),
                                      p64(self.snode), p64(self.xlindx), p64(self.lindx),
                                      p64(self.xlnz), p64(self.xunz), p64(self.perm),
                                      p64(self.invp), p64(self.colptr), p64(self.rowval))
        if rc == -1:
            raise RuntimeError("No space for matrix element.")
        check(rc, "spkhip_symbolic_upload")
        seg = as_i64(self.offsets + [self.n])
        check(L.spkhip_set_segments(self._dev.h, len(self.offsets), p64(seg)),
              "spkhip_set_segments")
        self._uploaded = True

    def set_values(self, values=None):
        """inmatrix for every system; values: list of per-system nzval arrays
        (CSC order of each system) or None for the matrices' own values."""
        self.upload()
        if values is not None:
            if len(values) != len(self.systems):
                raise ValueError(f"{len(values)} value arrays for {len(self.systems)} systems")
            for i, (v, (_, cp, *_r)) in enumerate(zip(values, self.systems)):
                if np.size(v) != int(cp[-1] - 1):
                    raise ValueError(f"system {i}: {np.size(v)} values for a pattern of "
                                     f"{int(cp[-1] - 1)} entries")
        nz = self.nzval if values is None else np.concatenate(
            [np.asarray(v, dtype=np.float64).ravel() for v in values])
        nz = np.ascontiguousarray(nz, dtype=np.float64)
        if nz.shape[0] != self.nnz:
            raise ValueError("value count does not match the batched pattern")
        check(self._dev.L.spkhip_inmatrix(self._dev.h, pd(nz)), "spkhip_inmatrix")

    def factor(self, precision=FP64):
        self.precision = _PREC[precision]
        rc = self._dev.L.spkhip_factor(self._dev.h, self.precision)
        if rc == -1:
            raise RuntimeError("Numerical factorization: zero pivot in a diagonal block.")
        check(rc, "spkhip_factor")

    def solve(self, rhs, refine=None, tol=1e-14, max_iter=10):
        """rhs: list of per-system arrays (n_i or n_i x k, the same k for all)
        or one stacked (n_total x k) array. Returns the solutions in the same
        form."""
        stacked = not isinstance(rhs, (list, tuple))
        B = np.asarray(rhs, dtype=np.float64) if stacked else np.concatenate(
            [np.asarray(b, dtype=np.float64).reshape(s[0], -1) for b, s in zip(rhs, self.systems)])
        vec = B.ndim == 1
        B = np.asfortranarray(B.reshape(self.n, -1))
        if refine is None:
            refine = self.precision != FP64
        L = self._dev.L
        if refine:
            rc = L.spkhip_solve_refine(self._dev.h, B.shape[1], pd(B), self.n, float(tol),
                                       int(max_iter))
            _lib.check_refine(rc, "spkhip_solve_refine",
                              self.stats().backward_error if rc == -5 else 0.0)
        else:
            check(L.spkhip_solve(self._dev.h, B.shape[1], pd(B), self.n), "spkhip_solve")
        if stacked:
            return B[:, 0] if vec else B
        out = []
        for (n, *_), off, b in zip(self.systems, self.offsets, rhs):
            x = B[off:off + n]
            out.append(x[:, 0] if np.ndim(b) == 1 else x)
        return out

    def solve_device(self, d_b, nrhs=1, refine=None, tol=1e-14, max_iter=10):
        """The stacked right-hand sides already in device memory, overwritten
        by the solutions: d_b is a device address (int) or an object with
        data_ptr() (a torch tensor), float64, column-major n x nrhs. The
        library orders its reads after the caller's stream (the legacy
        default stream unless set otherwise)."""
        ptr = ctypes.c_void_p(d_b.data_ptr() if hasattr(d_b, "data_ptr") else int(d_b))
        if refine is None:
            refine = self.precision != FP64
        L = self._dev.L
        if refine:
            rc = L.spkhip_solve_refine_device(self._dev.h, int(nrhs), ptr, self.n, float(tol),
                                              int(max_iter))
            _lib.check_refine(rc, "spkhip_solve_refine_device",
                              self.stats().backward_error if rc == -5 else 0.0)
        else:
            check(L.spkhip_solve_device(self._dev.h, int(nrhs), ptr, self.n),
                  "spkhip_solve_device")

    def autotune(self, precision=FP64, reps=2):
        """Plan knobs picked by timing this batch's factorisation (spkhip_autotune);
        leaves the batch factored with the winner. Returns the knobs."""
        self.precision = _PREC[precision]
        t = _lib.Tuning()
        check(self._dev.L.spkhip_autotune(self._dev.h, self.precision, int(reps),
                                          ctypes.byref(t)), "spkhip_autotune")
        return {k: getattr(t, k) for k in _lib.TUNING_KNOBS}

    def stats(self):
        st = _lib.Stats()
        check(self._dev.L.spkhip_stats(self._dev.h, ctypes.byref(st)), "spkhip_stats")
        return st

    def block_diagonal(self):
        """The batch as one scipy matrix (original numbering with offsets)."""
        return sp.block_diag([sp.csc_matrix((nz, rv - 1, cp - 1), shape=(n, n))
                              for n, cp, rv, nz, _ in self.systems], format="csr")
