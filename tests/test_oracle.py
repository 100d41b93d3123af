"""The CPU oracle (oracle/spk_oracle.cpp) pinned against the reference's
golden vectors, a dense textbook LU with pivoting confined to the diagonal
blocks, and direct dense solves (CPU only)."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle.oracle as orc
from helpers import (backward_error, blocked_lu_reference, csc1, panels_to_dense,
                     random_pattern, symbolic)
from sparspak_hip.gallery import convdiff3d, laplace2d
from sparspak_hip.problem import Problem
from sparspak_hip.symbolic import Graph, SymbolicFactor


def _tridiag_problem(n):
    p = Problem(n, n)
    for i in range(1, n):
        p.inaij(i + 1, i, -1.0)
        p.inaij(i, i, 4.0)
        p.inaij(i, i + 1, -1.0)
        p.inbi(i, 2.0 * i)
    p.inaij(n, n, 4.0)
    p.inbi(n, 3.0 * n + 1.0)
    return p


def _sym_from_problem(p, kind="lu"):
    n, cp, rv, nz = p.csc()
    s = SymbolicFactor(n, Graph.from_csc(n, cp, rv), kind=kind)
    s.findorder()
    s.symbolicfactor()
    return s, cp, rv, nz


def test_msprs005_inmatrix(fixtures):
    d = fixtures["tridiag11"]
    s, cp, rv, nz = _sym_from_problem(_tridiag_problem(11))
    lnz, unz, ipiv = orc.inmatrix(s, cp, rv, nz)
    assert lnz.tolist() == d["lnz_inmatrix"]
    assert unz.tolist() == d["unz_inmatrix"]


def test_msprs006_factor(fixtures):
    d = fixtures["tridiag11"]
    s, cp, rv, nz = _sym_from_problem(_tridiag_problem(11))
    lnz, unz, ipiv = orc.inmatrix(s, cp, rv, nz)
    rc, ff = orc.lufactor(s, lnz, unz, ipiv)
    assert rc == 0 and ff == 0
    gl = np.array(d["lnz_factor"])
    gu = np.array(d["unz_factor"])
    # reference tolerance 1e-6; the values agree to the last bits
    assert np.linalg.norm(lnz - gl) / np.linalg.norm(gl) < d["factor_rtol"]
    assert np.allclose(lnz, gl, rtol=4e-16, atol=0)
    assert np.array_equal(unz, gu)


@pytest.mark.parametrize("n", [11, 1101, 11000])
def test_tridiag_solve(n):
    # msprs007-009
    p = _tridiag_problem(n)
    s, cp, rv, nz = _sym_from_problem(p)
    lnz, unz, ipiv, rc, _ = orc.factor_csc(s, cp, rv, nz)
    x = orc.solve(s, lnz, unz, ipiv, p.rhs.copy())
    A = p.outsparse()
    xd = sp.linalg.spsolve(A.tocsc(), p.rhs)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-6
    assert backward_error(A, x, p.rhs) < 1e-15


def test_m31_solve(fixtures):
    # msprs010: fixed 31 x 31 matrix, rhs = 1:31
    d = fixtures["m31"]
    p = Problem(31, 31)
    p.insparse(d["I"], d["J"], d["V"])
    p.infullrhs(np.arange(1, 32, dtype=float))
    s, cp, rv, nz = _sym_from_problem(p)
    lnz, unz, ipiv, rc, _ = orc.factor_csc(s, cp, rv, nz)
    assert rc == 0
    x = orc.solve(s, lnz, unz, ipiv, p.rhs.copy())
    A = p.outsparse().toarray()
    xd = np.linalg.solve(A, p.rhs)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < d["rtol"]
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-14


def test_small4_solves(fixtures):
    # test_small.jl and test_structunsymm.jl fixed 4 x 4 matrices, b = e1
    for case in fixtures["small4"]:
        A = np.array(case["A"])
        p = Problem(4, 4)
        for i in range(4):
            for j in range(4):
                if A[i, j] != 0.0:
                    p.inaij(i + 1, j + 1, A[i, j])
        p.inbi(1, 1.0)
        s, cp, rv, nz = _sym_from_problem(p)
        lnz, unz, ipiv, rc, _ = orc.factor_csc(s, cp, rv, nz)
        x = orc.solve(s, lnz, unz, ipiv, p.rhs.copy())
        xd = np.linalg.solve(A, p.rhs)
        assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-14, case["source"]


def _dense_check(A, s, lnz, unz, ipiv, rtol):
    """The oracle's factor equals a dense right-looking LU (pivots confined
    to the supernode diagonal blocks) of P A P^T, entry by entry."""
    p = s.order.rperm - 1
    Ap = A.toarray()[np.ix_(p, p)]
    ref, rpiv = blocked_lu_reference(Ap, s.xsuper)
    got = panels_to_dense(s, lnz, unz)
    mask = got != 0
    assert np.array_equal(ipiv, rpiv)
    scale = np.abs(ref).max()
    assert np.max(np.abs(got - ref)) <= rtol * scale
    # nothing outside the symbolic structure is nonzero in the dense LU
    assert np.max(np.abs(ref[~mask]), initial=0.0) <= rtol * scale


@pytest.mark.parametrize("seed", range(10))
def test_lufactor_vs_dense_blocked_lu(seed):
    # weak diagonal -> interchanges inside the diagonal blocks happen
    A = random_pattern(40, 0.15, seed, symmetric=seed % 2 == 0, dominant=False)
    A = A + sp.diags(np.full(40, 0.3))
    s = symbolic(A)
    n, cp, rv, nz = csc1(A)
    lnz, unz, ipiv, rc, _ = orc.factor_csc(s, cp, rv, nz)
    _dense_check(A, s, lnz, unz, ipiv, 1e-11)
    b = np.arange(1.0, 41.0)
    x = orc.solve(s, lnz, unz, ipiv, b.copy())
    assert backward_error(A, x, b) < 1e-14


def test_lufactor_wide_supernodes_vs_dense():
    # dense blocks -> fundamental supernodes wider than 30 (split pieces)
    rng = np.random.default_rng(7)
    B = rng.uniform(-1, 1, (70, 70))
    A = sp.csc_matrix(B + np.diag(np.full(70, 0.5)))
    s = symbolic(A)
    assert s.nsuper > 1
    n, cp, rv, nz = csc1(A)
    lnz, unz, ipiv, rc, _ = orc.factor_csc(s, cp, rv, nz)
    _dense_check(A, s, lnz, unz, ipiv, 1e-10)
    assert (ipiv != np.concatenate([np.arange(1, w + 1) for w in np.diff(s.xsuper)])).any()


def test_lufactor_grid_problems():
    for A in (laplace2d(12), convdiff3d(6)):
        s = symbolic(A)
        n, cp, rv, nz = csc1(A)
        lnz, unz, ipiv, rc, _ = orc.factor_csc(s, cp, rv, nz)
        _dense_check(A, s, lnz, unz, ipiv, 1e-12)


def test_zero_pivot_flag():
    # structurally present but numerically singular diagonal block
    A = sp.csc_matrix(np.array([[0.0, 1.0], [1.0, 0.0]]) * 0 + np.array([[1.0, 1.0], [1.0, 1.0]]))
    s = symbolic(A)
    n, cp, rv, nz = csc1(A)
    lnz, unz, ipiv, rc, ff = orc.factor_csc(s, cp, rv, nz)
    assert rc == -1 and ff == 1


def test_float32_oracle_close_to_float64():
    A = convdiff3d(6)
    s = symbolic(A)
    n, cp, rv, nz = csc1(A)
    l64, u64, p64, _, _ = orc.factor_csc(s, cp, rv, nz, np.float64)
    l32, u32, p32, _, _ = orc.factor_csc(s, cp, rv, nz, np.float32)
    assert np.array_equal(p64, p32)
    assert np.max(np.abs(l32 - l64)) < 1e-5 * np.abs(l64).max()


# ---- LDL^T ----

def test_ldlt_reference_defect_two_by_two():
    # SpkLDLtFactor.jl:351-362 on A = [[4,-1],[-1,4]] (one 2-column
    # supernode): the reference computes d2 = 4 - (-1)(4)(-1) = 0
    A = sp.csc_matrix(np.array([[4.0, -1.0], [-1.0, 4.0]]))
    s = symbolic(A, kind="spd")
    assert s.nsuper == 1
    n, cp, rv, nz = csc1(A)
    lnz, _, _ = orc.inmatrix(s, cp, rv, nz)
    orc.ldltfactor(s, lnz, faithful=True)
    assert lnz[3] == 0.0  # d2 = 0: the reference's LDL^T breaks down
    lnz2, _, _ = orc.inmatrix(s, cp, rv, nz)
    orc.ldltfactor(s, lnz2)
    # corrected: d1 = 4, l21 = -1/4, d2 = 4 - 1/4
    assert np.allclose(lnz2, [4.0, -0.25, 0.0, 3.75])


@pytest.mark.parametrize("seed", range(6))
def test_ldlt_corrected_vs_dense(seed):
    A = random_pattern(45, 0.12, seed, symmetric=True)
    A = (A + A.T) * 0.5
    A = A + sp.diags(np.abs(A).sum(axis=1).A1 + 1.0)
    s = symbolic(A, kind="spd")
    n, cp, rv, nz = csc1(A)
    lnz, _, _ = orc.inmatrix(s, cp, rv, nz)
    orc.ldltfactor(s, lnz)
    p = s.order.rperm - 1
    Ap = A.toarray()[np.ix_(p, p)]
    Lc = np.linalg.cholesky(Ap)
    d = np.diag(Lc) ** 2
    Lunit = Lc / np.diag(Lc)
    got = panels_to_dense(s, lnz, None)
    assert np.allclose(np.diag(got), d, rtol=1e-12)
    low = np.tril(got, -1)
    assert np.allclose(low, np.tril(Lunit, -1), atol=1e-12)
    b = np.linspace(-1, 1, n)
    x = orc.solve(s, lnz, None, None, b.copy())
    assert backward_error(A, x, b) < 1e-14


def test_ldlt_wide_supernodes():
    rng = np.random.default_rng(3)
    B = rng.uniform(-1, 1, (130, 130))
    A = sp.csc_matrix(B @ B.T + 130 * np.eye(130))
    s = symbolic(A, kind="spd")
    assert s.nsuper >= 2
    n, cp, rv, nz = csc1(A)
    lnz, _, _ = orc.inmatrix(s, cp, rv, nz)
    orc.ldltfactor(s, lnz)
    b = np.ones(130)
    x = orc.solve(s, lnz, None, None, b.copy())
    assert backward_error(A, x, b) < 1e-14


def test_oracle_hager_estimate_pins():
    """The oracle's Hager restatement (ainverse_onenorm): exact on a diagonal
    matrix (||A^-1||_1 = max 1/|d_i|, found at the first unit-vector step),
    and a lower bound within 3x of the exact ||A^-1||_1 on symmetric
    matrices (dense inverses as the solve)."""
    import oracle.oracle as orc
    d = np.array([4.0, -0.5, 2.0, 8.0, 0.25, 3.0])
    est = orc.ainverse_onenorm(lambda b: b / d, d.size)
    assert est == pytest.approx(4.0, rel=1e-15)
    rng = np.random.default_rng(11)
    for n in (5, 17, 40):
        M = rng.uniform(-1, 1, (n, n))
        A = M + M.T + np.diag(np.abs(M).sum(axis=1) * 2 + 0.1)
        Ai = np.linalg.inv(A)
        exact = np.abs(Ai).sum(axis=0).max()
        est = orc.ainverse_onenorm(lambda b: Ai @ b, n)
        assert est <= exact * (1 + 1e-12)
        assert est >= exact / 3.0
