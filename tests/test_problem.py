"""The Problem store (SpkProblem.jl:105-251) and its walk into the CSC arrays
the device upload takes.

* A short input traced by hand through inaij! (SpkProblem.jl:204-251): a
  first entry, an entry appended at a list's end, one linked in before the
  head, a repeated entry (its value added), a second column: the exact
  head / link / rowsubs / values arrays and counters, and the CSC walk.
* Random input with repeats in random order: the walk equals scipy's COO ->
  CSC with duplicates summed (rows ascending in every column).
"""
import numpy as np
import scipy.sparse as sp

from sparspak_hip.problem import Problem, linked_to_csc


def test_inaij_hand_trace():
    p = Problem(3, 3, nnz=4)
    for r, c, v in [(2, 1, 1.0), (3, 1, 2.0), (1, 1, 3.0), (3, 1, 0.5), (2, 2, 4.0),
                    (1, 2, 5.0), (2, 2, 1.0)]:
        assert p.inaij(r, c, v)
    # entries e = 1..5 in input order: (2,1) (3,1) (1,1) (2,2) (1,2)
    assert list(p.head[:3]) == [3, 5, 0]
    assert list(p.link[:5]) == [2, 0, 1, 0, 4]
    assert list(p.rowsubs[:5]) == [2, 3, 1, 2, 1]
    assert list(p.values[:5]) == [1.0, 2.5, 3.0, 5.0, 5.0]
    assert (p.lastused, p.nnz, p.nedges, p.dnz, p.dedges) == (5, 5, 5, 2, 2)  # (1,1), (2,2)
    n, cp, rv, nz = p.csc()
    assert n == 3
    assert list(cp) == [1, 4, 6, 6]
    assert list(rv) == [1, 2, 3, 1, 2]
    assert list(nz) == [3.0, 1.0, 2.5, 5.0, 5.0]


def test_walk_equals_coo_sum():
    rng = np.random.default_rng(3)
    n = 40
    I = rng.integers(1, n + 1, 600)
    J = rng.integers(1, n + 1, 600)
    V = rng.uniform(-1, 1, 600)
    p = Problem(n, n, nnz=16)      # grows while input arrives
    for i, j, v in zip(I, J, V):
        p.inaij(int(i), int(j), float(v))
    m = sp.coo_matrix((V, (I - 1, J - 1)), shape=(n, n)).tocsc()
    m.sum_duplicates()
    m.sort_indices()
    cp, rv, nz = linked_to_csc(n, p.head, p.link, p.rowsubs, p.values)
    assert np.array_equal(cp, m.indptr + 1)
    assert np.array_equal(rv, m.indices + 1)
    assert np.allclose(nz, m.data, rtol=0, atol=1e-15)
    for c in range(n):
        assert np.all(np.diff(rv[cp[c] - 1:cp[c + 1] - 1]) > 0)
