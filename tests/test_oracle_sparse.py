"""The oracle's compressed-matrix core and orderings, pinned by the reference's own test
literals and properties (ybus_sparse/tests/factorization_test.rs, colamd_test.rs,
compressed_matrix_test.rs) and by brute force."""
import numpy as np
import pytest

import oracle as O
from tests.conftest import Lcg
from tests.fixtures import (banded, graph_matrix, grid_laplacian, is_permutation, random_spd)


def test_from_triplets_sums_duplicates_in_input_order_and_keeps_stored_zeros():
    # compressed_matrix.rs:234-249: duplicates summed in the order given, zeros stored
    big, small = 1e16, 1.0
    m = O.Matrix.from_triplets(2, 3, [(1, 2, big), (0, 1, 0.0), (1, 2, small), (1, 2, -big), (1, 0, 2.0)])
    assert list(m.ptr) == [0, 1, 3]
    assert list(m.idx) == [1, 0, 2]
    # (1e16 + 1) - 1e16 = 0 in input order; a different order would give 1
    assert list(m.data) == [0.0, 2.0, (big + small) - big]
    with pytest.raises(O.SparseError) as e:
        O.Matrix.from_triplets(2, 2, [(2, 0, 1.0)])
    assert e.value.kind == "OutOfBounds"


def test_reorientation_and_select_match_dense():
    a = random_spd(30, 20, 4)
    d = a.to_dense()
    assert np.array_equal(a.to_csc().to_dense(), d)
    rng = Lcg(3)
    perm = rng.permutation(30)
    assert np.array_equal(a.permute_symmetric(perm).to_dense(), d[np.ix_(perm, perm)])
    rows, cols = [3, 1, 1, 7], [0, 29, 5]
    assert np.array_equal(a.select(rows, cols).to_dense(), d[np.ix_(rows, cols)])


def test_mul_vec_sums_each_row_from_zero_in_index_order():
    a = random_spd(40, 10, 9)
    x = np.array([np.sin(0.3 * i) for i in range(40)])
    y = a.mul_vec(x)
    d = a.to_dense()
    for i in range(40):
        s = 0.0
        for j in range(40):
            if d[i, j] != 0.0 or (j in set(a.idx[a.ptr[i]:a.ptr[i + 1]])):
                s = s + d[i, j] * x[j]
        assert y[i] == s
    assert np.allclose(a.to_csc().mul_vec(x), d @ x, rtol=1e-14, atol=1e-14)


def test_natural_ordering_is_the_identity():
    assert list(grid_laplacian(3).ordering(O.NATURAL)) == list(range(9))


def profile(a: O.Matrix) -> int:
    d = a.to_dense()
    n = a.rows
    total = 0
    for i in range(n):
        nz = [j for j in range(i + 1) if d[i, j] != 0.0 or j == i]
        total += i - min(nz)
    return total


def bandwidth(a: O.Matrix) -> int:
    d = a.to_dense()
    ii, jj = np.nonzero(d)
    return int(np.max(np.abs(ii - jj))) if len(ii) else 0


def test_rcm_does_not_widen_banded_matrices():
    # factorization_test.rs:102-115
    for n, band in [(10, 1), (20, 3), (7, 6)]:
        a = banded(n, band)
        perm = a.ordering(O.RCM)
        assert is_permutation(perm, n)
        reordered = a.permute_symmetric(perm)
        assert bandwidth(reordered) <= band
        assert profile(reordered) <= profile(a)


def test_rcm_handles_components_and_isolated_nodes():
    # factorization_test.rs:140-152: two paths and an isolated node
    a = graph_matrix(6, [(0, 3), (3, 5), (1, 4)])
    perm = a.ordering(O.RCM)
    assert is_permutation(perm, 6)
    assert bandwidth(a.permute_symmetric(perm)) == 1


def test_amd_returns_a_permutation_that_depends_only_on_the_pattern():
    # factorization_test.rs:565-586
    for n, fill, seed in [(1, 0, 1), (2, 100, 2), (17, 30, 3), (80, 5, 4), (150, 2, 5)]:
        a = random_spd(n, fill, seed)
        perm = a.ordering(O.AMD)
        assert is_permutation(perm, n)
        assert list(a.to_csc().ordering(O.AMD)) == list(perm)
        scaled = O.Matrix(a.rows, a.cols, True, a.ptr, a.idx, a.data * -2.0)
        assert list(scaled.ordering(O.AMD)) == list(perm)
        assert list(a.ordering(O.AMD)) == list(perm)
    rng = Lcg(21)
    shuffled = grid_laplacian(9).permute_symmetric(rng.permutation(81))
    assert is_permutation(shuffled.ordering(O.AMD), 81)
    empty = O.Matrix(0, 0, True, np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0))
    assert len(empty.ordering(O.AMD)) == 0


def fill_in(a: O.Matrix, ordering: int) -> int:
    """factorization_test.rs:559-562."""
    lower = sum(1 for i in range(a.rows) for p in range(a.ptr[i], a.ptr[i + 1]) if a.idx[p] <= i)
    return O.Cholesky(a, ordering, symbolic_only=True).predicted_entries - lower


def test_amd_orders_trees_and_arrowheads_without_fill():
    # factorization_test.rs:589-613
    rng = Lcg(8)
    for n in [10, 100, 1000]:
        edges = [(i, rng.next() % i) for i in range(1, n)]
        tree = graph_matrix(n, edges).permute_symmetric(rng.permutation(n))
        assert fill_in(tree, O.AMD) == 0, n
    for n, hub in [(12, 5), (500, 123)]:
        edges = [(hub, i) for i in range(n) if i != hub]
        arrow = graph_matrix(n, edges)
        perm = list(arrow.ordering(O.AMD))
        assert hub in perm[n - 2:]
        if n > 160:
            assert perm[-1] == hub
        assert fill_in(arrow, O.AMD) == 0
        assert fill_in(arrow, O.NATURAL) > 0


def test_amd_sets_dense_rows_aside():
    # factorization_test.rs:616-637: the two dense nodes are ordered last, [900, 901]
    k = 30
    edges = []
    for r in range(k):
        for c in range(k):
            i = r * k + c
            if c + 1 < k:
                edges.append((i, i + 1))
            if r + 1 < k:
                edges.append((i, i + k))
            edges += [(i, 900), (i, 901)]
    perm = graph_matrix(902, edges).ordering(O.AMD)
    assert is_permutation(perm, 902)
    assert list(perm[900:]) == [900, 901]


def test_amd_fills_grids_less_than_rcm():
    # factorization_test.rs:640-649
    rng = Lcg(13)
    for k in [10, 20, 40]:
        grid = grid_laplacian(k).permute_symmetric(rng.permutation(k * k))
        assert fill_in(grid, O.AMD) < fill_in(grid, O.RCM)


def random_pattern(m, n, per_row, seed):
    """colamd_test.rs:25-35."""
    rng = Lcg(seed)
    t = []
    for row in range(m):
        for _ in range(per_row):
            col = rng.next() % n
            t.append((row, col, (float(rng.next() % 2001) - 1000.0) / 317.0))
    return t


def test_colamd_gives_a_column_permutation_of_any_shape():
    # colamd_test.rs:77-126
    for m, n, per_row, seed in [(50, 50, 3, 1), (80, 30, 4, 2), (30, 80, 6, 3), (200, 200, 2, 4)]:
        a = O.Matrix.from_triplets(m, n, random_pattern(m, n, per_row, seed))
        perm = a.ordering(O.COLAMD)
        assert is_permutation(perm, n)
        assert list(a.to_csc().ordering(O.COLAMD)) == list(perm)
        assert list(a.ordering(O.COLAMD)) == list(perm)
    n = 300
    t = [(i, i, 1.0) for i in range(n) if i % 10 != 3]
    t += [(5, j, 1.0) for j in range(n)]
    t += [(i, 7, 1.0) for i in range(n)]
    t += [(i, i + 1, 2.0) for i in range(n - 1)]
    a = O.Matrix.from_triplets(n, n, t, row_major=False)
    perm = a.ordering(O.COLAMD)
    assert is_permutation(perm, n)
    assert int(perm[-1]) == 7
    empty = O.Matrix.from_triplets(4, 3, [], row_major=False)
    assert is_permutation(empty.ordering(O.COLAMD), 3)
    with pytest.raises(O.SparseError) as e:
        empty.ordering(O.AMD)
    assert e.value.kind == "NotSquare"
