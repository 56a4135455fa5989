"""The oracle's LU refactorization (lu.rs:662-757), pinned by the reference's refactor tests
(ybus_sparse/tests/factorization_test.rs:955-1141), closed forms, and a dense textbook
elimination: left-looking Doolittle on P A Q with the factor's pivots, each column's updates
taken in increasing order with a separate multiply and subtract, which gives every stored entry
the same operations in the same order as the sparse loop (a skipped structural zero adds nothing
to a nonzero entry), so its values are the same bits."""
import numpy as np
import pytest

import oracle as O
from tests.conftest import Lcg
from tests.fixtures import random_square_sparse


def with_values(a: O.Matrix, data) -> O.Matrix:
    return O.Matrix(a.rows, a.cols, a.row_major, a.ptr.copy(), a.idx.copy(), np.asarray(data, dtype=np.float64))


def new_values(a: O.Matrix, seed: int) -> O.Matrix:
    """factorization_test.rs:955-964: every stored value times a factor in [0.5, 1.5)."""
    rng = Lcg(seed)
    data = a.data.copy()
    for p in range(data.shape[0]):
        data[p] *= 0.5 + (rng.next() % 1000) / 1000.0
    return with_values(a, data)


def ring_blocks(nodes: int, seed: int) -> O.Matrix:
    """A Newton-like matrix of dense 2 x 2 blocks, as factorization_test.rs:1066-1095 describes
    it: node i couples to i + 1 around a ring and, every third node, to the node a third of the
    way round; each diagonal block is dominant. CSC."""
    rng = Lcg(seed)

    def value():
        return 0.5 + (rng.next() % 1000) / 1000.0

    t = []

    def couple(i, j):
        for a, b in [(0, 0), (0, 1), (1, 0), (1, 1)]:
            t.append((2 * i + a, 2 * j + b, -value()))
            t.append((2 * j + a, 2 * i + b, -value()))

    for i in range(nodes):
        couple(i, (i + 1) % nodes)
        if i % 3 == 0:
            couple(i, (i + nodes // 3) % nodes)
    for i in range(nodes):
        for a, b in [(0, 0), (0, 1), (1, 0), (1, 1)]:
            t.append((2 * i + a, 2 * i + b, (20.0 if a == b else 0.0) + value()))
    return O.Matrix.from_triplets(2 * nodes, 2 * nodes, t, row_major=False)


def dense(m: O.Matrix) -> np.ndarray:
    return m.to_dense()


def reconstructs(a: O.Matrix, lu: O.Lu, tol: float = 1e-12) -> bool:
    """P A Q = L U to within tol relative to the largest entry."""
    pa = dense(a)[np.asarray(lu.row_perm)][:, np.asarray(lu.col_perm)]
    return np.abs(dense(lu.l) @ dense(lu.u) - pa).max() <= tol * max(1.0, np.abs(pa).max())


def textbook_refactor(a: O.Matrix, lu: O.Lu):
    """Dense left-looking Doolittle on P A Q, no pivot search: column k takes x[j+1:] -= L[j+1:, j]
    * x[j] for j = 0 .. k - 1 in order, then U[:k+1, k] = x[:k+1], L[k+1:, k] = x[k+1:] / x[k]."""
    pa = dense(a)[np.asarray(lu.row_perm)][:, np.asarray(lu.col_perm)]
    n = pa.shape[0]
    lower, upper = np.eye(n), np.zeros((n, n))
    for k in range(n):
        x = pa[:, k].copy()
        for j in range(k):
            product = lower[j + 1:, j] * x[j]
            x[j + 1:] = x[j + 1:] - product
        upper[:k + 1, k] = x[:k + 1]
        lower[k + 1:, k] = x[k + 1:] / x[k]
    return lower, upper


def stored(m: O.Matrix, full: np.ndarray) -> np.ndarray:
    """full's values at m's stored (CSC) positions."""
    cols = np.repeat(np.arange(m.cols), np.diff(m.ptr))
    return full[m.idx, cols]


def bits(lu: O.Lu):
    return np.concatenate([lu.l.data, lu.u.data]).view(np.uint64).copy()


def test_two_by_two_closed_form():
    a = O.Matrix.from_triplets(2, 2, [(0, 0, 2.0), (0, 1, 1.0), (1, 0, 1.0), (1, 1, 3.0)], row_major=False)
    lu = O.Lu(a, O.NATURAL)
    b = O.Matrix.from_triplets(2, 2, [(0, 0, 4.0), (0, 1, 1.0), (1, 0, 2.0), (1, 1, 5.0)], row_major=False)
    assert lu.refactor(b) == "reused"
    # L = [1 0; 1/2 1], U = [4 1; 0 5 - 1/2], every value exact
    assert list(lu.l.data) == [1.0, 0.5, 1.0]
    assert list(lu.u.data) == [4.0, 1.0, 4.5]
    assert list(lu.row_perm) == [0, 1]


def test_refactor_reuses_the_pivots_for_new_values():
    # factorization_test.rs:966-991
    for n, fill, scale, seed in [(20, 20, 1.5, 1), (90, 4, 3.0, 2)]:
        a = random_square_sparse(n, fill, scale, seed).to_csc()
        lu = O.Lu(a, O.AMD)
        l_idx, u_idx = lu.l.idx.copy(), lu.u.idx.copy()
        perms = (lu.row_perm.copy(), lu.col_perm.copy())
        for round_ in range(3):
            b = new_values(a, 10 + round_)
            assert lu.refactor(b) == "reused"
            assert np.array_equal(lu.l.idx, l_idx) and np.array_equal(lu.u.idx, u_idx)
            assert np.array_equal(lu.row_perm, perms[0]) and np.array_equal(lu.col_perm, perms[1])
            assert reconstructs(b, lu)
            x = np.array([1.0 + (i % 5) for i in range(n)])
            rhs = b.mul_vec(x)
            assert np.abs(lu.solve(rhs) - x).max() < 1e-9
            fresh = O.Lu(b, O.AMD)
            det = np.prod(lu.u.data[lu.u.ptr[1:] - 1])
            det_fresh = np.prod(fresh.u.data[fresh.u.ptr[1:] - 1])
            assert abs(abs(det) - abs(det_fresh)) <= 1e-9 * abs(det_fresh)


def test_refactor_repivots_when_a_pivot_becomes_too_small():
    # factorization_test.rs:993-1030
    t = [(0, 0, 4.0), (0, 1, 1.0), (1, 0, 1.0), (1, 1, 4.0), (1, 2, 1.0), (2, 1, 1.0), (2, 2, 4.0)]
    a = O.Matrix.from_triplets(3, 3, t, row_major=False)
    lu = O.Lu(a, O.NATURAL)
    assert list(lu.row_perm) == [0, 1, 2]
    p = 0  # (0, 0) is column 0's first stored entry
    assert a.idx[p] == 0
    data = a.data.copy()
    data[p] = 0.01
    b = with_values(a, data)
    assert lu.refactor(b) == "repivoted"
    assert lu.row_perm[0] == 1
    assert reconstructs(b, lu)
    data[p] = 0.0
    b = with_values(a, data)
    assert lu.refactor(b) == "reused"
    fresh = O.Lu(a, O.NATURAL)
    assert fresh.refactor(b) == "repivoted"
    assert reconstructs(b, fresh)
    assert np.abs(lu.solve([1.0, 2.0, 3.0]) - np.linalg.solve(dense(b), [1.0, 2.0, 3.0])).max() < 1e-12
    assert lu.refactor(a) == "repivoted"
    assert list(lu.row_perm) == [0, 1, 2]
    assert reconstructs(a, lu)


def test_refactor_errors_leave_the_factorization_unchanged():
    # factorization_test.rs:1048-1065 and :1129-1141
    a = O.Matrix.from_triplets(2, 2, [(0, 0, 2.0), (0, 1, 1.0), (1, 0, 1.0), (1, 1, 3.0)], row_major=False)
    lu = O.Lu(a, O.AMD)
    before = bits(lu)
    other = O.Matrix.from_triplets(2, 2, [(0, 0, 2.0), (1, 0, 1.0), (1, 1, 3.0)], row_major=False)
    with pytest.raises(O.SparseError) as e:
        lu.refactor(other)
    assert e.value.kind == "PatternMismatch"
    singular = O.Matrix.from_triplets(2, 2, [(0, 0, 1.0), (0, 1, 2.0), (1, 0, 2.0), (1, 1, 4.0)], row_major=False)
    with pytest.raises(O.SparseError) as e:
        lu.refactor(singular)
    assert e.value.kind == "SingularMatrix"
    assert np.array_equal(bits(lu), before)
    assert np.abs(lu.solve([3.0, 4.0]) - [1.0, 1.0]).max() < 1e-12
    b = O.Matrix.from_triplets(2, 2, [(0, 0, 4.0), (0, 1, 1.0), (1, 0, 2.0), (1, 1, 5.0)], row_major=False)
    assert lu.refactor(b) == "reused"
    first = bits(lu)
    with pytest.raises(O.SparseError):
        lu.refactor(singular)
    assert lu.refactor(a) == "reused"
    assert lu.refactor(b) == "reused"
    assert np.array_equal(bits(lu), first)


@pytest.mark.parametrize("nodes,seed", [(12, 5), (60, 6), (100, 7)])
def test_refactor_of_ring_blocks_repeats_bit_for_bit(nodes, seed):
    # factorization_test.rs:1097-1127
    a = ring_blocks(nodes, seed)
    lu = O.Lu(a, O.AMD)
    b = new_values(a, seed + 100)
    assert lu.refactor(b) == "reused"
    assert reconstructs(b, lu)
    n = 2 * nodes
    x = np.array([1.0 + (i % 7) for i in range(n)])
    assert np.abs(lu.solve(b.mul_vec(x)) - x).max() < 1e-9
    first = bits(lu)
    assert lu.refactor(a) == "reused"
    assert lu.refactor(b) == "reused"
    assert np.array_equal(bits(lu), first)


@pytest.mark.parametrize("case", ["ring", "random", "jacobian"])
def test_refactor_equals_the_dense_textbook_elimination_bit_for_bit(case):
    if case == "ring":
        a, ordering = ring_blocks(40, 3), O.AMD
        b = new_values(a, 77)
    elif case == "random":
        a, ordering = random_square_sparse(70, 5, 2.0, 9).to_csc(), O.COLAMD
        b = new_values(a, 77)
    else:
        # the Jacobian at a nearby operating point, as a Newton iterate refactors it
        grid = O.make_grid(80, 20261205)
        a, ordering = grid.jacobian(20261205)[0], O.COLAMD
        b = grid.jacobian(20261205, 0.01, 7)[0]
        assert np.array_equal(a.ptr, b.ptr) and np.array_equal(a.idx, b.idx)
        assert not np.array_equal(a.data, b.data)
    lu = O.Lu(a, ordering)
    assert lu.refactor(b) == "reused"
    lower, upper = textbook_refactor(b, lu)
    assert np.array_equal(lu.l.data.view(np.uint64), stored(lu.l, lower).view(np.uint64))
    assert np.array_equal(lu.u.data.view(np.uint64), stored(lu.u, upper).view(np.uint64))
    # and the structural zeros the sparse loop never touches stay zero in the dense one
    assert np.count_nonzero(lower) <= lu.l.nnz and np.count_nonzero(upper) <= lu.u.nnz


def test_refactor_of_a_reversed_update_order_is_caught():
    # The dense pin is sensitive to the order of each entry's updates: the same elimination with
    # the rows below the diagonal taking their terms in decreasing j (mathematically the same
    # factors) differs from the oracle in some bits
    a = ring_blocks(40, 3)
    lu = O.Lu(a, O.AMD)
    lu.refactor(new_values(a, 77))
    b = new_values(a, 77)
    pa = dense(b)[np.asarray(lu.row_perm)][:, np.asarray(lu.col_perm)]
    n = pa.shape[0]
    lower, upper = np.eye(n), np.zeros((n, n))
    for k in range(n):
        x = pa[:, k].copy()
        for j in range(k):
            x[j + 1:k + 1] = x[j + 1:k + 1] - lower[j + 1:k + 1, j] * x[j]
        for j in reversed(range(k)):
            x[k + 1:] = x[k + 1:] - lower[k + 1:, j] * x[j]
        upper[:k + 1, k] = x[:k + 1]
        lower[k + 1:, k] = x[k + 1:] / x[k]
    assert np.abs(stored(lu.l, lower) - lu.l.data).max() < 1e-12
    assert not np.array_equal(stored(lu.l, lower).view(np.uint64), lu.l.data.view(np.uint64))


def test_refactor_work_counts_the_multiply_subtracts():
    a = ring_blocks(30, 4)
    lu = O.Lu(a, O.AMD)
    lp, up, ui = lu.l.ptr, lu.u.ptr, lu.u.idx
    expected = sum(int(lp[j + 1] - lp[j] - 1) for k in range(lu.n) for j in ui[up[k]:up[k + 1] - 1])
    assert lu.refactor_work() == expected > 0
