"""The oracle's Cholesky and LU factorizations, solves, refinement and supernodes, pinned by the
reference's test literals (ybus_sparse/tests/factorization_test.rs, accuracy_test.rs,
rayon_test.rs, front_tree_test.rs), brute-force symbolic elimination, the pre-Phase-3
independent algorithms (tests/common/mod.rs) and dense numpy/scipy checks."""
import math

import numpy as np
import pytest
import scipy.linalg

import oracle as O
from tests.conftest import Lcg
from tests.fixtures import (brute_force_factor_pattern, graph_matrix, grid_laplacian, is_permutation,
                            random_spd, random_square_cuda, random_square_sparse, reference_cholesky,
                            reference_lu_perm, two_by_two_blocks)


@pytest.mark.parametrize("ordering", [O.NATURAL, O.RCM, O.AMD])
def test_analysis_matches_brute_force_elimination(ordering):
    # factorization_test.rs:246-274 (Natural, RCM; AMD added: brute force holds for any order)
    for n, fill, seed in [(1, 0, 1), (12, 20, 2), (40, 8, 3), (60, 4, 4), (30, 50, 5)]:
        a = random_spd(n, fill, seed)
        s = O.Cholesky(a, ordering, symbolic_only=True)
        assert s.n == n
        assert list(s.perm) == list(a.ordering(ordering))
        filled = brute_force_factor_pattern(a, s.perm)
        parent = [next((i for i in range(j + 1, n) if filled[i, j]), None) for j in range(n)]
        assert s.elimination_tree() == parent
        counts = [int(filled[j:, j].sum()) for j in range(n)]
        assert list(s.column_counts) == counts
        assert s.predicted_entries == sum(counts)
        # each row's pattern is its brute-force row, sorted, diagonal last
        for i in range(n):
            row = list(s.pattern_idx[s.pattern_ptr[i]:s.pattern_ptr[i + 1]])
            assert row == [j for j in range(i + 1) if filled[i, j]]


def test_analysis_literals():
    # factorization_test.rs:312-334: only the lower triangle stored, one entry a stored zero
    a = O.Matrix.from_triplets(3, 3, [(0, 0, 4.0), (1, 1, 4.0), (2, 2, 4.0), (2, 0, 0.0), (2, 1, 1.0)])
    s = O.Cholesky(a, O.NATURAL, symbolic_only=True)
    assert s.elimination_tree() == [2, 2, None]
    assert list(s.column_counts) == [2, 2, 1]
    ident = O.Matrix.from_triplets(3, 3, [(i, i, 1.0) for i in range(3)])
    forest = O.Cholesky(ident, O.RCM, symbolic_only=True)
    assert forest.elimination_tree() == [None, None, None]
    assert forest.predicted_entries == 3
    with pytest.raises(O.SparseError) as e:
        O.Cholesky(O.Matrix.from_triplets(2, 3, []), O.RCM, symbolic_only=True)
    assert e.value.kind == "NotSquare"


def test_postorder_visits_descendants_first():
    # factorization_test.rs:277-309
    a = grid_laplacian(7)
    for ordering in [O.NATURAL, O.RCM, O.AMD]:
        s = O.Cholesky(a, ordering, symbolic_only=True)
        order = list(s.postorder)
        assert is_permutation(order, 49)
        position = [0] * 49
        for k, node in enumerate(order):
            position[node] = k
        tree = s.elimination_tree()
        size = [1] * 49
        for node in range(49):
            if tree[node] is not None:
                assert position[node] < position[tree[node]]
                assert tree[node] > node
                size[tree[node]] += size[node]
        for node in range(49):
            first = position[node] + 1 - size[node]
            for member in order[first:position[node] + 1]:
                anc = member
                while anc is not None and anc != node:
                    anc = tree[anc]
                assert anc == node


def test_natural_ordering_matches_the_unordered_cholesky_bit_for_bit():
    # factorization_test.rs:396-417 against tests/common/mod.rs:12-62
    for n, fill, seed in [(1, 0, 1), (15, 30, 2), (50, 6, 3), (90, 3, 4)]:
        a = random_spd(n, fill, seed)
        f = O.Cholesky(a, O.NATURAL)
        ptr, idx, data = reference_cholesky(a)
        assert list(f.l.ptr) == list(ptr)
        assert list(f.l.idx) == list(idx)
        assert f.l.data.tobytes() == np.asarray(data, dtype=np.float64).tobytes()


@pytest.mark.parametrize("ordering", [O.NATURAL, O.RCM, O.AMD])
def test_ordered_cholesky_reconstructs_the_reordered_matrix(ordering):
    # factorization_test.rs:372-393
    a = random_spd(60, 6, 8)
    f = O.Cholesky(a, ordering)
    assert f.l.nnz == f.predicted_entries
    l = f.l.to_dense()
    assert np.all(np.triu(l, 1) == 0.0)
    assert np.all(np.diag(l) > 0.0)
    pap = a.permute_symmetric(f.perm).to_dense()
    assert np.max(np.abs(l @ l.T - pap)) <= 1e-12
    # L^T stored as the adjoint, diagonal first in each row
    assert np.array_equal(f.l_adjoint.to_dense(), l.T)
    for i in range(f.n):
        assert f.l_adjoint.idx[f.l_adjoint.ptr[i]] == i
        assert f.l.idx[f.l.ptr[i + 1] - 1] == i
    # values against scipy's dense Cholesky of the same reordering
    assert np.allclose(l, scipy.linalg.cholesky(pap, lower=True), rtol=1e-12, atol=1e-14)


def test_cholesky_keeps_a_diagonal_matrix_diagonal_and_reports_errors():
    # factorization_test.rs:531-541, 484-528
    d = O.Cholesky(O.Matrix.from_triplets(3, 3, [(0, 0, 4.0), (1, 1, 9.0), (2, 2, 16.0)]))
    assert d.l.nnz == 3
    assert sorted(d.l.data) == [2.0, 3.0, 4.0]
    with pytest.raises(O.SparseError) as e:
        O.Cholesky(O.Matrix.from_triplets(2, 2, [(0, 0, 1.0), (0, 1, 2.0), (1, 0, 2.0), (1, 1, 1.0)]), O.NATURAL)
    assert e.value.kind == "NotPositiveDefinite" and e.value.pivot == 1
    with pytest.raises(O.SparseError) as e:
        O.Cholesky(O.Matrix.from_triplets(2, 2, [(0, 0, 1.0), (0, 1, 2.0), (1, 0, 3.0), (1, 1, 9.0)]))
    assert e.value.kind == "NotSymmetric"
    with pytest.raises(O.SparseError) as e:
        O.Cholesky(O.Matrix.from_triplets(2, 2, [(0, 0, 1.0)]), O.COLAMD)
    assert e.value.kind == "UnsupportedOrdering"


def assert_lu_reconstructs(a_csc: O.Matrix, lu: O.Lu):
    # factorization_test.rs:698-711
    l, u = lu.l.to_dense(), lu.u.to_dense()
    assert np.all(np.triu(l, 1) == 0.0) and np.all(np.tril(u, -1) == 0.0)
    assert np.all(np.diag(l) == 1.0)
    paq = a_csc.select(lu.row_perm, lu.col_perm).to_dense()
    assert np.max(np.abs(l @ u - paq)) <= 1e-10 * max(1.0, np.max(np.abs(paq)))
    # column storage: L's unit diagonal first, U's diagonal last
    for j in range(lu.n):
        assert lu.l.idx[lu.l.ptr[j]] == j and lu.u.idx[lu.u.ptr[j + 1] - 1] == j


def test_lu_reconstructs_the_permuted_matrix():
    # factorization_test.rs:714-740
    for n, fill, scale, seed in [(1, 0, 1.0, 1), (12, 30, 0.1, 2), (40, 10, 2.0, 3), (60, 6, 0.01, 4)]:
        a = random_square_sparse(n, fill, scale, seed).to_csc()
        for ordering in [O.NATURAL, O.RCM, O.AMD, O.COLAMD]:
            for tau in [0.001, 0.1, 1.0]:
                lu = O.Lu(a, ordering, tau)
                assert list(lu.col_perm) == list(a.ordering(ordering))
                assert is_permutation(lu.row_perm, n)
                assert_lu_reconstructs(a, lu)


def test_lu_matches_textbook_partial_pivoting_under_natural_ordering():
    # factorization_test.rs:743-769 against tests/common/mod.rs:88-165
    for n, fill, scale, seed in [(30, 15, 0.05, 5), (80, 4, 1.5, 6), (120, 3, 0.02, 7)]:
        csr = random_square_sparse(n, fill, scale, seed)
        a = csr.to_csc()
        x = np.array([math.cos(i * 0.37) for i in range(n)])
        b = a.mul_vec(x)
        assert np.allclose(O.Lu(a).solve(b), x, rtol=0, atol=1e-9)
        partial = O.Lu(a, O.NATURAL, 1.0)
        assert list(partial.row_perm) == reference_lu_perm(csr)


def test_lu_threshold_decides_between_diagonal_and_largest():
    # factorization_test.rs:772-790: column 0's candidates are 0.5 (diagonal) and 1.0
    a = O.Matrix.from_triplets(2, 2, [(0, 0, 0.5), (0, 1, 1.0), (1, 0, 1.0), (1, 1, 1.0)], row_major=False)
    assert list(O.Lu(a, O.NATURAL, 0.1).row_perm) == [0, 1]
    assert list(O.Lu(a, O.NATURAL, 0.5).row_perm) == [0, 1]
    assert list(O.Lu(a, O.NATURAL, 0.6).row_perm) == [1, 0]
    assert list(O.Lu(a, O.NATURAL, 1.0).row_perm) == [1, 0]
    with pytest.raises(O.SparseError):
        O.Lu(a, O.NATURAL, 0.0)
    # a singular column names the column of A
    s = O.Matrix.from_triplets(2, 2, [(0, 0, 1.0), (1, 0, 1.0), (0, 1, 0.0)], row_major=False)
    with pytest.raises(O.SparseError) as e:
        O.Lu(s, O.NATURAL)
    assert e.value.kind == "SingularMatrix" and e.value.pivot == 1


def unsymmetric(n, seed):
    """colamd_test.rs:38-56."""
    rng = Lcg(seed)
    t = []
    for row in range(n):
        count = 1 + rng.next() % 6 + (8 if row % 7 == 0 else 0)
        total = 1.0
        for _ in range(count):
            col = rng.next() % n
            if col != row:
                value = (float(rng.next() % 2001) - 1000.0) / 500.0
                total += abs(value)
                t.append((row, col, value))
        t.append((row, row, total))
    return O.Matrix.from_triplets(n, n, t, row_major=False)


def test_lu_under_colamd_solves():
    # colamd_test.rs:129-145
    for n, seed in [(40, 5), (150, 6), (400, 7)]:
        a = unsymmetric(n, seed)
        lu = O.Lu(a, O.COLAMD)
        truth = np.array([math.sin(i * 0.37) for i in range(n)])
        assert np.max(np.abs(lu.solve(a.mul_vec(truth)) - truth)) <= 1e-10


def test_solve_many_is_the_serial_solve_and_par_solve_many_matches_it_bit_for_bit():
    # factorization.rs:52-92, rayon_test.rs:131-175
    a = random_spd(200, 3, 21)
    f = O.Cholesky(a)
    rhs = np.array([[Lcg(s).value() for _ in range(200)] for s in range(40)])
    serial = f.solve_many(rhs, threads=1)
    parallel = f.solve_many(rhs, threads=7)
    assert serial.tobytes() == parallel.tobytes()
    assert serial[3].tobytes() == f.solve(rhs[3]).tobytes()
    d = a.to_dense()
    assert np.max(np.abs(d @ serial.T - rhs.T)) <= 1e-12
    lu = O.Lu(random_square_cuda(150, 4, 22))
    rhs = rhs[:, :150]
    s1, s2 = lu.solve_many(rhs, 1), lu.solve_many(rhs, 5)
    assert s1.tobytes() == s2.tobytes()
    t1, t2 = lu.solve_transpose_many(rhs, 1), lu.solve_transpose_many(rhs, 5)
    assert t1.tobytes() == t2.tobytes()


def test_solve_transpose_solves_with_the_transpose():
    # accuracy_test.rs:91-123
    for n, fill, scale, seed in [(30, 10, 1.5, 1), (80, 4, 0.05, 2), (120, 2, 0.5, 3)]:
        a = random_square_sparse(n, fill, scale, seed).to_csc()
        lu = O.Lu(a, O.AMD)
        b = np.array([math.sin(i * 0.7) for i in range(n)])
        x = lu.solve_transpose(b)
        at = a.to_dense().T
        assert np.max(np.abs(at @ x - b)) <= 1e-10 * max(1.0, np.max(np.abs(b)))


def refine_normwise_py(a, b, max_steps, tol, solve):
    """A Python restatement of accuracy.rs:83-122, to pin the C++ loop's stop rule."""
    x = solve(b)
    r = b - a.mul_vec(x)
    def ratio_of(r):
        rn, bn = np.max(np.abs(r)), np.max(np.abs(b))
        return rn / bn if bn > 0 else (math.inf if rn > 0 else 0.0)
    ratio = ratio_of(r)
    best = (x.copy(), 0, ratio)
    for step in range(1, max_steps + 1):
        if ratio <= tol:
            break
        d = solve(r)
        x = x + d
        previous = ratio
        r = b - a.mul_vec(x)
        ratio = ratio_of(r)
        if ratio < best[2]:
            best = (x.copy(), step, ratio)
        if math.isnan(ratio) or ratio > 0.5 * previous:
            break
    return best


def test_refine_normwise_stop_rule_and_best_iterate():
    # accuracy.rs:83-122; an exact solve needs no step (accuracy_test.rs:226-239)
    a = O.Matrix.from_triplets(3, 3, [(0, 0, 2.0), (1, 1, 4.0), (2, 2, 8.0)], row_major=False)
    lu = O.Lu(a)
    x, steps, ratio = lu.solve_refined_normwise(a, [2.0, 4.0, 8.0], 5, 1e-14)
    assert list(x) == [1.0, 1.0, 1.0] and steps == 0 and ratio == 0.0
    for n, fill, scale, seed in [(150, 3, 0.11, 9), (80, 10, 0.01, 3)]:
        a = random_square_sparse(n, fill, scale, seed).to_csc()
        lu = O.Lu(a)
        b = np.array([1.0 + ((i * 37) % 11) for i in range(n)])
        for max_steps, tol in [(0, 0.0), (10, 0.0), (10, 1e-30), (3, 1e-15)]:
            x, steps, ratio = lu.solve_refined_normwise(a, b, max_steps, tol)
            px, psteps, pratio = refine_normwise_py(a, b, max_steps, tol, lu.solve)
            assert x.tobytes() == px.tobytes() and steps == psteps and ratio == pratio
        x0, steps0, _ = lu.solve_refined_normwise(a, b, 0, 0.0)
        assert steps0 == 0 and x0.tobytes() == lu.solve(b).tobytes()


def test_front_tree_literals():
    # front_tree_test.rs:14-65 (natural ordering, no amalgamation: fundamental supernodes)
    def tree(n, entries):
        a = O.Matrix.from_triplets(n, n, entries)
        return O.Cholesky(a, O.NATURAL, symbolic_only=True).supernodes()[1]
    block = lambda o: [(o, o, 4.0), (o, o + 1, 1.0), (o + 1, o, 1.0), (o + 1, o + 1, 4.0)]
    assert tree(4, block(0) + block(2)) == dict(fronts=2, roots=2, height=1, work=2, critical_work=1,
                                                solve_work=2, solve_critical_work=1)
    star = [(0, 0, 4.0), (1, 1, 4.0), (2, 2, 4.0), (0, 2, 1.0), (2, 0, 1.0), (1, 2, 1.0), (2, 1, 1.0)]
    assert tree(3, star) == dict(fronts=3, roots=1, height=2, work=2, critical_work=1, solve_work=2,
                                 solve_critical_work=1)


def test_supernodes_follow_their_definition():
    # multifrontal.rs:94-138 restated by brute force from the etree and column counts
    for a in [grid_laplacian(9), random_spd(120, 3, 2), graph_matrix(50, [(i, i + 1) for i in range(49)])]:
        for ordering in [O.NATURAL, O.AMD]:
            s = O.Cholesky(a, ordering, symbolic_only=True)
            nodes, _ = s.supernodes()
            n = s.n
            tree = s.elimination_tree()
            counts = list(s.column_counts)
            children = [0] * n
            for p in tree:
                if p is not None:
                    children[p] += 1
            expect, start = [], 0
            for j in range(n):
                cont = j + 1 < n and tree[j] == j + 1 and children[j + 1] == 1 and counts[j] == counts[j + 1] + 1
                if not cont:
                    expect.append((start, j + 1))
                    start = j + 1
            assert [(st, en) for st, en, _, _ in nodes] == expect
            node_of = {c: k for k, (st, en, _, _) in enumerate(nodes) for c in range(st, en)}
            filled = brute_force_factor_pattern(a, s.perm)
            for k, (st, en, below, parent) in enumerate(nodes):
                assert below == [r for r in range(en, n) if filled[r, st]]
                assert parent == (None if tree[en - 1] is None else node_of[tree[en - 1]])
