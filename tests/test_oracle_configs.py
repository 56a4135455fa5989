"""The configurations' workload builders and the oracle's 1-norm condition estimate.

The estimate restates accuracy.rs:207-279 and is pinned by the reference's own test of it
(accuracy_test.rs:125-169): a lower bound on the exact ||A^-1||_1 (every column solved), within
a factor of 3, exact on at least 80% of 120 random SPD and pivoting LU matrices, exact on a
diagonal matrix, and kappa_1 of the 6 x 6 Hilbert matrix between 1e7 and 1e8."""
import numpy as np
import pytest

import oracle as O
from oracle import configs
from tests.fixtures import random_spd, random_square_sparse


def exact_inverse_norm_1(n, solve):
    return max(float(np.sum(np.abs(solve(np.eye(n)[j])))) for j in range(n))


def norm_1(a):
    return float(np.max(np.sum(np.abs(a.to_dense()), axis=0)))


def test_the_condition_estimate_is_a_lower_bound_and_usually_exact():
    tried = exact = 0
    for seed in range(20):
        for n, fill in [(10, 30), (40, 8), (100, 3)]:
            chol = O.Cholesky(random_spd(n, fill, seed))
            truth = exact_inverse_norm_1(n, chol.solve)
            estimate = configs.inverse_norm_1_estimate(n, chol.solve, chol.solve)
            assert truth / 3.0 <= estimate <= truth * (1.0 + 1e-12)
            tried += 1
            exact += estimate >= truth * (1.0 - 1e-12)
            lu = O.Lu(random_square_sparse(n, fill, 0.2, seed + 100).to_csc())
            truth = exact_inverse_norm_1(n, lu.solve)
            estimate = configs.inverse_norm_1_estimate(n, lu.solve, lu.solve_transpose)
            assert truth / 3.0 <= estimate <= truth * (1.0 + 1e-12)
            tried += 1
            exact += estimate >= truth * (1.0 - 1e-12)
    assert exact * 10 >= tried * 8, (exact, tried)


def test_the_condition_estimate_on_closed_forms():
    diagonal = O.Matrix.from_triplets(3, 3, [(0, 0, 2.0), (1, 1, -0.5), (2, 2, 4.0)])
    lu = O.Lu(diagonal.to_csc())
    assert configs.inverse_norm_1_estimate(3, lu.solve, lu.solve_transpose) == 2.0
    assert norm_1(diagonal) * 2.0 == 8.0
    h = O.Matrix.from_triplets(6, 6, [(i, j, 1.0 / (i + j + 1)) for i in range(6) for j in range(6)])
    chol = O.Cholesky(h)
    kappa = norm_1(h) * configs.inverse_norm_1_estimate(6, chol.solve, chol.solve)
    truth = norm_1(h) * exact_inverse_norm_1(6, chol.solve)
    assert truth / 3.0 <= kappa <= truth * (1.0 + 1e-8)
    assert 1e7 < kappa < 1e8


def test_the_infinity_norm_estimate_swaps_the_solves():
    a = random_square_sparse(30, 10, 0.2, 7)
    lu = O.Lu(a.to_csc())
    truth = float(np.max(np.sum(np.abs(np.linalg.inv(a.to_dense())), axis=1)))
    estimate = configs.lu_inverse_norm_inf_estimate(lu)
    assert truth / 3.0 <= estimate <= truth * (1.0 + 1e-10)
    assert configs.norm_inf(a) == float(np.max(np.sum(np.abs(a.to_dense()), axis=1)))
    assert configs.norm_inf(a.to_csc()) == configs.norm_inf(a)
    assert configs.norm_1(a) == float(np.max(np.sum(np.abs(a.to_dense()), axis=0)))
    assert configs.norm_1(a.to_csc()) == configs.norm_1(a)


def test_dense_pairs_and_pair_lists():
    grid = O.make_grid(50, 3)
    plus, minus = configs.pair_lists(grid)
    assert len(plus) == grid.branches
    assert np.all((plus >= -1) & (plus < grid.buses - 1)) and np.all(plus != minus)
    dense = configs.dense_pairs(plus, minus, grid.buses - 1, 2, 9)
    for k in range(7):
        want = np.zeros(grid.buses - 1)
        if plus[2 + k] >= 0:
            want[plus[2 + k]] = 1.0
        if minus[2 + k] >= 0:
            want[minus[2 + k]] -= 1.0
        assert np.array_equal(dense[k], want)


@pytest.mark.slow
def test_the_configurations_are_the_ones_measured():
    """The builders give the workloads the measurements were quoted on (sizes and factor
    entries), so a change to a generator shows here before it changes a number."""
    _, bprime, chol, plus, minus = configs.ptdf()
    assert (chol.n, int(chol.l.nnz), len(plus)) == (9999, 27835, 13500)
    jac, lu, rhs = configs.jacobian()
    assert (jac.rows, int(lu.l.nnz + lu.u.nnz), rhs.shape) == (44998, 544539, (512, 44998))
    case = configs.adjoint()
    assert (case.chol.n, int(case.chol.l.nnz), case.bprime.nnz, case.p.shape) == (69999, 195210, 258991,
                                                                                  (64, 69999))


def test_supernode_widths_restate_the_cholesky_supernodes():
    """configs.supernode_widths on L's columns gives the pinned supernodes() widths."""
    for buses, seed in ((400, 3), (2000, 11)):
        chol = O.Cholesky(O.make_grid(buses, seed).dc()[0])
        csc = chol.l.to_csc()
        widths, entries = configs.supernode_widths(chol.n, csc.ptr, csc.idx)
        want = [end - start for start, end, _, _ in chol.supernodes()[0]]
        assert widths.tolist() == want
        assert int(entries.sum()) == int(chol.l.nnz)


@pytest.mark.slow
def test_the_factors_supernodes_are_too_narrow_for_dense_updates():
    """The basis of the choice not to build supernodal MFMA updates: in every configuration's
    factor the supernodes of eight or more columns hold at most 3.1 % of the stored entries (0 %
    for the grids' Cholesky factors), and none is wider than 16 columns."""
    def shares(n, ptr, idx):
        widths, entries = configs.supernode_widths(n, ptr, idx)
        return int(widths.max()), entries[widths >= 8].sum() / entries.sum()

    _, _, chol, _, _ = configs.ptdf()
    csc = chol.l.to_csc()
    assert shares(chol.n, csc.ptr, csc.idx) == (6, 0.0)
    _, lu, _ = configs.jacobian(count=1)
    widest, share = shares(lu.n, lu.l.ptr, lu.l.idx)
    assert widest == 12 and 0.025 < share < 0.027
    ut = lu.u.to_csr()
    widest, share = shares(lu.n, ut.ptr, ut.idx)
    assert widest == 16 and 0.030 < share < 0.032
