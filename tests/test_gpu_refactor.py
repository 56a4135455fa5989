"""The device LU refactorization (Lu::refactor, lu.rs:662-757) and value updates against the
oracle's CPU refactor: the new factor values bit for bit, then the solves (plain, transposed,
FP32-refined) with those values, on both kernel paths; the reference's refactor cases
(factorization_test.rs:966-1141): pivots that fail the threshold, repeated refactors, pattern
errors leaving the factor unchanged; and a wait given up inside the kernel (a tiny spin limit
set from the host), which must leave the factor unchanged."""
import numpy as np
import pytest

import oracle as O
import ybus_amd as Y
from tests.fixtures import random_square_sparse, right_hand_sides
from tests.gpu_common import gpu, path  # noqa: F401
from tests.test_oracle_refactor import new_values, ring_blocks, with_values

pytestmark = [pytest.mark.gpu]


def same_bits(a, b) -> bool:
    a, b = np.ascontiguousarray(a, dtype=np.float64), np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def cases():
    """(name, matrix, ordering, new matrix) with pivots that survive the new values."""
    ring = ring_blocks(60, 6)
    rand = random_square_sparse(90, 4, 3.0, 2).to_csc()
    grid = O.make_grid(600, 20261207)
    jac = grid.jacobian(20261207)[0]
    return [
        ("ring", ring, O.AMD, new_values(ring, 106)),
        ("random", rand, O.AMD, new_values(rand, 10)),
        ("jacobian", jac, O.COLAMD, grid.jacobian(20261207, 0.003, 5)[0]),
    ]


@pytest.mark.usefixtures("path")
@pytest.mark.parametrize("case", cases(), ids=lambda c: c[0])
def test_refactor_matches_the_cpu_bit_for_bit_and_solves_with_it(gpu, case):
    _, a, ordering, b = case
    lu = O.Lu(a, ordering)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    assert factor.refactor(b) == "reused"
    assert lu.refactor(b) == "reused"
    l_val, u_val = factor.lu_values()
    assert same_bits(l_val, lu.l.data) and same_bits(u_val, lu.u.data)
    n = lu.n
    rhs = right_hand_sides(n, 9, 3)
    assert same_bits(factor.solve_many(rhs), lu.solve_many(rhs))
    assert same_bits(factor.solve_transpose_many(rhs), lu.solve_transpose_many(rhs))
    # FP32 sweeps from the new values, refined in FP64 against the new matrix, to the floor the
    # CPU's FP64 refinement reaches with the same factor (about 1e-11 for the Jacobian)
    factor.set_matrix(b)
    x, steps, ratio = factor.solve_many_refined(rhs, 1e-12, 10)
    floor = np.array([lu.solve_refined_normwise(b, r, 10, 1e-12)[2] for r in rhs])
    assert np.all(ratio <= np.maximum(1e-12, 4.0 * floor)), (ratio, floor)
    stats = factor.refactor_stats()
    assert stats["calls"] == 1 and stats["reused"] == 1
    assert stats["multiply_subtracts"] == lu.refactor_work()


def test_refactor_repeats_bit_for_bit(gpu):
    # factorization_test.rs:1097-1127 on the device: b, then a, then b again
    a = ring_blocks(100, 7)
    b = new_values(a, 107)
    lu = O.Lu(a, O.AMD)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    assert factor.refactor(b) == "reused"
    first = factor.lu_values()
    assert factor.refactor(a) == "reused"
    cpu_a = O.Lu(a, O.AMD)
    assert cpu_a.refactor(a) == "reused"
    l_val, u_val = factor.lu_values()
    assert same_bits(l_val, cpu_a.l.data) and same_bits(u_val, cpu_a.u.data)
    assert factor.refactor(b) == "reused"
    again = factor.lu_values()
    assert same_bits(first[0], again[0]) and same_bits(first[1], again[1])


def test_a_pivot_below_the_threshold_leaves_the_factor_unchanged(gpu):
    # factorization_test.rs:993-1030: the first pivot shrinks below 0.1 of its column's largest
    t = [(0, 0, 4.0), (0, 1, 1.0), (1, 0, 1.0), (1, 1, 4.0), (1, 2, 1.0), (2, 1, 1.0), (2, 2, 4.0)]
    a = O.Matrix.from_triplets(3, 3, t, row_major=False)
    lu = O.Lu(a, O.NATURAL)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    before = factor.lu_values()
    for value in [0.01, 0.0]:
        data = a.data.copy()
        data[0] = value
        b = with_values(a, data)
        assert factor.refactor(b) == "needs_pivoting"
        assert O.Lu(a, O.NATURAL).refactor(b) == "repivoted"
        after = factor.lu_values()
        assert same_bits(before[0], after[0]) and same_bits(before[1], after[1])
        rhs = np.array([[1.0, 2.0, 3.0]])
        assert same_bits(factor.solve_many(rhs), lu.solve_many(rhs))
    stats = factor.refactor_stats()
    assert stats["calls"] == 2 and stats["reused"] == 0
    # the threshold is the caller's: a first pivot of 0.5 against 1 below it passes 0.1, not 1
    data = a.data.copy()
    data[0] = 0.5
    b = with_values(a, data)
    assert factor.refactor(b, 0.1) == "reused"
    kept = factor.lu_values()
    assert factor.refactor(b, 1.0) == "needs_pivoting"
    after = factor.lu_values()
    assert same_bits(kept[0], after[0]) and same_bits(kept[1], after[1])


def test_pattern_errors_leave_the_factor_unchanged(gpu):
    # factorization_test.rs:1048-1065
    a = ring_blocks(20, 3)
    lu = O.Lu(a, O.AMD)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    before = factor.lu_values()
    # an entry the factor's pattern has no place for (pivot step, column step) outside L and U
    pattern = set(zip(lu.l.idx.tolist(), np.repeat(np.arange(lu.n), np.diff(lu.l.ptr)).tolist()))
    pattern |= set(zip(lu.u.idx.tolist(), np.repeat(np.arange(lu.n), np.diff(lu.u.ptr)).tolist()))
    step_of, column_of = np.argsort(lu.row_perm), np.argsort(lu.col_perm)
    dense = a.to_dense()
    outside = next((int(i), int(j)) for i, j in np.argwhere(dense == 0.0)
                   if (int(step_of[i]), int(column_of[j])) not in pattern)
    entries = [(int(i), int(j), float(dense[i, j])) for i, j in np.argwhere(dense != 0.0)]
    grown = O.Matrix.from_triplets(a.rows, a.cols, entries + [(outside[0], outside[1], 1.0)], row_major=False)
    with pytest.raises(Y.PatternMismatch):
        factor.refactor(grown)
    after = factor.lu_values()
    assert same_bits(before[0], after[0]) and same_bits(before[1], after[1])
    # a good call fixes the pattern; a smaller one (inside the factor's) is then refused too
    assert factor.refactor(new_values(a, 3)) == "reused"
    fixed = factor.lu_values()
    fewer = O.Matrix.from_triplets(a.rows, a.cols, [e for e in entries if e[0] != e[1] or e[0] > 0],
                                   row_major=False)
    with pytest.raises(Y.PatternMismatch):
        factor.refactor(fewer)
    with pytest.raises(Y.InvalidArgument):
        factor.refactor(a, 0.0)
    after = factor.lu_values()
    assert same_bits(fixed[0], after[0]) and same_bits(fixed[1], after[1])
    # a Cholesky factor has no LU values
    spd = O.Cholesky(O.Matrix.from_triplets(2, 2, [(0, 0, 2.0), (1, 1, 3.0)]))
    chol = Y.GpuFactor.from_cholesky(gpu, spd)
    with pytest.raises(Y.InvalidArgument):
        chol.refactor(a)


@pytest.mark.usefixtures("path")
def test_value_updates_from_a_host_refactor(gpu):
    grid = O.make_grid(800, 20261208)
    a = grid.jacobian(20261208)[0]
    lu = O.Lu(a, O.COLAMD)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    rhs = right_hand_sides(lu.n, 5, 8)
    for step in range(3):
        b = grid.jacobian(20261208, 0.002, 30 + step)[0]
        assert lu.refactor(b) == "reused"
        factor.update_lu_values(lu.l.data, lu.u.data)
        assert same_bits(factor.solve_many(rhs), lu.solve_many(rhs))
        assert same_bits(factor.solve_transpose_many(rhs), lu.solve_transpose_many(rhs))
    l_val, u_val = factor.lu_values()
    assert same_bits(l_val, lu.l.data) and same_bits(u_val, lu.u.data)
    assert factor.refactor_stats()["updates"] == 3


def test_the_jacobian_of_25k_buses_refactors_bit_for_bit(gpu):
    grid = O.make_grid(25000, 20261205)
    a = grid.jacobian(20261205)[0]
    b = grid.jacobian(20261205, 0.001, 7)[0]
    lu = O.Lu(a, O.COLAMD)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    for round_ in range(2):
        assert factor.refactor(b) == "reused"
        assert factor.refactor(a) == "reused"
    assert factor.refactor(b) == "reused"
    assert lu.refactor(b) == "reused"
    l_val, u_val = factor.lu_values()
    assert same_bits(l_val, lu.l.data) and same_bits(u_val, lu.u.data)
    rhs = right_hand_sides(lu.n, 4, 2)
    assert same_bits(factor.solve_many(rhs), lu.solve_many(rhs))


def test_a_wait_given_up_leaves_the_factor_unchanged(gpu, monkeypatch):
    # a chain: column k of a tridiagonal matrix reads column k - 1, so with a one-poll spin limit
    # some wait is given up (the kernel still ends: every column sets its flag); the factor keeps
    # its values either way
    n = 3000
    t = [(i, i, 4.0) for i in range(n)] + [(i, i + 1, -1.0) for i in range(n - 1)] + \
        [(i + 1, i, -1.0) for i in range(n - 1)]
    a = O.Matrix.from_triplets(n, n, t, row_major=False)
    lu = O.Lu(a, O.NATURAL)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    before = factor.lu_values()
    b = new_values(a, 5)
    monkeypatch.setenv("YBUS_HIP_REFACTOR_SPIN", "1")
    try:
        outcome = factor.refactor(b)
    except Y.GpuError as e:
        assert e.status == Y.INTERNAL
        outcome = None
    if outcome is None:
        after = factor.lu_values()
        assert same_bits(before[0], after[0]) and same_bits(before[1], after[1])
    monkeypatch.delenv("YBUS_HIP_REFACTOR_SPIN")
    assert factor.refactor(b) == "reused"
    assert lu.refactor(b) == "reused"
    l_val, u_val = factor.lu_values()
    assert same_bits(l_val, lu.l.data) and same_bits(u_val, lu.u.data)


def test_an_attached_matrix_follows_the_refactor(gpu):
    # the matrix kept for FP32-refined solves (set_matrix) takes the refactored matrix's values
    # when it is that matrix (CSC, the same pattern); the device's residual of each refined
    # solution is then the CPU's residual against the new matrix, bit for bit
    grid = O.make_grid(600, 20261209)
    a = grid.jacobian(20261209)[0]
    b = grid.jacobian(20261209, 0.002, 4)[0]
    lu = O.Lu(a, O.COLAMD)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    factor.set_matrix(a)
    rhs = right_hand_sides(lu.n, 6, 12)
    assert factor.refactor(b) == "reused"
    x, steps, ratio = factor.solve_many_refined(rhs, 1e-12, 10)
    for k in range(rhs.shape[0]):
        assert ratio[k] == b.relative_residual(x[k], rhs[k])
    # a matrix attached in another orientation is left as it was
    other = Y.GpuFactor.from_lu(gpu, lu)
    other.set_matrix(a.to_csr())
    assert other.refactor(b) == "reused"
    x, steps, ratio = other.solve_many_refined(rhs, 1e-12, 10)
    for k in range(rhs.shape[0]):
        assert ratio[k] == a.relative_residual(x[k], rhs[k])


def arrow(n: int, seed: int) -> O.Matrix:
    """A dominant diagonal with a full first row and column: under natural ordering L and U fill
    completely, so L's first columns hold up to n - 1 entries below the diagonal (CSC)."""
    rng = np.random.default_rng(seed)
    t = [(i, i, 4.0 + n * 0.1 + rng.random()) for i in range(n)]
    t += [(0, j, 0.5 + rng.random()) for j in range(1, n)] + [(i, 0, 0.5 + rng.random()) for i in range(1, n)]
    return O.Matrix.from_triplets(n, n, t, row_major=False)


def test_columns_of_l_longer_than_a_wavefront(gpu):
    # steps with 65 to 149 entries (a wavefront's 64 lanes take them in more than one round) and
    # columns of thousands of entries (many batches)
    a = arrow(150, 1)
    lu = O.Lu(a, O.NATURAL)
    assert max(np.diff(lu.l.ptr)) - 1 == 149
    factor = Y.GpuFactor.from_lu(gpu, lu)
    b = new_values(a, 4)
    assert factor.refactor(b) == "reused"
    assert lu.refactor(b) == "reused"
    l_val, u_val = factor.lu_values()
    assert same_bits(l_val, lu.l.data) and same_bits(u_val, lu.u.data)
    rhs = right_hand_sides(150, 3, 1)
    assert same_bits(factor.solve_many(rhs), lu.solve_many(rhs))


def test_a_column_beyond_the_stage_is_refused(gpu):
    # L's first column has 599 entries below its diagonal, more than one batch stages (512): the
    # device refactor refuses the factor, which keeps its values and solves
    a = arrow(600, 2)
    lu = O.Lu(a, O.NATURAL)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    with pytest.raises(Y.TooLarge):
        factor.refactor(new_values(a, 5))
    rhs = right_hand_sides(600, 2, 3)
    assert same_bits(factor.solve_many(rhs), lu.solve_many(rhs))
    # a host refactor's values still go in
    b = new_values(a, 5)
    assert lu.refactor(b) == "reused"
    factor.update_lu_values(lu.l.data, lu.u.data)
    assert same_bits(factor.solve_many(rhs), lu.solve_many(rhs))


def test_one_unknown_and_a_diagonal(gpu):
    one = O.Matrix.from_triplets(1, 1, [(0, 0, 3.0)], row_major=False)
    lu = O.Lu(one, O.NATURAL)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    assert factor.refactor(O.Matrix.from_triplets(1, 1, [(0, 0, -2.0)], row_major=False)) == "reused"
    assert list(factor.lu_values()[1]) == [-2.0]
    assert factor.refactor(O.Matrix.from_triplets(1, 1, [(0, 0, 0.0)], row_major=False)) == "needs_pivoting"
    assert list(factor.lu_values()[1]) == [-2.0]
    d = O.Matrix.from_triplets(5, 5, [(i, i, 1.0 + i) for i in range(5)], row_major=False)
    lu = O.Lu(d, O.NATURAL)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    b = new_values(d, 2)
    assert factor.refactor(b) == "reused"
    assert lu.refactor(b) == "reused"
    assert same_bits(factor.lu_values()[1], lu.u.data)


# ---- complex LU factors (the reference's refactor is generic over the element type) ----


class ComplexCsc:
    """A complex CSC matrix as the refactor takes it: ptr, idx and complex data."""

    def __init__(self, ptr, idx, data):
        self.ptr, self.idx, self.data = ptr, idx, np.asarray(data, dtype=np.complex128)
        self.row_major = False


def complex_case(n: int, fill: int, seed: int):
    from tests.fixtures import complex_lu_case, random_square_cuda

    lu, a = complex_lu_case(n, fill, seed)
    real = random_square_cuda(n, fill, seed)
    rng = np.random.default_rng(seed)
    scale = 1.0 + 0.02 * (rng.random(real.nnz) - 0.5) + 0.01j * (rng.random(real.nnz) - 0.5)
    cols = np.repeat(np.arange(n), np.diff(real.ptr))
    return lu, ComplexCsc(real.ptr, real.idx, a[real.idx, cols] * scale)


def complex_jacobian_case(buses: int, seed: int):
    """A Newton Jacobian made complex (each value v as v (1 - 0.5 i)) over its real COLAMD LU, and
    the same pattern with every value turned by 0.9 + 0.05 i."""
    from oracle import complex_solve as Z

    jac = O.make_grid(buses, seed).jacobian(seed)[0]
    lu = Z.ComplexLu(jac.to_dense() * (1.0 - 0.5j), O.Lu(jac, O.COLAMD))
    return lu, ComplexCsc(jac.ptr, jac.idx, jac.data * (1.0 - 0.5j) * (0.9 + 0.05j))


@pytest.mark.parametrize("case", ["random-40", "random-120", "jacobian"])
def test_complex_refactor_matches_the_cpu_bit_for_bit(gpu, case):
    from oracle import complex_solve as Z
    from tests.fixtures import complex_right_hand_sides

    if case == "random-40":
        lu, b = complex_case(40, 8, 3)
    elif case == "random-120":
        lu, b = complex_case(120, 4, 11)
    else:
        lu, b = complex_jacobian_case(150, 20261212)
    n, seed = lu.n, 7
    factor = Y.GpuFactor.from_lu_complex(gpu, lu)
    assert factor.refactor(b) == "reused"
    assert Z.lu_refactor(lu, b.ptr, b.idx, b.data.real.copy(), b.data.imag.copy())
    l_val, u_val = factor.lu_values()
    assert same_bits(l_val.view(np.float64), lu.l_val.view(np.float64))
    assert same_bits(u_val.view(np.float64), lu.u_val.view(np.float64))
    rhs = complex_right_hand_sides(n, 4, seed)
    assert same_bits(factor.solve_many(rhs).view(np.float64), lu.solve_many(rhs).view(np.float64))
    assert same_bits(factor.solve_transpose_many(rhs).view(np.float64), lu.solve_transpose_many(rhs).view(np.float64))


def test_complex_refactor_of_a_grid_laplacian(gpu):
    # factorization_test.rs:1032-1046 on the device: a = (v, -2v), then b = (0.7v, -1.1v)
    from oracle import complex_solve as Z
    from tests.fixtures import grid_laplacian

    real = grid_laplacian(5).to_csc()
    lu = Z.ComplexLu(real.to_dense() * (1.0 - 2.0j), O.Lu(real, O.AMD))
    factor = Y.GpuFactor.from_lu_complex(gpu, lu)
    b = ComplexCsc(real.ptr, real.idx, real.data * (0.7 - 1.1j))
    assert factor.refactor(b) == "reused"
    x = np.array([complex(i, 1.0) for i in range(25)])
    solved = factor.solve_many(((real.to_dense() * (0.7 - 1.1j)) @ x)[None, :])[0]
    assert np.abs(solved - x).max() < 1e-10


def test_complex_refactor_refuses_a_small_pivot_and_the_other_kind(gpu):
    from oracle import complex_solve as Z

    real = O.Matrix.from_triplets(2, 2, [(0, 0, 2.0), (0, 1, 1.0), (1, 0, 1.0), (1, 1, 3.0)], row_major=False)
    lu = Z.ComplexLu(real.to_dense().astype(complex), O.Lu(real, O.NATURAL))
    factor = Y.GpuFactor.from_lu_complex(gpu, lu)
    before = factor.lu_values()
    small = ComplexCsc(real.ptr, real.idx, [0.01 + 0.01j, 1.0, 1.0, 3.0])
    assert factor.refactor(small) == "needs_pivoting"
    after = factor.lu_values()
    assert same_bits(before[0].view(np.float64), after[0].view(np.float64))
    assert same_bits(before[1].view(np.float64), after[1].view(np.float64))
    exact = ComplexCsc(real.ptr, real.idx, [2j, 1.0, 1.0, 3.0])
    assert factor.refactor(exact) == "reused"
    assert list(factor.lu_values()[0]) == [1.0, -0.5j, 1.0]
    assert list(factor.lu_values()[1]) == [2j, 1.0, 3 + 0.5j]
    # a real factor takes the real entry points only, and the reverse
    real_factor = Y.GpuFactor.from_lu(gpu, O.Lu(real, O.NATURAL))
    ones = np.ones(8)
    assert Y.library().ybus_hip_factor_update_lu_values_complex(real_factor._h, ones, ones) == Y.INVALID_ARGUMENT
    assert Y.library().ybus_hip_factor_update_lu_values(factor._h, ones, ones) == Y.INVALID_ARGUMENT
