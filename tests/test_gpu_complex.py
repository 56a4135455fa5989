"""Complex GPU solves against the complex oracle's CPU solves (Cholesky<Complex64> and
Lu<Complex64>, solve and transposed solve). The kernels keep the CPU's operation order and
num-complex's rounding of each product and quotient, so the solutions are checked bit for bit,
and against A itself by ||b - A x||inf / ||b||inf (accuracy.rs:57-78)."""
import numpy as np
import pytest

import ybus_amd as Y
from tests.fixtures import complex_lu_case, complex_relative_residual, complex_right_hand_sides, hermitian_case
from tests.gpu_common import gpu  # noqa: F401

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hermitian():
    return hermitian_case(300, 7)


@pytest.fixture(scope="module")
def complex_lu():
    return complex_lu_case(250, 4, 3)


def test_complex_cholesky_solves_match_the_cpu_at_every_block_width(gpu, hermitian):
    chol, a = hermitian
    b = complex_right_hand_sides(chol.n, 37, 11)
    cpu = chol.solve_many(b)
    assert complex_relative_residual(a, cpu, b) <= 1e-12
    factor = Y.GpuFactor.from_cholesky_complex(gpu, chol)
    assert factor.is_complex()
    for width in [1, 5, 37, 64, 1024]:
        factor.set_block_width(width)
        got = factor.solve_many(b)
        assert got.dtype == np.complex128 and got.shape == b.shape
        assert got.tobytes() == cpu.tobytes(), width


def test_complex_cholesky_transposed_solves_are_the_conjugate_identity(gpu, hermitian):
    chol, a = hermitian
    b = complex_right_hand_sides(chol.n, 20, 3)
    factor = Y.GpuFactor.from_cholesky_complex(gpu, chol)
    got = factor.solve_transpose_many(b)
    assert got.tobytes() == chol.solve_transpose_many(b).tobytes()
    assert complex_relative_residual(a.T, got, b) <= 1e-12


def test_complex_lu_solves_and_transposed_solves_match_the_cpu(gpu, complex_lu):
    lu, a = complex_lu
    b = complex_right_hand_sides(lu.n, 20, 5)
    factor = Y.GpuFactor.from_lu_complex(gpu, lu)
    factor.set_block_width(8)
    got = factor.solve_many(b)
    assert got.tobytes() == lu.solve_many(b).tobytes()
    assert complex_relative_residual(a, got, b) <= 1e-10
    got_t = factor.solve_transpose_many(b)
    assert got_t.tobytes() == lu.solve_transpose_many(b).tobytes()
    assert complex_relative_residual(a.T, got_t, b) <= 1e-10


@pytest.mark.parametrize("tile", ["1", "8", "64"])
def test_every_tile_width_gives_the_same_bits(gpu, hermitian, monkeypatch, tile):
    chol, _ = hermitian
    monkeypatch.setenv("YBUS_HIP_TILE", tile)
    b = complex_right_hand_sides(chol.n, 70, 9)
    factor = Y.GpuFactor.from_cholesky_complex(gpu, chol)
    assert factor.solve_many(b).tobytes() == chol.solve_many(b).tobytes()


def test_a_larger_grid_with_many_right_hand_sides(gpu):
    chol, a = hermitian_case(2000, 5)
    b = complex_right_hand_sides(chol.n, 300, 13)
    factor = Y.GpuFactor.from_cholesky_complex(gpu, chol)
    got = factor.solve_many(b)
    assert got.tobytes() == chol.solve_many(b).tobytes()
    assert complex_relative_residual(a, got, b) <= 1e-12
    stats = factor.kernel_stats()
    assert stats.launches >= 1 and stats.milliseconds > 0


def test_empty_and_mismatched_right_hand_sides(gpu, hermitian):
    chol, _ = hermitian
    factor = Y.GpuFactor.from_cholesky_complex(gpu, chol)
    assert factor.solve_many(np.zeros((0, chol.n), complex)).shape == (0, chol.n)
    with pytest.raises(Y.ShapeMismatch):
        factor.solve_many(np.zeros((2, chol.n + 1), complex))


def test_real_entry_points_reject_a_complex_factor_and_the_reverse(gpu, hermitian):
    chol, a = hermitian
    factor = Y.GpuFactor.from_cholesky_complex(gpu, chol)
    real_rhs = np.zeros((1, chol.n))
    with pytest.raises(Y.InvalidArgument):
        factor.solve_many_refined(real_rhs)
    with pytest.raises(Y.InvalidArgument):
        factor.pair_solves([Y.Pair(0, 1)])
    with pytest.raises(Y.InvalidArgument):
        Y.GpuBatch.new(gpu).add(factor)
    # a real factor refuses the complex solves
    real = Y.GpuFactor.from_cholesky_arrays(gpu, chol.n, chol.l_ptr, chol.l_idx, chol.l_re, chol.perm)
    assert not real.is_complex()
    b = np.zeros((1, 2 * chol.n))
    out = np.zeros_like(b)
    assert Y.library().ybus_hip_factor_solve_many_complex(real._h, b, 1, out) != 0


def test_one_unknown_and_a_pivoting_two_by_two(gpu):
    """Closed forms: a 1 x 1 Hermitian matrix (x = b / a with a real), and a 2 x 2 complex LU whose
    row permutation swaps the rows (A = P^T L U)."""
    from oracle import complex_solve as Z

    class One:
        n = 1
        perm = np.array([0])
        l_ptr = np.array([0, 1])
        l_idx = np.array([0])
        l_re = np.array([2.0])
        l_im = np.array([0.0])

        @property
        def l(self):
            return Z.Compressed(self.l_ptr, self.l_idx, np.array([2.0 + 0.0j]))

    factor = Y.GpuFactor.from_cholesky_complex(gpu, One())
    b = np.array([[4.0 + 8.0j], [-1.0 + 0.5j]])
    # A = L L^H = 4: x = b / 4 (num-complex divides twice by the diagonal 2, exactly)
    assert np.array_equal(factor.solve_many(b), b / 4.0)
    # P A Q = L U with P swapping the rows: A = [[3 + i, 1], [1 - 2i, 2]] has rows (1, 0) pivoted
    l_val = np.array([1.0 + 0.0j, (3.0 + 1.0j) / (1.0 - 2.0j), 1.0 + 0.0j])
    u_val = np.array([1.0 - 2.0j, 2.0 + 0.0j, 1.0 - (3.0 + 1.0j) / (1.0 - 2.0j) * 2.0])
    lu = Y.GpuFactor.from_lu_complex_arrays(gpu, 2, [0, 2, 3], [0, 1, 1], l_val, [0, 1, 3], [0, 0, 1], u_val,
                                            [1, 0], [0, 1])
    a = np.array([[3.0 + 1.0j, 1.0], [1.0 - 2.0j, 2.0]])
    rhs = np.array([[1.0 + 1.0j, -2.0 + 0.5j], [0.25j, 3.0]])
    x = lu.solve_many(rhs)
    assert np.max(np.abs(a @ x.T - rhs.T)) <= 1e-14
    xt = lu.solve_transpose_many(rhs)
    assert np.max(np.abs(a.T @ xt.T - rhs.T)) <= 1e-14
