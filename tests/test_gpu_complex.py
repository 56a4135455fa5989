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
        factor.stage(real_rhs)
    with pytest.raises(Y.InvalidArgument):
        factor.pair_solves([Y.Pair(0, 1)])
    # the complex refined solve needs the complex matrix first
    with pytest.raises(Y.InvalidArgument, match="set_matrix_complex"):
        factor.solve_many_refined(np.zeros((1, chol.n), complex))
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


# ---- FP32 sweeps refined in FP64, and vectors in device memory ----


def _sparse(a):
    import scipy.sparse as sp

    return a if sp.issparse(a) else sp.csr_matrix(a)


def _column_ratios(a, x, b):
    """Each column's ||b - A x||inf / ||b||inf (accuracy.rs:57-78), |z| the hypotenuse."""
    r = b.T - _sparse(a) @ x.T
    return np.max(np.abs(r), axis=0) / np.max(np.abs(b.T), axis=0)


def _ordered_ratios(a, x, b):
    """The same ratios with the residual formed as CsrMatrix::mul_vec forms it
    (product.rs:17-27): each row summed from zero over its entries in ascending column order,
    every complex product (re re - im im, re im + im re) and sum rounded on its own. That is the
    device's order, so only the hypotenuses may round apart."""
    m = _sparse(a).tocsr()
    m.sort_indices()
    ptr, idx, val = m.indptr, m.indices, m.data
    lens = np.diff(ptr)
    xr, xi = np.ascontiguousarray(x.real.T), np.ascontiguousarray(x.imag.T)
    acc_r, acc_i = np.zeros_like(xr), np.zeros_like(xi)
    for k in range(int(lens.max())):
        rows = np.flatnonzero(lens > k)
        at = ptr[rows] + k
        vr, vi, c = val[at].real[:, None], val[at].imag[:, None], idx[at]
        acc_r[rows] = acc_r[rows] + (vr * xr[c] - vi * xi[c])
        acc_i[rows] = acc_i[rows] + (vr * xi[c] + vi * xr[c])
    rn = np.max(np.hypot(b.real.T - acc_r, b.imag.T - acc_i), axis=0)
    return rn / np.max(np.hypot(b.real.T, b.imag.T), axis=0)


@pytest.mark.parametrize("case", ["cholesky", "lu"])
def test_complex_refined_solves_reach_the_fp64_floor(gpu, hermitian, complex_lu, case):
    """FP32 sweeps with FP64 residuals and corrections: every column reaches 1e-12 or, where no
    FP64 solution can, twice the CPU FP64 solution's own residual; the reported ratio is the
    residual of the returned x; and x is within a bound of the FP64 oracle's solution set by
    the residuals (1e-9 relative here)."""
    if case == "cholesky":
        chol, a = hermitian
        factor = Y.GpuFactor.from_cholesky_complex(gpu, chol)
        cpu_solver = chol
    else:
        lu, a = complex_lu
        factor = Y.GpuFactor.from_lu_complex(gpu, lu)
        cpu_solver = lu
    factor.set_matrix(_sparse(a).tocsc() if case == "lu" else _sparse(a))
    b = complex_right_hand_sides(cpu_solver.n, 33, 17)
    x, steps, ratio = factor.solve_many_refined(b, 1e-12, 10)
    assert x.dtype == np.complex128 and x.shape == b.shape
    cpu = cpu_solver.solve_many(b)
    floor = _column_ratios(a, cpu, b)
    got = _column_ratios(a, x, b)
    assert np.all(ratio <= np.maximum(1e-12, 2.0 * floor))
    # the reported ratio is the residual of the returned x, summed in the CPU's order (only the
    # hypotenuses, libm's and the device's, may differ in their last bits)
    ordered = _ordered_ratios(a, x, b)
    assert np.max(np.abs(ratio - ordered) / ordered) <= 1e-14
    assert np.all(got <= np.maximum(1e-12, 2.0 * floor) * 1.5)
    assert np.all(steps >= 1) and np.all(steps <= 10)
    scale = np.max(np.abs(cpu), axis=1)
    assert np.max(np.max(np.abs(x - cpu), axis=1) / scale) <= 1e-9
    # the first FP32 solve alone is far from that: the corrections did the work
    x0, steps0, ratio0 = factor.solve_many_refined(b, 1e-12, 0)
    assert np.all(steps0 == 0) and np.min(ratio0) > np.max(ratio)


def test_complex_refined_solves_follow_a_refactor(gpu):
    """A device refactor of a complex LU rewrites the FP32 copies of the sweeps too: the refined
    solve then converges on the new matrix, to the dense solution of it."""
    from tests.fixtures import random_square_cuda

    class ComplexCsc:
        def __init__(self, ptr, idx, data):
            self.ptr, self.idx, self.data, self.row_major = ptr, idx, np.asarray(data, np.complex128), False

    n, fill, seed = 120, 4, 11  # a case whose pivots pass the threshold (as in the refactor tests)
    lu, a = complex_lu_case(n, fill, seed)
    real = random_square_cuda(n, fill, seed)
    cols = np.repeat(np.arange(n), np.diff(real.ptr))
    moved = ComplexCsc(real.ptr, real.idx, a[real.idx, cols] * (1.0 + 0.01j))  # keeps every pivot
    dense = np.zeros((n, n), complex)
    dense[real.idx, cols] = moved.data
    factor = Y.GpuFactor.from_lu_complex(gpu, lu)
    assert factor.refactor(moved) == "reused"
    factor.set_matrix(moved)
    b = complex_right_hand_sides(n, 9, 3)
    x, _, ratio = factor.solve_many_refined(b, 1e-12, 10)
    assert np.all(ratio <= 1e-10)
    assert np.max(_column_ratios(dense, x, b)) <= 1e-10
    want = np.linalg.solve(dense, b.T).T
    assert np.max(np.abs(x - want)) <= 1e-9 * np.max(np.abs(want))


def test_complex_device_solves_match_the_host_solves(gpu, hermitian, complex_lu):
    """complex128 tensors on the GPU: FP64 and transposed solves give the host calls' bits, with
    contiguous and strided vectors; the refined solve gives the host refined solve's bits."""
    import torch

    for kind in ["cholesky", "lu"]:
        if kind == "cholesky":
            chol, a = hermitian
            factor = Y.GpuFactor.from_cholesky_complex(gpu, chol)
            n = chol.n
        else:
            lu, a = complex_lu
            factor = Y.GpuFactor.from_lu_complex(gpu, lu)
            n = lu.n
        factor.set_matrix(_sparse(a))
        b = complex_right_hand_sides(n, 21, 29)
        tb = torch.tensor(b, dtype=torch.complex128, device="cuda")
        assert factor.solve_device(tb).cpu().numpy().tobytes() == factor.solve_many(b).tobytes()
        got_t = factor.solve_device(tb, transpose=True).cpu().numpy()
        assert got_t.tobytes() == factor.solve_transpose_many(b).tobytes()
        wide_in = torch.zeros((21, n + 5), dtype=torch.complex128, device="cuda")
        wide_in[:, :n] = tb
        wide_out = torch.full((21, n + 9), 7.0 + 1.0j, dtype=torch.complex128, device="cuda")
        factor.solve_device(wide_in[:, :n], wide_out[:, :n])
        assert wide_out[:, :n].cpu().numpy().tobytes() == factor.solve_many(b).tobytes()
        assert torch.all(wide_out[:, n:] == 7.0 + 1.0j)
        refined = factor.solve_device(tb, mode=Y.FP32_REFINED).cpu().numpy()
        host, _, _ = factor.solve_many_refined(b, 1e-12, 10)
        assert refined.tobytes() == host.tobytes(), kind
        factor.set_block_width(4)  # several blocks
        assert factor.solve_device(tb).cpu().numpy().tobytes() == factor.solve_many(b).tobytes()


@pytest.mark.parametrize("name", ["lu", "lu_transpose", "cholesky"])
def test_closed_form_solves_return_the_correctly_rounded_exact_solutions(gpu, name):
    """2 x 2 factors whose solves round only in their last division (num-complex's quotient of
    exact parts is the correctly rounded exact quotient): the device must return the correctly
    rounded exact solutions of Cramer's rule in rational arithmetic, for every right-hand side."""
    from tests.fixtures import gaussian_closed_forms, rounded
    case = gaussian_closed_forms()[name]
    arrays = case["arrays"]
    if name == "cholesky":
        factor = Y.GpuFactor.from_cholesky_complex_arrays(gpu, 2, arrays["l_ptr"], arrays["l_idx"],
                                                          arrays["l_val"], arrays["perm"])
    else:
        factor = Y.GpuFactor.from_lu_complex_arrays(gpu, 2, arrays["l_ptr"], arrays["l_idx"], arrays["l_val"],
                                                    arrays["u_ptr"], arrays["u_idx"], arrays["u_val"],
                                                    arrays["row_perm"], arrays["col_perm"])
    got = factor.solve_transpose_many(case["b"]) if name == "lu_transpose" else factor.solve_many(case["b"])
    want = rounded(case["exact"])
    # equal bit for bit, a zero's sign aside
    assert (got + 0.0).tobytes() == (want + 0.0).tobytes()
