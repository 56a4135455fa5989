"""The adjoint of A x = b on the GPU: b_bar = A^-T x_bar and a_bar = -(sum of b_bar_s x_s^T) on A's
pattern. The reference has no GPU gradient; the checks are the CPU composition of the oracle's
transposed solve with the outer product (bit for bit), closed forms on 1x1 and 2x2 systems, and
central finite differences of a flow-limit loss on a synthetic DC grid, for both the injections
and A's stored values."""
import numpy as np
import pytest

import oracle as O
import ybus_amd as Y
from tests.fixtures import random_spd, random_square_cuda, right_hand_sides
from tests.gpu_common import gpu, path  # noqa: F401

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("path")]


def test_one_by_one_closed_form(gpu):
    a = O.Matrix.from_triplets(1, 1, [(0, 0, 4.0)], row_major=False)
    factor = Y.GpuFactor.from_lu(gpu, O.Lu(a))
    factor.set_matrix(a)
    b_bar, a_bar = factor.adjoint([[3.0]], [[0.5]])
    assert b_bar[0, 0] == 3.0 / 4.0
    assert a_bar[0] == -(0.75 * 0.5)


def test_two_by_two_against_dense_algebra(gpu):
    # unsymmetric, so a transposed solve or a swapped outer product would show
    a = O.Matrix.from_triplets(2, 2, [(0, 0, 2.0), (0, 1, 1.0), (1, 0, -3.0), (1, 1, 5.0)], row_major=True)
    factor = Y.GpuFactor.from_lu(gpu, O.Lu(a.to_csc()))
    factor.set_matrix(a)
    x_bar = np.array([[1.0, -2.0], [0.25, 4.0]])
    x = np.array([[3.0, 7.0], [-1.0, 0.5]])
    b_bar, a_bar = factor.adjoint(x_bar, x)
    dense = a.to_dense()
    want_b = np.linalg.solve(dense.T, x_bar.T).T
    assert np.allclose(b_bar, want_b, rtol=1e-14, atol=0)
    want_a = -(want_b.T @ x)  # (i, j) = -sum_s b_bar_s[i] x_s[j]
    got = np.zeros((2, 2))
    for p, (i, j) in enumerate([(0, 0), (0, 1), (1, 0), (1, 1)]):
        got[i, j] = a_bar[p]
    assert np.allclose(got, want_a, rtol=1e-14, atol=0)


@pytest.mark.parametrize("kind", ["lu", "cholesky"])
def test_adjoint_matches_the_cpu_composition_bit_for_bit(gpu, kind):
    if kind == "lu":
        a = random_square_cuda(180, 4, 21)
        cpu_factor = O.Lu(a)
        factor = Y.GpuFactor.from_lu(gpu, cpu_factor)
    else:
        a = random_spd(220, 3, 22)
        cpu_factor = O.Cholesky(a)
        factor = Y.GpuFactor.from_cholesky(gpu, cpu_factor)
    factor.set_matrix(a)
    x_bar = right_hand_sides(a.rows, 45, 5)
    x = right_hand_sides(a.rows, 45, 6)
    want_b, want_a = O.adjoint(cpu_factor, a, x_bar, x)
    for width in [45, 16, 7]:  # blocks of columns carry the sums from one to the next
        factor.set_block_width(width)
        b_bar, a_bar = factor.adjoint(x_bar, x)
        assert b_bar.tobytes() == want_b.tobytes(), width
        assert a_bar.tobytes() == want_a.tobytes(), width


def test_adjoint_needs_the_matrix_and_matching_shapes(gpu):
    a = random_spd(30, 10, 3)
    factor = Y.GpuFactor.from_cholesky(gpu, O.Cholesky(a))
    with pytest.raises(Y.InvalidArgument):
        factor.adjoint(np.ones((2, 30)), np.ones((2, 30)))
    factor.set_matrix(a)
    with pytest.raises(Y.ShapeMismatch):
        factor.adjoint(np.ones((2, 30)), np.ones((3, 30)))
    b_bar, a_bar = factor.adjoint(np.zeros((0, 30)), np.zeros((0, 30)))
    assert b_bar.shape == (0, 30) and np.all(a_bar == 0.0)


def flow_loss(a: O.Matrix, bf: np.ndarray, p: np.ndarray, limit: np.ndarray):
    """sum over scenarios and branches of relu(|f| - limit)^2, f = Bf A^-1 p (by LU on the CPU),
    and its gradient with respect to the solutions, x_bar = Bf^T (2 relu(|f| - limit) sign(f))."""
    lu = O.Lu(a.to_csc(), O.AMD)
    x = lu.solve_many(p)
    f = x @ bf.T
    over = np.maximum(np.abs(f) - limit, 0.0)
    x_bar = (2.0 * over * np.sign(f)) @ bf
    return float(np.sum(over * over)), x, x_bar


@pytest.mark.parametrize("kind", ["lu", "cholesky"])
def test_refined_adjoint_agrees_with_the_fp64_adjoint(gpu, kind):
    """FP32 transposed sweeps refined in FP64 against A^T: b_bar reaches the transposed system's
    FP64 floor, and both gradients agree with the FP64 adjoint to 1e-8 relative (the bar the
    configuration sets for the refined gradients)."""
    if kind == "lu":
        a, _ = O.make_grid(400, 8).jacobian(8)
        cpu_factor = O.Lu(a, O.COLAMD)
        factor = Y.GpuFactor.from_lu(gpu, cpu_factor)
    else:
        a, _, _ = O.make_grid(900, 9).dc()
        cpu_factor = O.Cholesky(a)
        factor = Y.GpuFactor.from_cholesky(gpu, cpu_factor)
    factor.set_matrix(a)
    x_bar = right_hand_sides(a.rows, 12, 3)
    x = right_hand_sides(a.rows, 12, 4)
    b64, a64 = factor.adjoint(x_bar, x)
    b32, a32 = factor.adjoint(x_bar, x, Y.FP32_REFINED, 1e-12, 10)
    assert np.max(np.abs(b32 - b64)) <= 1e-8 * np.max(np.abs(b64))
    assert np.max(np.abs(a32 - a64)) <= 1e-8 * np.max(np.abs(a64))
    at = a.to_dense().T
    for k in range(12):
        floor = np.max(np.abs(x_bar[k] - at @ b64[k])) / np.max(np.abs(x_bar[k]))
        ratio = np.max(np.abs(x_bar[k] - at @ b32[k])) / np.max(np.abs(x_bar[k]))
        assert ratio <= max(1e-12, 4.0 * floor), (k, ratio, floor)


@pytest.mark.parametrize("mode", ["fp64", "fp32_refined"])
def test_gradients_match_central_finite_differences(gpu, mode):
    g = O.make_grid(300, 20261207)
    bp, bf, _ = g.dc()
    bf = bf.to_dense() if hasattr(bf, "to_dense") else bf
    p = O.injections(20261207, 4, 300, 6)
    _, x0, _ = flow_loss(bp, bf, p, np.zeros(bf.shape[0]))
    limit = np.quantile(np.abs(x0 @ bf.T), 0.6, axis=0)  # about 40% of the flows over their limit
    loss, x, x_bar = flow_loss(bp, bf, p, limit)
    assert loss > 0.0
    factor = Y.GpuFactor.from_cholesky(gpu, O.Cholesky(bp))
    factor.set_matrix(bp)
    b_bar, a_bar = factor.adjoint(x_bar, x, Y.FP64 if mode == "fp64" else Y.FP32_REFINED, 1e-12, 10)
    h = 1e-6
    # a central difference carries the loss's own rounding, a few units of 1e-16 * loss each, over
    # 2h (the loss is piecewise quadratic in p, so that is all of its error there)
    noise = 1e-8 * loss
    rng = np.random.default_rng(7)
    # the injections: dL/dp_s = b_bar_s
    for _ in range(40):
        s, i = int(rng.integers(p.shape[0])), int(rng.integers(p.shape[1]))
        plus, minus = p.copy(), p.copy()
        plus[s, i] += h
        minus[s, i] -= h
        fd = (flow_loss(bp, bf, plus, limit)[0] - flow_loss(bp, bf, minus, limit)[0]) / (2 * h)
        assert abs(fd - b_bar[s, i]) <= 1e-6 * abs(b_bar[s, i]) + noise, (s, i, fd, b_bar[s, i], loss)
    # A's stored values, one at a time (so the perturbed matrix is unsymmetric: LU). Here the
    # difference carries the solves' rounding, amplified by B''s conditioning (about 1.4e5 on
    # this grid): measured on the CPU, central differences at h = 1e-6 agree with the exact
    # gradient to 1.3e-5 relative at worst, so the bar is 1e-4 (a wrong sign, index or
    # transposition is off by O(1))
    for q in rng.choice(bp.nnz, size=200, replace=False):
        plus = O.Matrix(bp.rows, bp.cols, bp.row_major, bp.ptr, bp.idx, bp.data.copy())
        minus = O.Matrix(bp.rows, bp.cols, bp.row_major, bp.ptr, bp.idx, bp.data.copy())
        plus.data[q] += h
        minus.data[q] -= h
        fd = (flow_loss(plus, bf, p, limit)[0] - flow_loss(minus, bf, p, limit)[0]) / (2 * h)
        assert abs(fd - a_bar[q]) <= 1e-4 * max(1.0, abs(a_bar[q])), (q, fd, a_bar[q], loss)


def test_adjoint_over_partitioned_sweeps_matches_the_cpu(gpu, monkeypatch):
    """A Cholesky factor swept as independent elimination subtrees (forward, transposed and the
    refinement's FP64 residual path alike) gives the CPU composition's bits."""
    monkeypatch.setenv("YBUS_HIP_PARTS", "8")
    g = O.make_grid(3000, 99)
    a, _, _ = g.dc()
    cpu_factor = O.Cholesky(a)
    factor = Y.GpuFactor.from_cholesky(gpu, cpu_factor)
    factor.set_matrix(a)
    x_bar = right_hand_sides(a.rows, 24, 7)
    x = right_hand_sides(a.rows, 24, 8)
    want_b, want_a = O.adjoint(cpu_factor, a, x_bar, x)
    b_bar, a_bar = factor.adjoint(x_bar, x)
    assert b_bar.tobytes() == want_b.tobytes()
    assert a_bar.tobytes() == want_a.tobytes()
