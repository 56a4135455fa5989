"""The oracle's adjoint composition (the CPU check of the GPU adjoint), pinned on its own: dense
algebra on a small unsymmetric system, and central finite differences of a quadratic loss."""
import numpy as np

import oracle as O
from tests.fixtures import random_square_cuda, right_hand_sides


def test_composition_is_the_dense_gradient():
    a = random_square_cuda(12, 30, 4)
    lu = O.Lu(a)
    x_bar = right_hand_sides(12, 3, 1)
    x = right_hand_sides(12, 3, 2)
    b_bar, a_bar = O.adjoint(lu, a, x_bar, x)
    dense = a.to_dense()
    assert np.allclose(b_bar, np.linalg.solve(dense.T, x_bar.T).T, rtol=1e-12, atol=1e-14)
    full = -(b_bar.T @ x)
    major = np.repeat(np.arange(a.cols), np.diff(a.ptr))  # CSC: columns
    assert np.allclose(a_bar, full[a.idx, major], rtol=1e-12, atol=1e-14)


def test_composition_matches_finite_differences():
    a = random_square_cuda(10, 40, 9)
    b = right_hand_sides(10, 2, 3)
    w = right_hand_sides(10, 2, 4)

    def loss(m):
        x = O.Lu(m).solve_many(b)
        return 0.5 * float(np.sum((x - w) ** 2)), x

    _, x = loss(a)
    b_bar, a_bar = O.adjoint(O.Lu(a), a, x - w, x)
    h = 1e-6
    for q in range(a.nnz):
        up = O.Matrix(a.rows, a.cols, a.row_major, a.ptr, a.idx, a.data.copy())
        dn = O.Matrix(a.rows, a.cols, a.row_major, a.ptr, a.idx, a.data.copy())
        up.data[q] += h
        dn.data[q] -= h
        fd = (loss(up)[0] - loss(dn)[0]) / (2 * h)
        assert abs(fd - a_bar[q]) <= 1e-6 * max(1.0, abs(a_bar[q])), q
