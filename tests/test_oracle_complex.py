"""The complex oracle (oracle/complex_solve.py) pinned where the reference leaves it open: its
complex tests check tolerances only (accuracy_test.rs:105-115, sparse_rhs_test.rs:239), so these
pin the restatement by what the mathematics fixes. The complex product and quotient agree with
exact rational arithmetic to within their rounding. The Cholesky factor satisfies L L^H = P A P^T
on the real factor's pattern, and its L^H is L's conjugate transpose with each diagonal kept. The
solves agree with a dense LAPACK solve, and the transposed ones with a solve of A^T (the plain
transpose). A factor of a real matrix agrees with the real oracle's. A conjugation dropped from
L^H fails the dense check."""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from oracle import complex_solve as Z
from tests.fixtures import (complex_lu_case, complex_relative_residual, complex_right_hand_sides,
                            grid_laplacian, hermitian_case, random_spd, random_square_cuda)


def _exact(v):
    return Fraction(float(v))


def test_complex_product_and_quotient_are_within_rounding_of_exact():
    rng = np.random.default_rng(5)
    eps = np.finfo(float).eps
    for _ in range(300):
        ar, ai, br, bi = (float(v) for v in rng.standard_normal(4) * 10.0 ** rng.integers(-3, 4, 4))
        pr, pi = Z.zmul(ar, ai, br, bi)
        er = _exact(ar) * _exact(br) - _exact(ai) * _exact(bi)
        ei = _exact(ar) * _exact(bi) + _exact(ai) * _exact(br)
        scale = abs(ar * br) + abs(ai * bi) + abs(ar * bi) + abs(ai * br)
        assert abs(Fraction(pr) - er) <= Fraction(2 * eps * scale)
        assert abs(Fraction(pi) - ei) <= Fraction(2 * eps * scale)
        qr, qi = Z.zdiv(ar, ai, br, bi)
        s = _exact(br) ** 2 + _exact(bi) ** 2
        xr = (_exact(ar) * _exact(br) + _exact(ai) * _exact(bi)) / s
        xi = (_exact(ai) * _exact(br) - _exact(ar) * _exact(bi)) / s
        size = (abs(ar) + abs(ai)) / float(np.hypot(br, bi))
        assert abs(Fraction(qr) - xr) <= Fraction(8 * eps * size)
        assert abs(Fraction(qi) - xi) <= Fraction(8 * eps * size)
    # exact cases: i * i = -1, (3 + 4i) / (3 + 4i) = 1, (1 + 2i) / 2 = 0.5 + i
    assert Z.zmul(0.0, 1.0, 0.0, 1.0) == (-1.0, 0.0)
    assert Z.zdiv(3.0, 4.0, 3.0, 4.0) == (1.0, 0.0)
    assert Z.zdiv(1.0, 2.0, 2.0, 0.0) == (0.5, 1.0)


@pytest.mark.parametrize("buses,seed", [(60, 3), (300, 7)])
def test_complex_cholesky_factors_a_hermitian_matrix(buses, seed):
    chol, a = hermitian_case(buses, seed)
    n = chol.n
    dense = a.toarray()
    assert np.array_equal(dense, dense.conj().T), "the fixture is Hermitian"
    lower = np.zeros((n, n), complex)
    lower[np.repeat(np.arange(n), np.diff(chol.l_ptr)), chol.l_idx] = chol.l_values
    pap = dense[np.ix_(chol.perm, chol.perm)]
    assert np.max(np.abs(lower @ lower.conj().T - pap)) <= 1e-14 * np.max(np.abs(pap))
    # each row ends with a real, positive diagonal
    last = chol.l_ptr[1:] - 1
    assert np.array_equal(chol.l_idx[last], np.arange(n))
    assert np.all(chol.l_re[last] > 0) and np.all(chol.l_im[last] == 0)
    # L^H: L's columns, diagonal first and kept, the rest conjugated
    upper = np.zeros((n, n), complex)
    upper[np.repeat(np.arange(n), np.diff(chol.lt_ptr)), chol.lt_idx] = chol.lt_re + 1j * chol.lt_im
    assert np.array_equal(upper, lower.conj().T)
    assert np.array_equal(chol.lt_idx[chol.lt_ptr[:-1]], np.arange(n))


def test_complex_cholesky_of_a_real_matrix_agrees_with_the_real_factor():
    a = random_spd(120, 5, 9)
    real = O.Cholesky(a)
    chol = Z.ComplexCholesky(a, np.asarray(a.data, float), np.zeros(a.nnz), real)
    assert np.array_equal(chol.l_idx, real.l.idx)
    assert np.all(chol.l_im == 0)
    assert np.max(np.abs(chol.l_re - real.l.data)) <= 1e-14 * np.max(np.abs(real.l.data))


@pytest.mark.parametrize("buses,seed", [(60, 3), (300, 7)])
def test_complex_cholesky_solves_agree_with_dense_solves(buses, seed):
    chol, a = hermitian_case(buses, seed)
    dense = a.toarray()
    b = complex_right_hand_sides(chol.n, 6, 2)
    x = chol.solve_many(b)
    want = np.linalg.solve(dense, b.T).T
    assert np.max(np.abs(x - want)) <= 1e-12 * np.max(np.abs(want))
    assert complex_relative_residual(a, x, b) <= 1e-13
    xt = chol.solve_transpose_many(b)
    assert complex_relative_residual(dense.T, xt, b) <= 1e-13
    # the transposed solve is the conjugate identity, exactly
    assert np.array_equal(xt, np.conj(chol.solve_many(np.conj(b))))


def test_a_dropped_conjugation_fails_the_dense_check():
    chol, a = hermitian_case(60, 3)
    b = complex_right_hand_sides(chol.n, 2, 4)
    want = np.linalg.solve(a.toarray(), b.T).T
    chol.lt_im = -chol.lt_im  # L^T where L^H belongs
    assert np.max(np.abs(chol.solve_many(b) - want)) > 1e-6 * np.max(np.abs(want))


def test_complex_lu_solves_agree_with_dense_solves():
    lu, a = complex_lu_case(120, 4, 3)
    assert np.max(np.abs(lu.dense() - a)) <= 1e-13 * np.max(np.abs(a))
    assert any(i != p for i, p in enumerate(lu.row_perm)), "the fixture should pivot off the diagonal"
    b = complex_right_hand_sides(lu.n, 5, 8)
    x = lu.solve_many(b)
    assert np.max(np.abs(x - np.linalg.solve(a, b.T).T)) <= 1e-11 * np.max(np.abs(x))
    # accuracy_test.rs:111-114's bar for this kind of matrix
    assert complex_relative_residual(a, x, b) <= 1e-10
    xt = lu.solve_transpose_many(b)
    assert np.max(np.abs(xt - np.linalg.solve(a.T, b.T).T)) <= 1e-11 * np.max(np.abs(xt))
    assert complex_relative_residual(a.T, xt, b) <= 1e-10
    # not the conjugate transpose
    assert complex_relative_residual(a.conj().T, xt, b) > 1e-3


def test_complex_lu_of_a_real_matrix_solves_as_the_real_lu():
    matrix = random_square_cuda(80, 5, 11)
    real = O.Lu(matrix)
    lu = Z.ComplexLu(matrix.to_dense().astype(complex), real)
    assert np.all(lu.l_val.imag == 0) and np.all(lu.u_val.imag == 0)
    assert np.max(np.abs(lu.u_val.real - real.u.data)) <= 1e-12 * np.max(np.abs(real.u.data))
    assert np.max(np.abs(lu.l_val.real - real.l.data)) <= 1e-12 * np.max(np.abs(real.l.data))
    b = np.random.default_rng(3).standard_normal((3, 80))
    x = lu.solve_many(b.astype(complex))
    assert np.all(x.imag == 0)
    assert np.max(np.abs(x.real - real.solve_many(b))) <= 1e-12 * np.max(np.abs(x.real))
    xt = lu.solve_transpose_many(b.astype(complex))
    assert np.max(np.abs(xt.real - real.solve_transpose_many(b))) <= 1e-12 * np.max(np.abs(xt.real))


# ---- the complex LU refactor (lu.rs:662-757 for Complex<f64>) ----


def complex_csc(real: O.Matrix, re_scale: float, im_scale: float):
    """The pattern of a real CSC matrix with each stored v made re_scale v + im_scale v i."""
    csc = real if not real.row_major else real.to_csc()
    return csc, re_scale * csc.data, im_scale * csc.data


def textbook_complex_refactor(dense: np.ndarray, lu: Z.ComplexLu):
    """Dense left-looking elimination of P A Q with lu's pivots in explicit real arithmetic:
    column k takes x[j+1:] = x[j+1:] - L[j+1:, j] x[j] for j = 0 .. k - 1 in order (num-complex's
    product, then the difference), then L[k+1:, k] = x[k+1:] / x[k] (num-complex's quotient)."""
    pa = dense[np.ix_(lu.row_perm, lu.col_perm)]
    n = pa.shape[0]
    lr, li = np.eye(n), np.zeros((n, n))
    ur, ui = np.zeros((n, n)), np.zeros((n, n))
    for k in range(n):
        xr, xi = pa[:, k].real.copy(), pa[:, k].imag.copy()
        for j in range(k):
            pr, pi = Z.zmul(lr[j + 1:, j], li[j + 1:, j], xr[j], xi[j])
            xr[j + 1:], xi[j + 1:] = xr[j + 1:] - pr, xi[j + 1:] - pi
        ur[:k + 1, k], ui[:k + 1, k] = xr[:k + 1], xi[:k + 1]
        lr[k + 1:, k], li[k + 1:, k] = Z.zdiv(xr[k + 1:], xi[k + 1:], xr[k], xi[k])
    return lr, li, ur, ui


def stored(ptr, idx, full):
    cols = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
    return full[idx, cols]


def test_complex_refactor_closed_form():
    # [[2i, 1], [1, 3]] with natural pivots: l21 = 1 / 2i = -i/2, u22 = 3 - (-i/2) 1 = 3 + i/2
    real = O.Matrix.from_triplets(2, 2, [(0, 0, 2.0), (0, 1, 1.0), (1, 0, 1.0), (1, 1, 3.0)], row_major=False)
    lu = Z.ComplexLu(real.to_dense().astype(complex), O.Lu(real, O.NATURAL))
    re = np.array([0.0, 1.0, 1.0, 3.0])
    im = np.array([2.0, 0.0, 0.0, 0.0])
    assert Z.lu_refactor(lu, real.ptr, real.idx, re, im)
    assert list(lu.l_val) == [1.0, -0.5j, 1.0]
    assert list(lu.u_val) == [2j, 1.0, 3 + 0.5j]


def test_complex_refactor_of_a_grid_laplacian():
    # factorization_test.rs:1032-1046: a = (v, -2v), b = (0.7v, -1.1v) on a 5 x 5 grid Laplacian
    real = grid_laplacian(5).to_csc()
    lu_real = O.Lu(real, O.AMD)
    a = real.to_dense() * (1.0 - 2.0j)
    lu = Z.ComplexLu(a, lu_real)
    csc, re, im = complex_csc(real, 0.7, -1.1)
    assert Z.lu_refactor(lu, csc.ptr, csc.idx, re, im)
    b = real.to_dense() * (0.7 - 1.1j)
    x = np.array([complex(i, 1.0) for i in range(25)])
    solved = lu.solve_many((b @ x)[None, :])[0]
    assert np.abs(solved - x).max() < 1e-10


@pytest.mark.parametrize("n,fill,seed", [(40, 8, 3), (70, 5, 9)])
def test_complex_refactor_equals_the_dense_textbook_elimination_bit_for_bit(n, fill, seed):
    lu, a = complex_lu_case(n, fill, seed)
    real = random_square_cuda(n, fill, seed)
    # new values on the same pattern: each stored entry scaled by a complex factor near 1
    csc = real
    rng = np.random.default_rng(seed)
    scale = 1.0 + 0.02 * (rng.random(csc.nnz) - 0.5) + 0.01j * (rng.random(csc.nnz) - 0.5)
    cols = np.repeat(np.arange(n), np.diff(csc.ptr))
    values = a[csc.idx, cols] * scale
    assert Z.lu_refactor(lu, csc.ptr, csc.idx, values.real.copy(), values.imag.copy())
    dense = np.zeros((n, n), complex)
    dense[csc.idx, cols] = values
    lr, li, ur, ui = textbook_complex_refactor(dense, lu)
    for got, want in [(lu.l_val.real, stored(lu.l_ptr, lu.l_idx, lr)), (lu.l_val.imag, stored(lu.l_ptr, lu.l_idx, li)),
                      (lu.u_val.real, stored(lu.u_ptr, lu.u_idx, ur)), (lu.u_val.imag, stored(lu.u_ptr, lu.u_idx, ui))]:
        assert got.tobytes() == want.tobytes()
    # and it factors the new matrix
    pa = dense[np.ix_(lu.row_perm, lu.col_perm)]
    l_full = np.zeros((n, n), complex)
    l_full[lu.l_idx, np.repeat(np.arange(n), np.diff(lu.l_ptr))] = lu.l_val
    u_full = np.zeros((n, n), complex)
    u_full[lu.u_idx, np.repeat(np.arange(n), np.diff(lu.u_ptr))] = lu.u_val
    assert np.abs(l_full @ u_full - pa).max() < 1e-12 * np.abs(pa).max()


def test_complex_refactor_refuses_a_small_pivot_and_keeps_the_factor():
    real = O.Matrix.from_triplets(2, 2, [(0, 0, 2.0), (0, 1, 1.0), (1, 0, 1.0), (1, 1, 3.0)], row_major=False)
    lu = Z.ComplexLu(real.to_dense().astype(complex), O.Lu(real, O.NATURAL))
    before = (lu.l_val.copy(), lu.u_val.copy())
    # the first pivot 0.01 + 0.01i against 1 below it
    assert not Z.lu_refactor(lu, real.ptr, real.idx, np.array([0.01, 1.0, 1.0, 3.0]), np.array([0.01, 0, 0, 0]))
    assert np.array_equal(lu.l_val, before[0]) and np.array_equal(lu.u_val, before[1])
    assert not Z.lu_refactor(lu, real.ptr, real.idx, np.array([0.0, 1.0, 1.0, 3.0]), np.zeros(4))
