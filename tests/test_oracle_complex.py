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
                            hermitian_case, random_spd, random_square_cuda)


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
