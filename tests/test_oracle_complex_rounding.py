"""The complex oracle's rounding (oracle/complex_solve.py) pinned bit for bit where the mathematics
fixes it. num-complex 0.4's Mul is (a.re b.re - a.im b.im, a.re b.im + a.im b.re) and its Div is
((a.re b.re + a.im b.im) / s, (a.im b.re - a.re b.im) / s) with s = b.re b.re + b.im b.im, each
operation a binary64 operation rounded to nearest on its own (rustc contracts nothing into a
fused multiply-add). Three mistakes a restatement could plausibly make give other bits here:
- a fused multiply-add in a product or a quotient's numerator (the cancelling cases below);
- a scaled (Smith) division, as CPython's complex quotient is, or a product with the reciprocal:
  a quotient of Gaussian integers in num-complex's formula has exact numerators and divisor, so
  it is the correctly rounded exact quotient, and those two are not on about half of them;
- a conjugation dropped from L^H (the Cholesky closed form).
The solves are pinned the same way on 2 x 2 factors whose solves round only in their last
division (tests/fixtures.gaussian_closed_forms): they must return the correctly rounded exact
solutions, found by Cramer's rule in exact rational arithmetic."""
from fractions import Fraction

import numpy as np
import pytest

from oracle import complex_solve as Z
from tests.fixtures import exact_solve_2x2, gaussian_closed_forms, rounded

H = 2.0 ** -30
A = 1.0 + H  # A * A = 1 + 2^-29 + 2^-60, which rounds to B
B = 1.0 + 2 * H


def test_the_cancelling_products_are_exact_only_when_each_is_rounded():
    assert Fraction(A) * Fraction(A) - Fraction(B) == Fraction(1, 2 ** 60)
    assert A * A == B


def test_a_product_rounds_each_real_product_on_its_own():
    # real part A A - B 1, imaginary part A A + (-B) 1: zero when rounded apart, 2^-60 when fused
    assert Z.zmul(A, B, A, 1.0)[0] == 0.0
    assert Z.zmul(A, -B, 1.0, A)[1] == 0.0
    # the same through numpy arrays, as the solves apply it to many right-hand sides at once
    re, im = Z.zmul(np.array([A, A]), np.array([B, -B]), np.array([A, 1.0]), np.array([1.0, A]))
    assert re[0] == 0.0 and im[1] == 0.0


def test_a_quotient_rounds_each_numerator_product_on_its_own():
    # numerators A A + (-B) 1 and A A - B 1
    assert Z.zdiv(A, -B, A, 1.0)[0] == 0.0
    assert Z.zdiv(B, A, A, 1.0)[1] == 0.0
    re, im = Z.zdiv(np.array([A, B]), np.array([-B, A]), np.array([A, A]), np.array([1.0, 1.0]))
    assert re[0] == 0.0 and im[1] == 0.0


def test_a_quotient_of_gaussian_integers_is_correctly_rounded():
    rng = np.random.default_rng(11)
    smith_apart = reciprocal_apart = cases = 0
    for a, b, c, d in rng.integers(-20, 21, (3000, 4)).tolist():
        if c == 0 and d == 0:
            continue
        cases += 1
        s = c * c + d * d
        want = (float(Fraction(a * c + b * d, s)), float(Fraction(b * c - a * d, s)))
        assert Z.zdiv(float(a), float(b), float(c), float(d)) == want, (a, b, c, d)
        q = complex(a, b) / complex(c, d)
        smith_apart += (q.real, q.imag) != want
        reciprocal_apart += Z.zmul(float(a), float(b), c / s, -d / s) != want
    # the two other quotients miss the correctly rounded one on a large share of these cases
    assert smith_apart > cases // 4 and reciprocal_apart > cases // 4


def _oracle_lu(arrays):
    lu = Z.ComplexLu.__new__(Z.ComplexLu)
    lu.n = 2
    for key in ("l_ptr", "l_idx", "u_ptr", "u_idx", "row_perm", "col_perm"):
        setattr(lu, key, np.asarray(arrays[key], dtype=np.int64))
    lu.l_val = np.asarray(arrays["l_val"])
    lu.u_val = np.asarray(arrays["u_val"])
    return lu


def _oracle_cholesky(arrays):
    ch = Z.ComplexCholesky.__new__(Z.ComplexCholesky)
    ch.n = 2
    ch.perm = np.asarray(arrays["perm"], dtype=np.int64)
    ch.l_ptr = np.asarray(arrays["l_ptr"], dtype=np.int64)
    ch.l_idx = np.asarray(arrays["l_idx"], dtype=np.int64)
    val = np.asarray(arrays["l_val"])
    ch.l_re, ch.l_im = val.real.copy(), val.imag.copy()
    # L^H in CSR, rows starting with their diagonal: [l00, conj(l10); 0, l11]
    ch.lt_ptr = np.array([0, 2, 3], dtype=np.int64)
    ch.lt_idx = np.array([0, 1, 1], dtype=np.int64)
    ch.lt_re = np.array([val[0].real, val[1].real, val[2].real])
    ch.lt_im = np.array([val[0].imag, -val[1].imag, val[2].imag])
    return ch


def oracle_closed_form_solve(name, case):
    arrays, b = case["arrays"], case["b"]
    if name == "lu":
        return _oracle_lu(arrays).solve_many(b)
    if name == "lu_transpose":
        return _oracle_lu(arrays).solve_transpose_many(b)
    return _oracle_cholesky(arrays).solve_many(b)


def same_bits(x, y) -> bool:
    """Equal bit for bit, a zero's sign aside (adding +0.0 turns -0.0 into +0.0)."""
    return (np.asarray(x) + 0.0).tobytes() == (np.asarray(y) + 0.0).tobytes()


@pytest.mark.parametrize("name", ["lu", "lu_transpose", "cholesky"])
def test_closed_form_solves_return_the_correctly_rounded_exact_solutions(name):
    case = gaussian_closed_forms()[name]
    want = rounded(case["exact"])
    assert same_bits(oracle_closed_form_solve(name, case), want)
    # only the last division rounds: the other component is exact
    other = 1 - case["last"]
    for x in case["exact"]:
        assert all(v.denominator in (1, 2) for v in x[other])


@pytest.mark.parametrize("name", ["lu", "lu_transpose"])
def test_the_closed_forms_tell_a_scaled_division_apart(name):
    case = gaussian_closed_forms()[name]
    want = rounded(case["exact"])
    last, divisor = case["last"], case["divisor"]
    apart = 0
    for k, x in enumerate(case["exact"]):
        # the last division's numerator: exact, and representable by construction
        num = (Fraction(divisor.real) * x[last][0] - Fraction(divisor.imag) * x[last][1],
               Fraction(divisor.real) * x[last][1] + Fraction(divisor.imag) * x[last][0])
        numerator = complex(float(num[0]), float(num[1]))
        assert Fraction(numerator.real) == num[0] and Fraction(numerator.imag) == num[1]
        apart += numerator / divisor != want[k, last]
    assert apart >= len(want) // 4


def test_the_cholesky_closed_form_tells_a_dropped_conjugation_apart():
    case = gaussian_closed_forms()["cholesky"]
    lo = np.array([[3, 0], [2 - 1j, 1]], dtype=np.complex128)
    plain = rounded(exact_solve_2x2(lo @ lo.T, case["b"]))
    assert not same_bits(plain, rounded(case["exact"]))
