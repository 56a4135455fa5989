"""Complex factors and their CPU solves: test infrastructure only.

A pure-Python restatement of ybus_sparse's complex (Complex64) solve path, for checking the GPU's
complex solves. Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU leg may import
it; the product never does.

Arithmetic. The reference's element type is num-complex 0.4's ``Complex<f64>`` (a crates.io
dependency, ``ybus_sparse/Cargo.toml``, absent here and not pinned by a lock file). Its published
formulas are restated in ``zmul`` and ``zdiv``, each product and sum rounded on its own (Rust does
not contract them). Every function below works on real and imaginary parts held apart, as
Python floats or as float64 numpy arrays over many right-hand sides at once (numpy rounds each
elementwise operation on its own, so the two agree bit for bit).

Parity. The reference's complex tests check tolerances only (accuracy_test.rs:105-115,
sparse_rhs_test.rs:239), so no golden vector pins the rounding of a complex product or quotient.
What pins them: the factor satisfies L L^H = P A P^T, the solves agree with a dense LAPACK solve
to 1e-12, the transposed solves with one of A^T (plain, not conjugated), and the Cholesky
transposed solve is the conjugate identity (tests/test_oracle_complex.py). Bit for bit, where the
mathematics fixes the bits (tests/test_oracle_complex_rounding.py): products whose real products
cancel only when each is rounded on its own (no fused multiply-add), quotients of Gaussian
integers, which num-complex's formula rounds correctly and a scaled (Smith) or reciprocal
division does not, and 2 x 2 LU and Cholesky solves that round only in their last division, whose
results must be the correctly rounded exact solutions.

Factors. ``ComplexCholesky`` restates the up-looking numeric factorization
(SymbolicCholesky::factor, cholesky.rs:206-311) over the symbolic analysis (ordering and pattern)
of the oracle's real Cholesky of a matrix with the same pattern. ``ComplexLu`` takes the real LU's
ordering, pivot sequence and patterns and computes the complex values by dense elimination
without further pivoting: a genuine factorization P A Q = L U of the complex matrix, but not the
reference's own pivot choice for it (the GPU path starts from an uploaded factor; the factorization
itself is host work outside it).
"""
from __future__ import annotations

import math

import numpy as np


# ---- num-complex 0.4 arithmetic ----


def zmul(ar, ai, br, bi):
    """(a.re b.re - a.im b.im, a.re b.im + a.im b.re): num-complex's Mul for Complex<T>."""
    return ar * br - ai * bi, ar * bi + ai * br


def zdiv(ar, ai, br, bi):
    """num-complex's Div: s = b.re b.re + b.im b.im, then ((a.re b.re + a.im b.im) / s,
    (a.im b.re - a.re b.im) / s)."""
    s = br * br + bi * bi
    return (ar * br + ai * bi) / s, (ai * br - ar * bi) / s


def magnitude_squared(re, im):
    """properties.rs:141-143: re re + im im."""
    return re * re + im * im


# ---- test matrices ----


def hermitian_from(a, seed: int = 1):
    """A Hermitian positive definite matrix with the pattern of a real symmetric positive definite
    one (CSR ``Matrix`` with nonpositive off-diagonal entries, like B'): entry (i, j) off the
    diagonal is a_ij e^{-i t_ij} with a phase t_ij = -t_ji, so x^H A x >= |x|^T A |x| > 0.
    Returns (re, im) data arrays on ``a``'s pattern."""
    ptr, idx, val = np.asarray(a.ptr), np.asarray(a.idx), np.asarray(a.data, dtype=np.float64)
    re = np.empty(len(val))
    im = np.empty(len(val))
    for i in range(a.rows):
        for p in range(ptr[i], ptr[i + 1]):
            j = int(idx[p])
            if j == i:
                re[p], im[p] = val[p], 0.0
                continue
            lo, hi = min(i, j), max(i, j)
            t = 0.7 * math.sin(seed + 1.3 * lo + 0.37 * hi)
            if i > j:
                t = -t
            re[p] = val[p] * math.cos(t)
            im[p] = -val[p] * math.sin(t)
    return re, im


class Compressed:
    """A compressed (CSR or CSC) complex matrix's arrays: ptr, idx and complex data."""

    def __init__(self, ptr, idx, data):
        self.ptr, self.idx, self.data = ptr, idx, data


# ---- complex Cholesky (cholesky.rs) ----


class ComplexCholesky:
    """P A P^T = L L^H for a Hermitian positive definite A (CSR, ``a`` with ``re``/``im`` values),
    over ``symbolic``: the oracle's real ``Cholesky`` of a matrix with A's pattern (its ``perm``
    and L's pattern). Holds L (CSR, rows ending with their diagonal) and L^H (CSR, rows starting
    with their diagonal) as ptr/idx/re/im arrays, as the reference's Cholesky does."""

    def __init__(self, a, re, im, symbolic):
        n = a.rows
        perm = np.asarray(symbolic.perm, dtype=np.int64)
        pinv = np.empty(n, dtype=np.int64)
        pinv[perm] = np.arange(n)
        l_ptr = np.asarray(symbolic.l.ptr, dtype=np.int64)
        l_idx = np.asarray(symbolic.l.idx, dtype=np.int64)
        a_ptr, a_idx = np.asarray(a.ptr), np.asarray(a.idx)
        # column j's entries below the diagonal, from rows already done
        col_rows = [[] for _ in range(n)]
        col_re = [[] for _ in range(n)]
        col_im = [[] for _ in range(n)]
        l_re = np.zeros(len(l_idx))
        l_im = np.zeros(len(l_idx))
        diagonal = [0.0] * n
        work_re = [0.0] * n
        work_im = [0.0] * n
        for i in range(n):
            # row i of P A P^T left of the diagonal into the work vector, and its (real) diagonal
            r = int(perm[i])
            remaining = None
            for p in range(a_ptr[r], a_ptr[r + 1]):
                col = int(pinv[a_idx[p]])
                if col < i:
                    work_re[col], work_im[col] = float(re[p]), float(im[p])
                elif col == i:
                    remaining = float(re[p])
            if remaining is None:
                remaining = 0.0
            diag_slot = int(l_ptr[i + 1]) - 1
            for slot in range(int(l_ptr[i]), diag_slot):
                j = int(l_idx[slot])
                # l_ij = work[j] / diagonal[j], the diagonal a complex value with zero imaginary part
                lr, li = zdiv(work_re[j], work_im[j], diagonal[j], 0.0)
                work_re[j], work_im[j] = 0.0, 0.0
                for m, cr, ci in zip(col_rows[j], col_re[j], col_im[j]):
                    # work[m] = work[m] - conj(l_mj) * l_ij
                    pr, pi = zmul(cr, -ci, lr, li)
                    work_re[m], work_im[m] = work_re[m] - pr, work_im[m] - pi
                col_rows[j].append(i)
                col_re[j].append(lr)
                col_im[j].append(li)
                l_re[slot], l_im[slot] = lr, li
                remaining = remaining - magnitude_squared(lr, li)
            if math.isnan(remaining) or remaining <= 0.0:
                raise ValueError(f"not positive definite at row {int(perm[i])}")
            diagonal[i] = math.sqrt(remaining)
            l_re[diag_slot], l_im[diag_slot] = diagonal[i], 0.0
        self.n = n
        self.perm = perm
        self.l_ptr, self.l_idx, self.l_re, self.l_im = l_ptr, l_idx, l_re, l_im
        # L^H in CSR: each column's diagonal, then its entries below, conjugated
        lt_ptr, lt_idx, lt_re, lt_im = [0], [], [], []
        for j in range(n):
            lt_idx.append(j)
            lt_re.append(diagonal[j])
            lt_im.append(0.0)
            lt_idx.extend(col_rows[j])
            lt_re.extend(col_re[j])
            lt_im.extend(-v for v in col_im[j])
            lt_ptr.append(len(lt_idx))
        self.lt_ptr = np.array(lt_ptr, dtype=np.int64)
        self.lt_idx = np.array(lt_idx, dtype=np.int64)
        self.lt_re = np.array(lt_re)
        self.lt_im = np.array(lt_im)

    @property
    def l_values(self) -> np.ndarray:
        """L's values as complex128, each part exactly as stored (signed zeros kept)."""
        v = np.empty(len(self.l_re), dtype=np.complex128)
        v.real = self.l_re
        v.imag = self.l_im
        return v

    @property
    def l(self) -> Compressed:
        """L in CSR, as Cholesky::l() (cholesky.rs:342-344)."""
        return Compressed(self.l_ptr, self.l_idx, self.l_values)

    def solve_many(self, b: np.ndarray) -> np.ndarray:
        """A x = b for each row of the (count, n) complex array b: y = P b, forward_by_rows with
        L, back_by_rows with L^H (cholesky.rs:468-501), x = P^T y."""
        yr, yi = _columns(b, self.perm)
        _forward_by_rows(self.l_ptr, self.l_idx, self.l_re, self.l_im, yr, yi)
        _back_by_rows(self.lt_ptr, self.lt_idx, self.lt_re, self.lt_im, yr, yi)
        return _scatter(yr, yi, self.perm)

    def solve_transpose_many(self, b: np.ndarray) -> np.ndarray:
        """A^T x = b: A^T = conj(A) for a Hermitian A, so x = conj(A^-1 conj(b))."""
        return np.conj(self.solve_many(np.conj(b)))


def _columns(b, gather):
    """Right-hand sides as (n, count) real and imaginary arrays, row i taken from b[:, gather[i]]."""
    b = np.asarray(b, dtype=np.complex128)
    return np.ascontiguousarray(b.real.T[gather]), np.ascontiguousarray(b.imag.T[gather])


def _scatter(yr, yi, where):
    """x[:, where[k]] = y[k], as a (count, n) complex array."""
    n, count = yr.shape
    x = np.empty((count, n), dtype=np.complex128)
    x.real[:, where] = yr.T
    x.imag[:, where] = yi.T
    return x


def _forward_by_rows(ptr, idx, vr, vi, yr, yi):
    """cholesky.rs:468-483: each row's remainder less its entries left of the diagonal in stored
    order, divided by the diagonal (the first entry at or right of it)."""
    n = yr.shape[0]
    for i in range(n):
        rr, ri = yr[i].copy(), yi[i].copy()
        dr, di = 1.0, 0.0
        for p in range(ptr[i], ptr[i + 1]):
            col = idx[p]
            if col < i:
                pr, pi = zmul(vr[p], vi[p], yr[col], yi[col])
                rr, ri = rr - pr, ri - pi
            else:
                dr, di = vr[p], vi[p]
                break
        yr[i], yi[i] = zdiv(rr, ri, dr, di)


def _back_by_rows(ptr, idx, vr, vi, yr, yi):
    """cholesky.rs:487-501: rows from the last, each row's remainder less its entries right of the
    diagonal in stored order, divided by the diagonal (found anywhere in the row)."""
    n = yr.shape[0]
    for i in range(n - 1, -1, -1):
        rr, ri = yr[i].copy(), yi[i].copy()
        dr, di = 1.0, 0.0
        for p in range(ptr[i], ptr[i + 1]):
            col = idx[p]
            if col == i:
                dr, di = vr[p], vi[p]
            elif col > i:
                pr, pi = zmul(vr[p], vi[p], yr[col], yi[col])
                rr, ri = rr - pr, ri - pi
        yr[i], yi[i] = zdiv(rr, ri, dr, di)


# ---- complex LU (lu.rs) ----


class ComplexLu:
    """P A Q = L U of a complex A (dense ``a``, complex128) over ``real_lu``: the oracle's real
    ``Lu`` of a matrix with A's pattern, whose row and column permutations and L and U patterns
    are kept; the values come from dense elimination of P A Q without further pivoting. L and U
    are CSC with complex ``l_val`` / ``u_val`` (L's unit diagonal first, U's diagonal last)."""

    def __init__(self, a: np.ndarray, real_lu):
        n = a.shape[0]
        self.n = n
        self.row_perm = np.asarray(real_lu.row_perm, dtype=np.int64)
        self.col_perm = np.asarray(real_lu.col_perm, dtype=np.int64)
        m = np.array(a, dtype=np.complex128)[np.ix_(self.row_perm, self.col_perm)]
        for k in range(n):
            m[k + 1:, k] /= m[k, k]
            m[k + 1:, k + 1:] -= np.outer(m[k + 1:, k], m[k, k + 1:])
        self.l_ptr = np.asarray(real_lu.l.ptr, dtype=np.int64)
        self.l_idx = np.asarray(real_lu.l.idx, dtype=np.int64)
        self.u_ptr = np.asarray(real_lu.u.ptr, dtype=np.int64)
        self.u_idx = np.asarray(real_lu.u.idx, dtype=np.int64)
        lower = np.tril(m, -1) + np.eye(n)
        upper = np.triu(m)
        lcols = np.repeat(np.arange(n), np.diff(self.l_ptr))
        ucols = np.repeat(np.arange(n), np.diff(self.u_ptr))
        self.l_val = lower[self.l_idx, lcols]
        self.u_val = upper[self.u_idx, ucols]
        # the symbolic patterns hold every entry elimination can fill
        lmask = np.zeros((n, n), bool)
        lmask[self.l_idx, lcols] = True
        umask = np.zeros((n, n), bool)
        umask[self.u_idx, ucols] = True
        if np.any(lower[~lmask] != 0) or np.any(upper[~umask] != 0):
            raise ValueError("an entry falls outside the real factor's pattern")

    @property
    def l(self) -> Compressed:
        return Compressed(self.l_ptr, self.l_idx, self.l_val)

    @property
    def u(self) -> Compressed:
        return Compressed(self.u_ptr, self.u_idx, self.u_val)

    def solve_many(self, b: np.ndarray) -> np.ndarray:
        """Substitute (lu.rs:770-785): y = P b, L by columns, back_substitute with U
        (lu.rs:640-649), then x[Q[k]] = y[k]."""
        yr, yi = _columns(b, self.row_perm)
        lr, li = self.l_val.real, self.l_val.imag
        ur, ui = self.u_val.real, self.u_val.imag
        n = self.n
        for j in range(n):
            ar, ai = yr[j].copy(), yi[j].copy()
            for p in range(self.l_ptr[j] + 1, self.l_ptr[j + 1]):
                r = self.l_idx[p]
                pr, pi = zmul(lr[p], li[p], ar, ai)
                yr[r], yi[r] = yr[r] - pr, yi[r] - pi
        for j in range(n - 1, -1, -1):
            d = self.u_ptr[j + 1] - 1
            yr[j], yi[j] = zdiv(yr[j], yi[j], ur[d], ui[d])
            ar, ai = yr[j].copy(), yi[j].copy()
            for p in range(self.u_ptr[j], d):
                r = self.u_idx[p]
                pr, pi = zmul(ur[p], ui[p], ar, ai)
                yr[r], yi[r] = yr[r] - pr, yi[r] - pi
        return _scatter(yr, yi, self.col_perm)

    def solve_transpose_many(self, b: np.ndarray) -> np.ndarray:
        """Lu::solve_transpose (lu.rs:502-533), the plain transpose: y = Q^T b, U^T forward by
        its stored columns, unit L^T back, x[P[k]] = y[k]."""
        yr, yi = _columns(b, self.col_perm)
        lr, li = self.l_val.real, self.l_val.imag
        ur, ui = self.u_val.real, self.u_val.imag
        n = self.n
        for j in range(n):
            d = self.u_ptr[j + 1] - 1
            rr, ri = yr[j].copy(), yi[j].copy()
            for p in range(self.u_ptr[j], d):
                c = self.u_idx[p]
                pr, pi = zmul(ur[p], ui[p], yr[c], yi[c])
                rr, ri = rr - pr, ri - pi
            yr[j], yi[j] = zdiv(rr, ri, ur[d], ui[d])
        for j in range(n - 1, -1, -1):
            rr, ri = yr[j].copy(), yi[j].copy()
            for p in range(self.l_ptr[j] + 1, self.l_ptr[j + 1]):
                c = self.l_idx[p]
                pr, pi = zmul(lr[p], li[p], yr[c], yi[c])
                rr, ri = rr - pr, ri - pi
            yr[j], yi[j] = rr, ri
        return _scatter(yr, yi, self.row_perm)

    def dense(self) -> np.ndarray:
        """A = P^T L U Q^T from the stored factors."""
        n = self.n
        lower = np.zeros((n, n), np.complex128)
        upper = np.zeros((n, n), np.complex128)
        lower[self.l_idx, np.repeat(np.arange(n), np.diff(self.l_ptr))] = self.l_val
        upper[self.u_idx, np.repeat(np.arange(n), np.diff(self.u_ptr))] = self.u_val
        a = np.zeros((n, n), np.complex128)
        a[np.ix_(self.row_perm, self.col_perm)] = lower @ upper
        return a


def lu_refactor(lu: ComplexLu, a_ptr, a_idx, a_re, a_im, tau: float = 0.1) -> bool:
    """Lu::refactor (lu.rs:662-757) for Complex<f64>: the matrix (CSC, entries' real and imaginary
    parts apart) refactored with ``lu``'s pivots and patterns. Column k of P A Q is scattered into x
    by pivot step; for each row j above the diagonal in U(:, k), in increasing order, u_jk = x[j]
    and x[i] = x[i] - l_ij u_jk (num-complex's product, then the difference) over L(:, j) below its
    diagonal; the pivot x[k] is reused when it is nonzero and its magnitude (num-complex's
    ``abs``, the hypotenuse of its parts) is at least tau times the largest magnitude below it, in
    the reference's fold (b replaces a only when b > a); then l = x / pivot (num-complex's
    quotient). Returns True and replaces ``lu``'s values when every pivot passes ("Reused");
    False, leaving ``lu`` as it was, when one fails (the reference would factor afresh).
    The magnitudes are ``math.hypot``'s, so a pivot within a rounding of the threshold is where
    this and num-complex (libm's hypot) could decide apart: unpinned at such ties."""
    n = lu.n
    step_of = np.empty(n, dtype=np.int64)
    step_of[lu.row_perm] = np.arange(n)
    l_re, l_im = np.zeros(len(lu.l_val)), np.zeros(len(lu.l_val))
    u_re, u_im = np.zeros(len(lu.u_val)), np.zeros(len(lu.u_val))
    xr, xi = np.zeros(n), np.zeros(n)
    for k in range(n):
        col = lu.col_perm[k]
        for p in range(a_ptr[col], a_ptr[col + 1]):
            s = step_of[a_idx[p]]
            xr[s], xi[s] = a_re[p], a_im[p]
        diagonal = lu.u_ptr[k + 1] - 1
        for p in range(lu.u_ptr[k], diagonal):
            j = lu.u_idx[p]
            ur, ui = xr[j], xi[j]
            xr[j], xi[j] = 0.0, 0.0
            u_re[p], u_im[p] = ur, ui
            q = slice(lu.l_ptr[j] + 1, lu.l_ptr[j + 1])
            rows = lu.l_idx[q]
            pr, pi = zmul(l_re[q], l_im[q], ur, ui)
            xr[rows], xi[rows] = xr[rows] - pr, xi[rows] - pi
        pivot_r, pivot_i = xr[k], xi[k]
        largest = math.hypot(pivot_r, pivot_i)
        below = slice(lu.l_ptr[k] + 1, lu.l_ptr[k + 1])
        for row in lu.l_idx[below]:
            b = math.hypot(xr[row], xi[row])
            if b > largest:
                largest = b
        if (pivot_r == 0.0 and pivot_i == 0.0) or math.hypot(pivot_r, pivot_i) < tau * largest:
            return False
        u_re[diagonal], u_im[diagonal] = pivot_r, pivot_i
        xr[k], xi[k] = 0.0, 0.0
        l_re[lu.l_ptr[k]], l_im[lu.l_ptr[k]] = 1.0, 0.0
        rows = lu.l_idx[below]
        l_re[below], l_im[below] = zdiv(xr[rows], xi[rows], pivot_r, pivot_i)
        xr[rows], xi[rows] = 0.0, 0.0
    lu.l_val = l_re + 1j * l_im
    lu.u_val = u_re + 1j * u_im
    return True
