"""Deterministic test matrices and independent reference algorithms, restating the reference's
test helpers (ybus_cuda/tests/common/mod.rs, ybus_sparse/tests/common/mod.rs,
ybus_sparse/tests/factorization_test.rs). Pure Python, so they check the C++ oracle
independently of it."""
from __future__ import annotations

import math

import numpy as np

from tests.conftest import Lcg
import oracle as O


# ---- ybus_cuda/tests/common/mod.rs ----

def random_spd(n: int, fill_percent: int, seed: int) -> O.Matrix:
    """common/mod.rs:35-53 (and factorization_test.rs:198-216, the same generator)."""
    rng = Lcg(seed)
    triplets = []
    row_sums = [0.0] * n
    for i in range(n):
        for j in range(i):
            if rng.next() % 100 < fill_percent:
                value = -(float(rng.next() % 1000) + 1.0) / 997.0
                triplets += [(i, j, value), (j, i, value)]
                row_sums[i] -= value
                row_sums[j] -= value
    for i, s in enumerate(row_sums):
        triplets.append((i, i, s + 1.0 + i / n))
    return O.Matrix.from_triplets(n, n, triplets)


def random_square_cuda(n: int, fill_percent: int, seed: int) -> O.Matrix:
    """common/mod.rs:56-68: unsymmetric with a small diagonal, so the LU pivots off it (CSC)."""
    rng = Lcg(seed)
    triplets = []
    for i in range(n):
        for j in range(n):
            if i != j and rng.next() % 100 < fill_percent:
                triplets.append((i, j, 3.0 * rng.value()))
        triplets.append((i, i, 0.1 + abs(rng.value())))
    return O.Matrix.from_triplets(n, n, triplets, row_major=False)


def two_by_two_blocks(blocks: int) -> O.Matrix:
    """common/mod.rs:72-83 (CSC)."""
    t = []
    for b in range(blocks):
        p, q = 2 * b, 2 * b + 1
        t += [(p, p, 4.0), (p, q, 1.0), (q, p, -0.5), (q, q, 3.0)]
        if b + 1 < blocks:
            r, s = p + 2, q + 2
            t += [(p, r, -1.0), (r, p, -1.0), (q, s, -0.7), (s, q, -0.8)]
    return O.Matrix.from_triplets(2 * blocks, 2 * blocks, t, row_major=False)


def right_hand_sides(n: int, count: int, seed: int) -> np.ndarray:
    """common/mod.rs:85-90: (count, n)."""
    rng = Lcg(seed)
    return np.array([[rng.value() for _ in range(n)] for _ in range(count)], dtype=np.float64).reshape(count, n)


def worst_relative_difference(gpu: np.ndarray, cpu: np.ndarray) -> float:
    """common/mod.rs:94-109."""
    worst = 0.0
    for g, c in zip(gpu, cpu):
        scale = max(float(np.max(np.abs(c))) if len(c) else 0.0, np.finfo(float).tiny)
        worst = max(worst, float(np.max(np.abs(g - c))) / scale if len(c) else 0.0)
    return worst


# ---- ybus_sparse/tests/factorization_test.rs ----

def grid_laplacian(k: int) -> O.Matrix:
    """factorization_test.rs:38-53."""
    t = []
    for r in range(k):
        for c in range(k):
            i = r * k + c
            t.append((i, i, 4.0))
            if c + 1 < k:
                t += [(i, i + 1, -1.0), (i + 1, i, -1.0)]
            if r + 1 < k:
                t += [(i, i + k, -1.0), (i + k, i, -1.0)]
    return O.Matrix.from_triplets(k * k, k * k, t)


def banded(n: int, band: int) -> O.Matrix:
    """factorization_test.rs:56-65."""
    t = []
    for i in range(n):
        t.append((i, i, 2.0 * band + 1.0))
        for d in range(1, min(band, n - 1 - i) + 1):
            t += [(i, i + d, -1.0), (i + d, i, -1.0)]
    return O.Matrix.from_triplets(n, n, t)


def graph_matrix(n: int, edges) -> O.Matrix:
    """factorization_test.rs:546-556: Laplacian plus identity."""
    degree = [0.0] * n
    t = []
    for i, j in edges:
        t += [(i, j, -1.0), (j, i, -1.0)]
        degree[i] += 1.0
        degree[j] += 1.0
    t += [(i, i, d + 1.0) for i, d in enumerate(degree)]
    return O.Matrix.from_triplets(n, n, t)


def random_square_sparse(n: int, fill_percent: int, diagonal_scale: float, seed: int) -> O.Matrix:
    """factorization_test.rs:671-696 (CSR)."""
    rng = Lcg(seed)
    t = []
    for i in range(n):
        largest = 1.0
        for j in range(n):
            if i != j and rng.next() % 100 < fill_percent:
                value = (float(rng.next() % 2001) - 1000.0) / 313.0
                largest = max(largest, abs(value))
                t.append((i, j, value))
        sign = 1.0 if rng.next() % 2 == 0 else -1.0
        t.append((i, i, sign * diagonal_scale * largest * (1.0 + i / n)))
    return O.Matrix.from_triplets(n, n, t)


def brute_force_factor_pattern(a: O.Matrix, perm) -> np.ndarray:
    """factorization_test.rs:220-244: dense symbolic elimination of the reordered pattern."""
    n = len(perm)
    inverse = np.zeros(n, dtype=np.int64)
    inverse[np.asarray(perm)] = np.arange(n)
    filled = np.zeros((n, n), dtype=bool)
    for major in range(a.rows):
        for p in range(a.ptr[major], a.ptr[major + 1]):
            i, j = inverse[major], inverse[a.idx[p]]
            filled[max(i, j), min(i, j)] = True
    np.fill_diagonal(filled, True)
    for k in range(n):
        below = [i for i in range(k + 1, n) if filled[i, k]]
        for x, i in enumerate(below):
            for j in below[: x + 1]:
                filled[i, j] = True
    return filled


def is_permutation(perm, n: int) -> bool:
    return len(perm) == n and sorted(int(p) for p in perm) == list(range(n))


# ---- ybus_sparse/tests/common/mod.rs: the pre-Phase-3 algorithms, independent references ----

def reference_cholesky(a: O.Matrix):
    """common/mod.rs:12-62: unordered up-looking Cholesky with a sorted work map, dropping exact
    zeros. Returns (ptr, idx, data) of L in CSR."""
    n = a.rows
    diagonal = [0.0] * n
    columns = [[] for _ in range(n)]
    ptr, idx, data = [0], [], []
    for i in range(n):
        work = {}
        a_ii = 0.0
        for p in range(a.ptr[i], a.ptr[i + 1]):
            col = int(a.idx[p])
            if col < i:
                work[col] = float(a.data[p])
            else:
                if col == i:
                    a_ii = float(a.data[p])
                break
        remaining = a_ii
        while work:
            j = min(work)
            w_j = work.pop(j)
            l_ij = w_j / diagonal[j]
            if l_ij == 0.0:
                continue
            for m, l_mj in columns[j]:
                current = work.get(m, 0.0)
                work[m] = current - l_mj * l_ij
            columns[j].append((i, l_ij))
            idx.append(j)
            data.append(l_ij)
            remaining -= l_ij * l_ij
        if math.isnan(remaining) or remaining <= 0.0:
            raise ValueError(f"not positive definite at {i}")
        diagonal[i] = math.sqrt(remaining)
        idx.append(i)
        data.append(diagonal[i])
        ptr.append(len(idx))
    return np.array(ptr), np.array(idx), np.array(data)


def reference_lu_perm(a: O.Matrix):
    """common/mod.rs:88-165: dense right-looking elimination with true partial pivoting (the
    first row of largest magnitude); returns the row permutation."""
    n = a.rows
    dense = a.to_csr().to_dense()
    rows = [dict((j, dense[i, j]) for j in range(n) if dense[i, j] != 0.0) for i in range(n)]
    perm = list(range(n))
    for k in range(n):
        def mag(row):
            return abs(row.get(k, 0.0))
        best, best_mag = k, mag(rows[k])
        for i in range(k + 1, n):
            if mag(rows[i]) > best_mag:
                best, best_mag = i, mag(rows[i])
        if best_mag == 0.0:
            raise ValueError("singular")
        if best != k:
            rows[k], rows[best] = rows[best], rows[k]
            perm[k], perm[best] = perm[best], perm[k]
        pivot = rows[k][k]
        pivot_row = [(j, v) for j, v in sorted(rows[k].items()) if j > k]
        for row in rows[k + 1:]:
            if k not in row:
                continue
            multiplier = row[k] / pivot
            if multiplier == 0.0:
                del row[k]
                continue
            row[k] = multiplier
            for j, u_kj in pivot_row:
                updated = row.get(j, 0.0) - multiplier * u_kj
                if updated == 0.0:
                    row.pop(j, None)
                else:
                    row[j] = updated
    return perm


# ---- complex cases (accuracy_test.rs:105-115 complexifies a real matrix the same way) ----

def hermitian_case(buses: int, seed: int):
    """A Hermitian positive definite matrix on a synthetic grid's B' pattern and its complex
    Cholesky factor: (ComplexCholesky, A as a scipy CSR matrix)."""
    import scipy.sparse as sp

    from oracle import complex_solve as Z

    grid = O.make_grid(buses, seed)
    bprime, _, _ = grid.dc()
    re, im = Z.hermitian_from(bprime, seed)
    chol = Z.ComplexCholesky(bprime, re, im, O.Cholesky(bprime))
    a = sp.csr_matrix((re + 1j * im, np.asarray(bprime.idx), np.asarray(bprime.ptr)), shape=bprime.shape)
    return chol, a


def complex_lu_case(n: int, fill_percent: int, seed: int):
    """random_square_cuda's matrix with each stored v made v + (0.4 v - 0.1) i, and its complex LU
    over the real matrix's pivot sequence: (ComplexLu, A dense)."""
    from oracle import complex_solve as Z

    real = random_square_cuda(n, fill_percent, seed)
    lu = O.Lu(real)
    dense = real.to_dense()
    a = dense + 1j * np.where(dense != 0, 0.4 * dense - 0.1, 0.0)
    return Z.ComplexLu(a, lu), a


def complex_right_hand_sides(n: int, count: int, seed: int) -> np.ndarray:
    """(count, n) complex: right_hand_sides' values as real parts, a second stream's as imaginary."""
    return right_hand_sides(n, count, seed) + 1j * right_hand_sides(n, count, seed + 1000003)


def complex_relative_residual(a, x: np.ndarray, b: np.ndarray) -> float:
    """The worst column's ||b - A x||inf / ||b||inf (accuracy.rs:57-78), A dense or scipy sparse."""
    r = b.T - a @ x.T
    return float(np.max(np.max(np.abs(r), axis=0) / np.max(np.abs(b.T), axis=0)))


def _gauss(z):
    """A complex number's parts as exact Fractions."""
    from fractions import Fraction
    return Fraction(float(z.real)), Fraction(float(z.imag))


def _gauss_mul(a, b):
    return a[0] * b[0] - a[1] * b[1], a[0] * b[1] + a[1] * b[0]


def _gauss_div(a, b):
    s = b[0] * b[0] + b[1] * b[1]
    return (a[0] * b[0] + a[1] * b[1]) / s, (a[1] * b[0] - a[0] * b[1]) / s


def _gauss_sub(a, b):
    return a[0] - b[0], a[1] - b[1]


def exact_solve_2x2(a: np.ndarray, b: np.ndarray):
    """The exact solution of the 2 x 2 complex system a x = b (Cramer's rule in Fractions), for each
    row of the (count, 2) complex b: a list of [(re, im), (re, im)] pairs of Fractions."""
    m = [[_gauss(a[i, j]) for j in range(2)] for i in range(2)]
    det = _gauss_sub(_gauss_mul(m[0][0], m[1][1]), _gauss_mul(m[0][1], m[1][0]))
    out = []
    for row in b:
        b0, b1 = _gauss(row[0]), _gauss(row[1])
        x0 = _gauss_div(_gauss_sub(_gauss_mul(b0, m[1][1]), _gauss_mul(m[0][1], b1)), det)
        x1 = _gauss_div(_gauss_sub(_gauss_mul(m[0][0], b1), _gauss_mul(m[1][0], b0)), det)
        out.append([x0, x1])
    return out


def rounded(exact) -> np.ndarray:
    """exact_solve_2x2's solutions rounded to the nearest doubles, part by part, as (count, 2)
    complex128 (float() of a Fraction rounds correctly)."""
    x = np.empty((len(exact), 2), dtype=np.complex128)
    for k, (x0, x1) in enumerate(exact):
        x[k, 0] = complex(float(x0[0]), float(x0[1]))
        x[k, 1] = complex(float(x1[0]), float(x1[1]))
    return x


def gaussian_closed_forms(count: int = 16, seed: int = 5):
    """Three 2 x 2 complex factors in natural order with Gaussian-integer entries, each solved so
    that only its last division rounds: every product and sum before it is of small integers or
    half-integers, so it is exact, and a quotient whose parts' numerators and the divisor's squared
    magnitude are exact is the correctly rounded exact quotient. Any solve in num-complex's
    arithmetic, in the reference's operation order, must return the correctly rounded exact
    solution; a scaled (Smith) division or a product with the reciprocal returns other bits for
    most of these right-hand sides.
    - "lu": L = [1, 0; 2 - i, 1], U = [3 + i, 1 + 2i; 0, 1 + i], solving A x = b: x1 = y1 / (1 + i)
      is a half-integer, x0 = (y0 - (1 + 2i) x1) / (3 + i) rounds once.
    - "lu_transpose": L = I, U = [1 + i, 2 - 3i; 0, 3 + i], solving A^T x = b: x0 = b0 / (1 + i)
      is a half-integer, x1 = (b1 - (2 - 3i) x0) / (3 + i) rounds once.
    - "cholesky": L = [3, 0; 2 - i, 1] (a real diagonal), solving L L^H x = b with b0's parts
      multiples of 3: y0 = b0 / 3 and x1 = y1 are integers, x0 = (y0 - (2 + i) x1) / 3 rounds once.
    Each case: the factor's arrays as the uploads take them, the dense A (or A^T) the solve
    inverts, the (count, 2) right-hand sides, the exact solutions, the divisor of the rounded
    component and that component's index."""
    rng = np.random.default_rng(seed)

    def gaussian(k, scale=1):
        parts = rng.integers(-9, 10, (k, 2))
        return scale * (parts[:, 0] + 1j * parts[:, 1]).astype(np.complex128)

    cases = {}
    l10, u00, u01, u11 = 2 - 1j, 3 + 1j, 1 + 2j, 1 + 1j
    b = np.stack([gaussian(count), gaussian(count)], axis=1)
    lo = np.array([[1, 0], [l10, 1]], dtype=np.complex128)
    up = np.array([[u00, u01], [0, u11]], dtype=np.complex128)
    cases["lu"] = dict(
        arrays=dict(l_ptr=[0, 2, 3], l_idx=[0, 1, 1], l_val=np.array([1, l10, 1], np.complex128),
                    u_ptr=[0, 1, 3], u_idx=[0, 0, 1], u_val=np.array([u00, u01, u11], np.complex128),
                    row_perm=[0, 1], col_perm=[0, 1]),
        system=lo @ up, b=b, divisor=u00, last=0)
    u00, u01, u11 = 1 + 1j, 2 - 3j, 3 + 1j
    b = np.stack([gaussian(count), gaussian(count)], axis=1)
    up = np.array([[u00, u01], [0, u11]], dtype=np.complex128)
    cases["lu_transpose"] = dict(
        arrays=dict(l_ptr=[0, 1, 2], l_idx=[0, 1], l_val=np.array([1, 1], np.complex128),
                    u_ptr=[0, 1, 3], u_idx=[0, 0, 1], u_val=np.array([u00, u01, u11], np.complex128),
                    row_perm=[0, 1], col_perm=[0, 1]),
        system=up.T, b=b, divisor=u11, last=1)
    l00, l10, l11 = 3 + 0j, 2 - 1j, 1 + 0j
    b = np.stack([gaussian(count, 3), gaussian(count)], axis=1)
    lo = np.array([[l00, 0], [l10, l11]], dtype=np.complex128)
    cases["cholesky"] = dict(
        arrays=dict(l_ptr=[0, 1, 3], l_idx=[0, 0, 1], l_val=np.array([l00, l10, l11], np.complex128),
                    perm=[0, 1]),
        system=lo @ lo.conj().T, b=b, divisor=l00, last=0)
    for case in cases.values():
        case["exact"] = exact_solve_2x2(case["system"], case["b"])
    return cases
