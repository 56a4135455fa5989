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
