"""The synthetic workloads of the five configurations, built on the host as ybus_sparse would
build them (test infrastructure, like the rest of this package: the benchmark's setup and the GPU
tests take their inputs from here, and the timed GPU region never calls it).

Each builder is deterministic: the grids come from the oracle's generator (``make_grid``), the
random draws from its seeded streams, so the benchmark and the tests see the same matrices,
factors and right-hand sides.

- ``dc_single``: 2,000 buses, AMD Cholesky of B', one zero-sum N(0,1) injection vector.
- ``ptdf``: 10,000 buses, AMD Cholesky of B', one pair right-hand side per branch (+1 at the from
  bus, -1 at the to bus, the slack bus 0 dropped), as ``pair_solves`` takes them
  (sensitivity.rs:212,355; dc/gpu.rs:506-611).
- ``jacobian``: 25,000 buses, the polar AC Newton Jacobian (ac.rs:1297-1427), COLAMD LU with
  threshold 0.1, 512 N(0,1) right-hand sides.
- ``batch``: 256 grids, sizes log-uniform in [500, 20,000] buses, each with one random non-bridge
  branch out of service, AMD Cholesky of each B', r_i in 1..64 zero-sum N(0,1) scenarios.
- ``adjoint``: 70,000 buses, AMD Cholesky of B', 64 zero-sum N(0,1) scenarios, branch limits
  U[0.5, 2.0] pu, loss = sum relu(|Bf theta| - limit)^2.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

import oracle as O

SEED = 20261203


def pair_lists(grid):
    """One pair per branch: +1 at its from bus, -1 at its to bus, bus 0 (the slack) dropped, as
    (plus, minus) unknown indices with -1 for none."""
    plus = np.array([-1 if f == 0 else int(f) - 1 for f in grid.branch_from], dtype=np.int64)
    minus = np.array([-1 if t == 0 else int(t) - 1 for t in grid.branch_to], dtype=np.int64)
    return plus, minus


def dense_pairs(plus, minus, n: int, start: int = 0, stop: int | None = None) -> np.ndarray:
    """Pairs start..stop as dense vectors (what the CPU path solves)."""
    stop = len(plus) if stop is None else stop
    out = np.zeros((stop - start, n))
    k = np.arange(stop - start)
    p, m = plus[start:stop], minus[start:stop]
    out[k[p >= 0], p[p >= 0]] = 1.0
    out[k[m >= 0], m[m >= 0]] = -1.0
    return out


def dc_single(buses: int = 2000, seed: int = SEED):
    grid = O.make_grid(buses, seed)
    bprime, bf, _ = grid.dc()
    chol = O.Cholesky(bprime)
    rhs = O.injections(seed, 4, buses, 1)
    return grid, bprime, chol, rhs


def ptdf(buses: int = 10000, seed: int = SEED + 1):
    """(grid, B', Cholesky, plus, minus)."""
    grid = O.make_grid(buses, seed)
    bprime, _, _ = grid.dc()
    chol = O.Cholesky(bprime)
    plus, minus = pair_lists(grid)
    return grid, bprime, chol, plus, minus


def jacobian(buses: int = 25000, seed: int = SEED + 2, count: int = 512):
    """(Jacobian CSC, COLAMD LU, right-hand sides (count, n))."""
    grid = O.make_grid(buses, seed)
    jac, _ = grid.jacobian(seed)
    lu = O.Lu(jac, O.COLAMD)
    rhs = O.normals(seed, 9, count * jac.rows).reshape(count, jac.rows)
    return jac, lu, rhs


@dataclass
class BatchCase:
    sizes: np.ndarray          # buses per grid
    counts: np.ndarray         # right-hand sides per grid
    matrices: list             # each grid's B' (CSR)
    factors: list              # each grid's AMD Cholesky
    rhs: list                  # each grid's (counts[i], n_i) injections


def batch(grids: int = 256, seed: int = SEED + 3) -> BatchCase:
    sizes = np.round(np.exp(np.log(500.0) + O.uniforms(seed, 1, grids) * np.log(20000.0 / 500.0))).astype(int)
    counts = (1 + O.integers(seed, 2, grids) % 64).astype(int)
    picks = O.uniforms(seed, 3, grids)
    mats, factors, rhs = [], [], []
    for i in range(grids):
        g = O.make_grid(int(sizes[i]), seed * 1000 + i)
        candidates = np.flatnonzero(~g.bridges())
        mask = np.ones(g.branches, dtype=bool)
        mask[candidates[int(picks[i] * len(candidates))]] = False
        bprime, _, _ = g.dc(mask)
        mats.append(bprime)
        factors.append(O.Cholesky(bprime))
        rhs.append(O.injections(seed * 1000 + i, 4, int(sizes[i]), int(counts[i])))
    return BatchCase(sizes, counts, mats, factors, rhs)


@dataclass
class AdjointCase:
    grid: object
    bprime: object             # B' (CSR)
    bf: object                 # Bf (CSR, branches x unknowns)
    chol: object
    p: np.ndarray              # (64, n) injections
    limit: np.ndarray          # one limit per Bf row


def adjoint(buses: int = 70000, seed: int = SEED + 4, scenarios: int = 64) -> AdjointCase:
    grid = O.make_grid(buses, seed)
    bprime, bf, _ = grid.dc()
    chol = O.Cholesky(bprime)
    p = O.injections(seed, 4, buses, scenarios)
    limit = 0.5 + 1.5 * O.uniforms(seed, 5, bf.rows)
    return AdjointCase(grid, bprime, bf, chol, p, limit)


def flow_loss(bf_csr, x: np.ndarray, limit: np.ndarray, gradient: bool = True):
    """Per-scenario losses sum_k relu(|f_k| - limit_k)^2 with f = Bf x (scipy product), and (when
    `gradient`) the gradient with respect to the solutions, x_bar = Bf^T (2 relu(|f| - limit)
    sign(f)). bf_csr is a scipy.sparse matrix; x is (scenarios, n)."""
    f = np.asarray(bf_csr @ np.ascontiguousarray(x.T))  # branches x scenarios
    over = np.maximum(np.abs(f) - limit[:, None], 0.0)
    losses = np.sum(over * over, axis=0)
    if not gradient:
        return losses, None
    x_bar = np.ascontiguousarray(np.asarray(bf_csr.T @ (2.0 * over * np.sign(f))).T)
    return losses, x_bar


def supernode_widths(n: int, ptr, idx):
    """The fundamental supernodes of a lower triangular factor stored by columns (CSC, each
    column's rows ascending, the diagonal first), as (widths, stored entries of each): column
    j + 1 continues column j's supernode when j + 1 is j's parent in the elimination tree (the
    first row below j's diagonal), j + 1 has no other child, and column j + 1's rows below its
    diagonal are column j's rows below j + 1 (multifrontal.rs:94-138's rule, which Cholesky's
    supernodes() restates; this form also reads an LU factor's L, or U's rows as the columns of
    U^T)."""
    ptr = np.asarray(ptr, dtype=np.int64)
    idx = np.asarray(idx, dtype=np.int64)
    parent = np.full(n, -1, dtype=np.int64)
    below = [None] * n
    for j in range(n):
        rows = idx[ptr[j]:ptr[j + 1]]
        rows = rows[rows > j]
        below[j] = rows
        if len(rows):
            parent[j] = rows.min()
    children = np.bincount(parent[parent >= 0], minlength=n)
    starts = [0]
    for j in range(n - 1):
        a, b = below[j], below[j + 1]
        joined = (parent[j] == j + 1 and children[j + 1] == 1 and len(a) == len(b) + 1
                  and np.array_equal(a[1:], b))
        if not joined:
            starts.append(j + 1)
    widths = np.diff(np.append(starts, n))
    entries = np.add.reduceat(np.diff(ptr), starts) if n else np.zeros(0, dtype=np.int64)
    return widths, entries


def inverse_norm_1_estimate(n: int, solve, solve_transpose) -> float:
    """Hager's and Higham's estimate of ||A^-1||_1 from solves with A and A^T, keeping the
    largest estimate seen (accuracy.rs:207-279, LAPACK's xLACON): never more than the true value.
    Swapping the two solves estimates ||A^-T||_1 = ||A^-1||_inf."""
    if n == 0:
        return 0.0
    y = solve(np.full(n, 1.0 / n))
    estimate = float(np.sum(np.abs(y)))
    if n == 1:
        return estimate

    def sign(v):
        return np.where(v >= 0.0, 1.0, -1.0)

    def argmax_abs(v):
        # the first index of the largest magnitude (a strict > scan)
        return int(np.argmax(np.abs(v)))

    signs = sign(y)
    z = solve_transpose(signs)
    j = argmax_abs(z)
    for _ in range(2, 6):
        unit = np.zeros(n)
        unit[j] = 1.0
        y = solve(unit)
        previous = estimate
        current = float(np.sum(np.abs(y)))
        estimate = max(estimate, current)
        new_signs = sign(y)
        if np.array_equal(new_signs, signs) or current <= previous:
            break
        signs = new_signs
        z = solve_transpose(signs)
        last = j
        j = argmax_abs(z)
        if abs(z[last]) == abs(z[j]):
            break
    i = np.arange(n)
    alternating = np.where(i % 2 == 0, 1.0, -1.0) * (1.0 + i / (n - 1))
    extra = 2.0 * float(np.sum(np.abs(solve(alternating)))) / (3.0 * n)
    return max(estimate, extra)


def lu_inverse_norm_inf_estimate(lu) -> float:
    """An estimate of ||A^-1||_inf (= ||A^-T||_1) for an LU factor."""
    return inverse_norm_1_estimate(lu.n, lu.solve_transpose, lu.solve)


def norm_inf(a) -> float:
    """The largest absolute row sum of an oracle matrix."""
    major = np.repeat(np.arange(len(a.ptr) - 1), np.diff(np.asarray(a.ptr, dtype=np.int64)))
    rows = major if a.row_major else np.asarray(a.idx, dtype=np.int64)
    sums = np.zeros(a.rows)
    np.add.at(sums, rows, np.abs(a.data))
    return float(sums.max()) if a.rows else 0.0


def norm_1(a) -> float:
    """The largest absolute column sum of an oracle matrix (CompressedMatrix::norm_1)."""
    major = np.repeat(np.arange(len(a.ptr) - 1), np.diff(np.asarray(a.ptr, dtype=np.int64)))
    cols = np.asarray(a.idx, dtype=np.int64) if a.row_major else major
    sums = np.zeros(a.cols)
    np.add.at(sums, cols, np.abs(a.data))
    return float(sums.max()) if a.cols else 0.0
