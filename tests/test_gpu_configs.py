"""Every configuration at its own size on the GPU, against the oracle on the same inputs. The
reference checks its GPU path on real-size factors the same way (B' and Jacobian factors with
100 right-hand sides, ybus_powerflow/tests/cuda_factor_test.rs:93-136; PTDF and outage workloads,
cuda_dc_test.rs:184-236,312-415), there to 1e-10; here the FP64 paths give the CPU's bits, so
every solution is compared byte for byte, and the refined paths against bounds derived below.

- 2,000-bus DC solve (one right-hand side) and the 10,000-bus PTDF block: all 13,500 pair
  columns on the autotuned configuration, each compared with the CPU's solve of its dense vector.
- 25,000-bus Jacobian, COLAMD LU, 512 right-hand sides: FP64 bit for bit; FP32 sweeps refined in
  FP64: each column's reported ratio is the CPU's relative residual of the returned x, it reaches
  1e-12 or the FP64 floor (twice the CPU FP64 solution's own residual, or 32 u || |A| |x| || /
  || b ||), and the distance to the FP64 solution is within ||A^-1||_inf times the two residuals.
- 256-grid batch: every factor's block bit for bit in FP64, every column refined to its floor.
- 70,000-bus adjoint, 64 scenarios: theta, b_bar and A_bar bit for bit against the CPU
  composition, the FP32-refined gradients within 1e-8, and central differences (h = 1e-6) on
  200 injections and 200 stored values of B'.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
import ybus_amd as Y
from oracle import configs
from tests.gpu_common import gpu  # noqa: F401

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))
U = 2.0**-53


@pytest.fixture(autouse=True)
def library_defaults(monkeypatch):
    """The library's own path choice, schedules and tuning (what the benchmark runs)."""
    for name in list(os.environ):
        if name.startswith("YBUS_HIP_") and name != "YBUS_HIP_ALLOW_NO_DEVICE":
            monkeypatch.delenv(name)


def residual_floor(a, x, b):
    """32 unit roundoffs of || |A| |x| ||inf / ||b||inf: the rounding of A x itself, below which
    no FP64 solution's relative residual can be computed."""
    abs_ax = O.Matrix(a.rows, a.cols, a.row_major, a.ptr, a.idx, np.abs(a.data)).mul_vec(np.abs(x))
    return 32 * U * float(np.max(abs_ax)) / float(np.max(np.abs(b)))


def test_dc_power_flow_2000_buses(gpu):
    grid, bprime, chol, rhs = configs.dc_single()
    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    got = factor.solve_many(rhs)
    assert got.tobytes() == chol.solve_many(rhs).tobytes()
    assert bprime.relative_residual(got[0], rhs[0]) <= max(1e-12, residual_floor(bprime, got[0], rhs[0]))


def test_ptdf_block_10k_every_column_bit_for_bit(gpu):
    grid, bprime, chol, plus, minus = configs.ptdf()
    n, count = chol.n, len(plus)
    assert count == 13500
    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    factor.set_block_width(count)
    pairs = [Y.Pair(None if p < 0 else int(p), None if m < 0 else int(m)) for p, m in zip(plus, minus)]
    factor.autotune(count, pairs)
    factor.stage_pairs(pairs)
    with factor.solve_staged(Y.FP64) as block:
        assert block.columns() == count
        got = block.solutions()
    step = 1500
    for start in range(0, count, step):
        stop = min(count, start + step)
        rhs = configs.dense_pairs(plus, minus, n, start, stop)
        want = chol.solve_many(rhs, THREADS)
        assert got[start:stop].tobytes() == want.tobytes(), (start, stop)
        for k in range(0, stop - start, 250):
            assert bprime.relative_residual(got[start + k], rhs[k]) <= max(
                1e-12, residual_floor(bprime, got[start + k], rhs[k]))
    # and the same columns through pair_solves (the reference's own entry point)
    with factor.pair_solves(pairs[:2048]) as block:
        assert block.solutions().tobytes() == got[:2048].tobytes()


def inverse_norm_inf(lu, chunk: int = 2048) -> float:
    """||A^-1||_inf, the largest row sum of |A^-1|, from the CPU's transposed solves of every unit
    vector (column j of A^-T is row j of A^-1)."""
    n = lu.n
    best = 0.0
    for start in range(0, n, chunk):
        stop = min(n, start + chunk)
        e = np.zeros((stop - start, n))
        e[np.arange(stop - start), np.arange(start, stop)] = 1.0
        rows = lu.solve_transpose_many(e, THREADS)
        best = max(best, float(np.max(np.sum(np.abs(rows), axis=1))))
    return best


def test_jacobian_25k_fp64_and_fp32_refined(gpu):
    jac, lu, rhs = configs.jacobian()
    count, n = rhs.shape
    factor = Y.GpuFactor.from_lu(gpu, lu)
    factor.set_matrix(jac)
    factor.set_block_width(count)
    factor.autotune(count)
    factor.stage(rhs)
    cpu = lu.solve_many(rhs, THREADS)
    with factor.solve_staged(Y.FP64) as block:
        x64 = block.solutions()
    assert x64.tobytes() == cpu.tobytes()
    with factor.solve_staged(Y.FP32_REFINED, 1e-12, 10) as block:
        x32 = block.solutions()
        steps, ratio = block.refinement()
    assert np.all((steps >= 1) & (steps <= 10))
    # ||x32 - x64||inf = ||A^-1 (r64 - r32)||inf <= ||A^-1||inf (||r32||inf + ||r64||inf + e), with
    # r the computed residuals b - fl(A x) and e the rounding of the two residuals themselves
    # (at most the floor's 32 u || |A| |x| || each)
    inv = inverse_norm_inf(lu)
    abs_a = O.Matrix(jac.rows, jac.cols, jac.row_major, jac.ptr, jac.idx, np.abs(jac.data))
    worst = 0.0
    for k in range(count):
        r32 = rhs[k] - jac.mul_vec(x32[k])
        r64 = rhs[k] - jac.mul_vec(cpu[k])
        bnorm = float(np.max(np.abs(rhs[k])))
        assert ratio[k] == jac.relative_residual(x32[k], rhs[k])
        floor = float(np.max(np.abs(r64))) / bnorm
        assert ratio[k] <= max(1e-12, 2.0 * floor, residual_floor(jac, x32[k], rhs[k])), (k, ratio[k], floor)
        e = 32 * U * (float(np.max(abs_a.mul_vec(np.abs(x32[k])))) + float(np.max(abs_a.mul_vec(np.abs(cpu[k])))))
        bound = inv * (float(np.max(np.abs(r32))) + float(np.max(np.abs(r64))) + e)
        diff = float(np.max(np.abs(x32[k] - cpu[k])))
        assert diff <= bound, (k, diff, bound)
        worst = max(worst, diff / float(np.max(np.abs(cpu[k]))))
    print(f"jacobian 25k: ||A^-1||inf = {inv:.3e}, FP32-refined vs FP64 max relative difference {worst:.3e}, "
          f"mean corrections {steps.mean():.2f}", flush=True)


def test_batch_of_256_grids(gpu):
    case = configs.batch()
    factors = []
    batch = Y.GpuBatch.new(gpu)
    for chol, m in zip(case.factors, case.matrices):
        f = Y.GpuFactor.from_cholesky(gpu, chol)
        f.set_matrix(m)
        batch.add(f)
        factors.append(f)
    assert batch.factors() == 256
    batch.stage(case.rhs)
    batch.solve_staged(Y.FP64)
    got = [x.copy() for x in batch.solutions()]
    want = [chol.solve_many(b, THREADS) for chol, b in zip(case.factors, case.rhs)]
    for i, (x, w) in enumerate(zip(got, want)):
        assert x.shape == w.shape
        assert x.tobytes() == w.tobytes(), i
    batch.solve_staged(Y.FP32_REFINED, 1e-12, 10)
    refined = batch.solutions()
    worst = 0.0
    for i, ((steps, ratio), x, w, m, b) in enumerate(zip(batch.refinement(), refined, want, case.matrices, case.rhs)):
        for k in range(len(b)):
            assert ratio[k] == m.relative_residual(x[k], b[k]), (i, k)
            # 1e-12, or the FP64 floor: twice the CPU FP64 solution's own residual, or the
            # rounding of A x itself (refine_normwise stops where a step fails to halve the
            # ratio, anywhere at that floor)
            floor = max(2.0 * m.relative_residual(w[k], b[k]), residual_floor(m, x[k], b[k]))
            assert ratio[k] <= max(1e-12, floor), (i, k, ratio[k], floor)
            assert 1 <= steps[k] <= 10
        worst = max(worst, float(np.max(np.abs(x - w)) / np.max(np.abs(w))))
    # a replay after the refined one gives the FP64 bits again
    batch.solve_staged(Y.FP64)
    for i, (x, w) in enumerate(zip(batch.solutions(), want)):
        assert x.tobytes() == w.tobytes(), i
    print(f"batch: FP32-refined vs FP64 max relative difference {worst:.3e}", flush=True)


@pytest.fixture(scope="module")
def adjoint_case(gpu):
    case = configs.adjoint()
    bf = sp.csr_matrix((case.bf.data, case.bf.idx, case.bf.ptr), shape=case.bf.shape)
    factor = Y.GpuFactor.from_cholesky(gpu, case.chol)
    factor.set_matrix(case.bprime)
    factor.set_block_width(case.p.shape[0])
    factor.autotune(case.p.shape[0])
    theta = factor.solve_many(case.p)
    losses, x_bar = configs.flow_loss(bf, theta, case.limit)
    b_bar, a_bar = factor.adjoint(x_bar, theta)
    return dict(case=case, bf=bf, factor=factor, theta=theta, losses=losses, x_bar=x_bar, b_bar=b_bar,
                a_bar=a_bar)


def test_adjoint_70k_bit_for_bit_and_refined(adjoint_case):
    c = adjoint_case
    case, factor = c["case"], c["factor"]
    assert c["theta"].tobytes() == case.chol.solve_many(case.p, THREADS).tobytes()
    want_b, want_a = O.adjoint(case.chol, case.bprime, c["x_bar"], c["theta"])
    assert c["b_bar"].tobytes() == want_b.tobytes()
    assert c["a_bar"].tobytes() == want_a.tobytes()
    # FP32 sweeps refined in FP64, forward and transposed: within 1e-8 of the FP64 gradients
    theta32, _, ratio = factor.solve_many_refined(case.p, 1e-12, 10)
    b32, a32 = factor.adjoint(c["x_bar"], theta32, Y.FP32_REFINED, 1e-12, 10)
    rel = lambda got, want: float(np.max(np.abs(got - want)) / np.max(np.abs(want)))  # noqa: E731
    assert rel(theta32, c["theta"]) <= 1e-8
    assert rel(b32, want_b) <= 1e-8
    assert rel(a32, want_a) <= 1e-8
    print(f"adjoint 70k FP32-refined: theta {rel(theta32, c['theta']):.2e}, b_bar {rel(b32, want_b):.2e}, "
          f"a_bar {rel(a32, want_a):.2e}", flush=True)


H = 1e-6


def test_adjoint_70k_injection_gradient_by_central_differences(adjoint_case):
    """dL/dp_s = b_bar_s. Perturbing p_s[i] moves only scenario s's flows, so the central
    difference of the total loss is the difference of scenario s's own loss, evaluated apart.
    The loss is piecewise quadratic in p, so the difference has no truncation error away from a
    kink; what remains is the solves' and the loss's rounding over 2h, measured on the CPU at
    this size at 1.3e-6 of the scenario's largest gradient entry at worst. The bar is 1e-5 of it
    plus 1e-6 of the entry; a wrong sign, index or scenario is off by O(1)."""
    c = adjoint_case
    case, chol, bf, b_bar = c["case"], c["case"].chol, c["bf"], c["b_bar"]
    rng = np.random.default_rng(20261207)
    for _ in range(200):
        s, i = int(rng.integers(case.p.shape[0])), int(rng.integers(case.p.shape[1]))
        values = []
        for sign in (1.0, -1.0):
            q = case.p[s].copy()
            q[i] += sign * H
            values.append(configs.flow_loss(bf, chol.solve(q)[None, :], case.limit, False)[0][0])
        fd = (values[0] - values[1]) / (2 * H)
        bar = 1e-6 * abs(b_bar[s, i]) + 1e-5 * float(np.max(np.abs(b_bar[s])))
        assert abs(fd - b_bar[s, i]) <= bar, (s, i, fd, b_bar[s, i])


def test_adjoint_70k_susceptance_gradient_by_central_differences(adjoint_case):
    """dL/dA_q = a_bar_q for 200 stored entries q of B' perturbed one at a time (the perturbed
    matrix is unsymmetric, so it is factored by the oracle's LU with AMD, refactored with its
    pivots for each perturbation). The loss is smooth in A away from kinks, so the difference's
    error is the rounding of 64 solves of the perturbed systems over 2h, measured on the CPU at
    this size at 2.4e-7 of the largest gradient entry at worst: the bar is 2e-6 of it plus 1e-6
    of the entry. (That noise is of the order of the difference between a_bar and its transposition
    on a symmetric B', so the transposition is ruled out by the bit-exact CPU composition and its
    unsymmetric dense checks, not here; a wrong sign, scale or entry is off by O(1).)"""
    c = adjoint_case
    case, bf, a_bar = c["case"], c["bf"], c["a_bar"]
    bp = case.bprime
    n = bp.rows
    rows = np.repeat(np.arange(n), np.diff(bp.ptr))
    cols = np.asarray(bp.idx, dtype=np.int64)
    csc = bp.to_csc()
    # position of each CSR entry in the CSC arrays (columns ascending, rows ascending in each)
    order = np.lexsort((rows, cols))
    where = np.empty_like(order)
    where[order] = np.arange(len(order))
    assert np.array_equal(np.asarray(csc.idx)[where], rows)
    lu = O.Lu(csc, O.AMD)
    scale = float(np.max(np.abs(a_bar)))
    rng = np.random.default_rng(20261208)
    for q in rng.choice(bp.nnz, size=200, replace=False):
        values = []
        for sign in (1.0, -1.0):
            data = np.array(csc.data, copy=True)
            data[where[q]] += sign * H
            lu.refactor(O.Matrix(n, n, False, csc.ptr, csc.idx, data), read=False)
            values.append(float(np.sum(configs.flow_loss(bf, lu.solve_many(case.p, THREADS), case.limit, False)[0])))
        fd = (values[0] - values[1]) / (2 * H)
        bar = 1e-6 * abs(a_bar[q]) + 2e-6 * scale
        assert abs(fd - a_bar[q]) <= bar, (q, fd, a_bar[q])
