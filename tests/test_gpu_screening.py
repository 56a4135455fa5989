"""The screening caller's suspect-column contract (ybus_powerflow/src/dc/gpu.rs:386-430): an outage
whose update is singular, whose angles are not finite or exceed 1e5, or whose loadings are not
finite goes to the caller's own solve, which decides its result; every other outage's most
loaded branch is the device's. The workload takes every branch of a 3,000-bus grid out alone,
bridges included (their removal islands a bus, so the update is singular or its angles blow up),
and plants three more suspects through the inputs: an update whose weight cancels its
coefficient exactly (singular), one that leaves it a hair from zero (angles past 1e5), and one
with a NaN alpha (angles not finite). The suspect set must be exactly those, and it must be the
set the CPU composition (the oracle's solves and the same update formula) flags."""
import numpy as np
import pytest

import oracle as O
import ybus_amd as Y
from oracle import configs
from tests.gpu_common import gpu  # noqa: F401
from ybus_amd import screening

pytestmark = pytest.mark.gpu


def cpu_updates(chol, theta, plus, minus, alpha, inverse_weight, rows):
    """The rank-one updates on the CPU as the device forms them (pairs.rs:209-253): y = B'^-1 c,
    w = inverse_weight - c^T y, z = c^T x / w with x = theta + alpha y, and x + z y (x alone when w
    is 0): (updated angles, w)."""
    dense = configs.dense_pairs(plus, minus, chol.n)[rows]
    y = chol.solve_many(dense, 4)
    p, m = plus[rows], minus[rows]

    def at(v, idx):
        return np.where(idx >= 0, v[np.arange(len(idx)), np.maximum(idx, 0)], 0.0)

    x = theta[None, :] + alpha[rows][:, None] * y
    w = inverse_weight[rows] - (at(y, p) - at(y, m))
    with np.errstate(divide="ignore", invalid="ignore"):
        z = (at(x, p) - at(x, m)) / w
        updated = np.where((w == 0.0)[:, None], x, x + z[:, None] * y)
    return updated, w


def bf_product(bf, x):
    out = np.zeros((x.shape[0], bf.rows))
    ptr, idx, val = np.asarray(bf.ptr), np.asarray(bf.idx), np.asarray(bf.data)
    for slot in range(int(np.max(np.diff(ptr)))):
        has = np.diff(ptr) > slot
        at = ptr[:-1][has] + slot
        out[:, has] += val[at][None, :] * x[:, idx[at]]
    return out


def test_suspect_outages_go_to_the_callers_solve(gpu):
    grid = O.make_grid(3000, 4242)
    bprime, bf, _ = grid.dc()
    chol = O.Cholesky(bprime)
    theta = chol.solve(O.injections(4242, 4, 3000, 1)[0])
    plus, minus = configs.pair_lists(grid)
    inverse_weight = np.asarray(grid.x, dtype=np.float64).copy()
    alpha = np.zeros(grid.branches)
    bridges = grid.bridges()
    assert bridges.any() and (~bridges).sum() > 100
    # three planted suspects among the branches whose removal leaves the grid connected
    inner = np.flatnonzero(~bridges & (plus >= 0) & (minus >= 0))
    j_singular, j_large, j_nan = (int(v) for v in inner[[5, 17, 29]])
    y = chol.solve_many(configs.dense_pairs(plus, minus, chol.n)[[j_singular, j_large]], 2)
    cty = [y[0, plus[j_singular]] - y[0, minus[j_singular]], y[1, plus[j_large]] - y[1, minus[j_large]]]
    inverse_weight[j_singular] = cty[0]            # w = 0 exactly
    inverse_weight[j_large] = cty[1] * (1 + 1e-9)  # w a hair from zero
    alpha[j_nan] = np.nan
    planted = {j_singular, j_large, j_nan}
    limit = 0.5 + 1.5 * O.uniforms(4242, 5, bf.rows)
    weights = 1.0 / limit
    zero = np.arange(grid.branches, dtype=np.int64)

    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    factor.set_block_width(700)  # several blocks, the last one partial
    op = factor.operator(bf)
    op.set_offset(np.zeros(bf.rows))
    op.set_weights(weights)
    resolved = []

    def resolve(j):
        resolved.append(j)
        return ("cpu", j)

    results = screening.outage_loadings(factor, op, theta, plus, minus, alpha, inverse_weight, zero, resolve)
    assert len(results) == grid.branches
    expected = set(np.flatnonzero(bridges).tolist()) | planted
    assert set(resolved) == expected and len(resolved) == len(expected)
    for j in resolved:
        assert results[j] == ("cpu", j)

    # the CPU composition flags the same outages under the same rule
    everything = np.arange(grid.branches)
    updated, w = cpu_updates(chol, theta, plus, minus, alpha, inverse_weight, everything)
    with np.errstate(invalid="ignore", over="ignore"):
        flows = bf_product(bf, updated)
    flows[everything, zero] = 0.0
    with np.errstate(invalid="ignore", over="ignore"):
        loading = np.abs(flows) * weights[None, :]
    largest = np.where(np.isfinite(updated), np.abs(updated), 0.0).max(axis=1)
    cpu_suspect = ((w == 0.0) | ~np.isfinite(updated).all(axis=1) | (largest > screening.SUSPECT_ANGLE)
                   | ~np.isfinite(loading).all(axis=1))
    assert set(np.flatnonzero(cpu_suspect).tolist()) == expected
    assert w[j_singular] == 0.0 and np.abs(updated[j_large]).max() > screening.SUSPECT_ANGLE

    # every other outage: the device's most loaded branch and loading, as the CPU's
    worst = 0.0
    for j in np.flatnonzero(~cpu_suspect):
        got = results[j]
        assert isinstance(got, screening.OutageLoading), j
        want = loading[j].max()
        worst = max(worst, abs(got.loading - want) / want)
        if got.branch != int(np.argmax(loading[j])):
            assert abs(loading[j][got.branch] - want) <= 1e-12 * want, j
    assert worst <= 1e-12
