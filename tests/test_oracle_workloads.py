"""The synthetic workload generator and the restated B' and AC Jacobian builders
(ybus_powerflow/src/dc.rs:79-112, ac.rs:137-151, 1311-1427), pinned by their defining
properties: connectivity and branch counts, brute-force bridges, B' as a reduced Laplacian, and
the Jacobian against central finite differences of the power-flow mismatch."""
import math

import numpy as np

import oracle as O


def components(n, edges):
    parent = list(range(n))
    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    for i, j in edges:
        parent[find(i)] = find(j)
    return len({find(i) for i in range(n)})


def test_grid_shape_and_reactances():
    for n in [2, 3, 50, 400, 2000]:
        g = O.make_grid(n, 7 + n)
        assert g.branches == int(math.floor(1.35 * n + 0.5)) or n < 4
        assert np.all(g.branch_from < g.branch_to)
        pairs = list(zip(g.branch_from, g.branch_to))
        assert pairs == sorted(set(pairs))
        assert components(n, pairs) == 1
        assert np.all((g.x >= 0.005) & (g.x <= 0.2))
    a, b = O.make_grid(300, 11), O.make_grid(300, 11)
    assert a.x.tobytes() == b.x.tobytes() and np.array_equal(a.branch_from, b.branch_from)


def test_bridges_match_brute_force():
    g = O.make_grid(120, 5)
    pairs = list(zip(g.branch_from, g.branch_to))
    flags = g.bridges()
    for k in range(g.branches):
        rest = pairs[:k] + pairs[k + 1:]
        assert flags[k] == (components(g.buses, rest) > 1)
    assert flags.any() and not flags.all()


def test_bprime_is_the_reduced_susceptance_laplacian():
    g = O.make_grid(200, 3)
    bp, bf, rows = g.dc()
    n = g.buses
    full = np.zeros((n, n))
    flows = np.zeros((g.branches, n))
    for k, (f, t) in enumerate(zip(g.branch_from, g.branch_to)):
        b = 1.0 / g.x[k]
        full[f, f] += b; full[t, t] += b; full[f, t] -= b; full[t, f] -= b
        flows[k, f] = b; flows[k, t] = -b
    assert np.allclose(bp.to_dense(), full[1:, 1:], rtol=1e-15, atol=0)
    assert np.allclose(bf.to_dense(), flows[:, 1:], rtol=1e-15, atol=0)
    assert list(rows) == list(range(g.branches))
    # exactly symmetric (cholesky requires it) and positive definite
    d = bp.to_dense()
    assert np.array_equal(d, d.T)
    assert np.all(np.linalg.eigvalsh(d) > 0)
    # an outage drops its branch's stamps
    mask = np.ones(g.branches, bool)
    k = int(np.flatnonzero(~g.bridges())[0])
    mask[k] = False
    bp2, bf2, rows2 = g.dc(mask)
    assert bf2.rows == g.branches - 1 and k not in set(rows2)
    f, t = g.branch_from[k], g.branch_to[k]
    b = 1.0 / g.x[k]
    expect = full.copy()
    expect[f, f] -= b; expect[t, t] -= b; expect[f, t] += b; expect[t, f] += b
    assert np.allclose(bp2.to_dense(), expect[1:, 1:], rtol=1e-12, atol=1e-9)


def test_injections_are_zero_sum_with_the_slack_dropped():
    p = O.injections(99, 4, 50, 3)
    assert p.shape == (3, 49)
    z = O.normals(99, 4, 150).reshape(3, 50)
    full = z - z.sum(axis=1, keepdims=True) / 50
    assert np.allclose(p, full[:, 1:], rtol=0, atol=1e-15)
    u = O.uniforms(1, 2, 1000)
    assert 0.4 < u.mean() < 0.6 and u.min() >= 0 and u.max() < 1


def test_jacobian_matches_finite_differences_of_the_mismatch():
    g = O.make_grid(60, 13)
    jac, op = g.jacobian(13)
    n = g.buses
    y = np.zeros((n, n), complex)
    for k, (f, t) in enumerate(zip(g.branch_from, g.branch_to)):
        ys = 1.0 / complex(op["r"][k], g.x[k])
        ytt = ys + 1j * op["bc"][k] / 2
        y[f, f] += ytt; y[t, t] += ytt; y[f, t] -= ys; y[t, f] -= ys
    pv, pq = list(op["pv"]), list(op["pq"])
    pvpq = pv + pq
    assert sorted(pvpq) == list(range(1, n)) and len(pv) == round(0.2 * (n - 1))
    assert np.all((op["vm"] >= 0.95) & (op["vm"] <= 1.05)) and op["va"][0] == 0.0

    def mismatch(theta, vm):
        v = vm * np.exp(1j * theta)
        s = v * np.conj(y @ v)
        return np.concatenate([s.real[pvpq], s.imag[pq]])

    h = 1e-6
    cols = []
    for k, bus in enumerate(pvpq + pq):
        th, vm = op["va"].copy(), op["vm"].copy()
        target = th if k < len(pvpq) else vm
        target[bus] += h
        up = mismatch(th, vm)
        target[bus] -= 2 * h
        down = mismatch(th, vm)
        cols.append((up - down) / (2 * h))
    fd = np.array(cols).T
    dense = jac.to_dense()
    assert dense.shape == fd.shape == (len(pvpq) + len(pq),) * 2
    assert np.max(np.abs(dense - fd)) <= 1e-6 * np.max(np.abs(fd))
    # the pattern is Ybus's, stored zeros included: every stamped position is stored
    assert jac.nnz >= np.count_nonzero(dense)
