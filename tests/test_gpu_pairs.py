"""Blocks built, solved and updated on the device against the same formulas on the host from the
oracle's CPU solves (ybus_cuda/tests/gpu_pairs_test.rs restated): pair solves and maxima,
rank-one updates including a singular one, operator products with entries zeroed and weighted
maxima, and the single-row width cap."""
import numpy as np
import pytest

import oracle as O
import ybus_amd as Y
from tests.conftest import Lcg
from tests.fixtures import random_spd, random_square_cuda
from tests.gpu_common import gpu, path  # noqa: F401

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("path")]

BAR = 1e-12


def pair_vector(n, pair):
    c = np.zeros(n)
    if pair.plus is not None:
        c[pair.plus] = 1.0
    if pair.minus is not None:
        c[pair.minus] = -1.0
    return c


def dot(pair, v):
    return (v[pair.plus] if pair.plus is not None else 0.0) - (v[pair.minus] if pair.minus is not None else 0.0)


def host_update(cpu, base, update):
    """gpu_pairs_test.rs:48-62."""
    n = len(base)
    y = cpu.solve(pair_vector(n, update.pair))
    x = base + update.alpha * y
    w = -dot(update.pair, y) + update.inverse_weight
    if w == 0.0:
        return x, True
    z = dot(update.pair, x) / w
    return x + z * y, False


def max_difference(a, b):
    scale = max(1.0, float(np.max(np.abs(b))))
    return float(np.max(np.abs(a - b))) / scale


def pairs(n, count, seed):
    """gpu_pairs_test.rs:72-84."""
    rng = Lcg(seed)
    out = []
    for k in range(count):
        plus = rng.next() % n if k % 5 != 3 else None
        minus = rng.next() % n if k % 7 != 2 else None
        if minus is not None and minus == plus:
            minus = None
        out.append(Y.Pair(plus, minus))
    return out


def operator_matrix(rows, n, seed):
    rng = Lcg(seed)
    t = []
    for i in range(rows):
        for _ in range(3):
            t.append((i, rng.next() % n, rng.value()))
    return O.Matrix.from_triplets(rows, n, t)


def factors(n):
    return [("cholesky", O.Cholesky(random_spd(n, 3, 21))), ("lu", O.Lu(random_square_cuda(n, 4, 22)))]


def upload(gpu, cpu):
    return Y.GpuFactor.from_cholesky(gpu, cpu) if isinstance(cpu, O.Cholesky) else Y.GpuFactor.from_lu(gpu, cpu)


def test_pair_solves_and_their_maxima_match_the_host(gpu):
    n = 200
    for name, cpu in factors(n):
        factor = upload(gpu, cpu)
        factor.set_block_width(32)
        ps = pairs(n, 29, 5)
        block = factor.pair_solves(ps)
        assert block.columns() == 29
        assert not any(block.singular())
        solutions = block.solutions()
        maxima = block.maxima()
        for pair, solution, mx in zip(ps, solutions, maxima):
            expected = cpu.solve(pair_vector(n, pair))
            assert max_difference(solution, expected) < BAR, name
            # the right-hand side is built as the host builds it: the same bits
            assert solution.tobytes() == expected.tobytes(), name
            largest = float(np.max(np.abs(expected)))
            assert abs(mx.value - largest) <= BAR * largest
            if mx.row is not None:
                assert abs(abs(solution[mx.row]) - mx.value) <= BAR * largest
                # the smallest unknown on a tie
                assert mx.row == int(np.flatnonzero(np.abs(solution) == mx.value)[0])
            else:
                assert pair.plus is None and pair.minus is None
            assert not mx.not_finite
        block.free()


def test_rank_one_updates_match_the_host_including_a_singular_one(gpu):
    n = 150
    rng = Lcg(9)
    base = np.array([rng.value() for _ in range(n)])
    for name, cpu in factors(n):
        factor = upload(gpu, cpu)
        factor.set_block_width(16)
        updates = [Y.PairUpdate(p, rng.value(), 2.0 + rng.value()) for p in pairs(n, 15, 6)]
        updates.append(Y.PairUpdate(Y.Pair(None, None), 0.5, 0.0))
        block = factor.pair_updates(base, updates)
        solutions = block.solutions()
        singular = block.singular()
        for k, (update, solution) in enumerate(zip(updates, solutions)):
            expected, sing = host_update(cpu, base, update)
            assert singular[k] == sing, (name, k)
            assert max_difference(solution, expected) < BAR, (name, k)
        assert singular[15]
        block.free()


def test_operator_products_and_maxima_match_the_host(gpu):
    n, rows = 120, 90
    rng = Lcg(31)
    base = np.array([rng.value() for _ in range(n)])
    offset = np.array([rng.value() for _ in range(rows)])
    weights = np.array([0.0 if i % 4 == 0 else 1.0 + abs(rng.value()) for i in range(rows)])
    m = operator_matrix(rows, n, 32)
    for name, cpu in factors(n):
        factor = upload(gpu, cpu)
        factor.set_block_width(8)
        op = factor.operator(m)
        op.set_offset(offset)
        op.set_weights(weights)
        updates = [Y.PairUpdate(p, 0.25, 3.0) for p in pairs(n, 7, 33)]
        zero = [k * 11 if k % 2 == 0 else None for k in range(7)]
        block = factor.pair_updates(base, updates)
        products = block.product(op, zero)
        maxima = block.product_maxima(op, zero)
        for k in range(7):
            theta, _ = host_update(cpu, base, updates[k])
            expected = m.mul_vec(theta) + offset
            if zero[k] is not None:
                expected[zero[k]] = 0.0
                assert products[k][zero[k]] == 0.0
            assert max_difference(products[k], expected) < BAR, (name, k)
            row, value = None, 0.0
            for i, (e, w) in enumerate(zip(products[k], weights)):
                if abs(e) * w > value:
                    row, value = i, abs(e) * w
            assert maxima[k] == Y.ColumnMax(row, value, False), (name, k)
        block.free()


def test_a_single_row_factor_keeps_its_blocks_narrow(gpu):
    lu = O.Lu(O.Matrix.from_triplets(1, 1, [(0, 0, 4.0)], row_major=False))
    factor = Y.GpuFactor.from_lu(gpu, lu)
    factor.set_block_width(1024)
    assert factor.block_width() == 64
    solutions = factor.pair_solves([Y.Pair(0, None)] * 64).solutions()
    assert all(list(s) == [0.25] for s in solutions)


def test_caller_bugs_are_invalid_arguments(gpu):
    factor = Y.GpuFactor.from_cholesky(gpu, O.Cholesky(random_spd(20, 10, 2)))
    factor.set_block_width(4)
    with pytest.raises(Y.InvalidArgument):
        factor.pair_solves([Y.Pair(0, 1)] * 5)
    with pytest.raises(Y.InvalidArgument):
        factor.pair_solves([Y.Pair(20, None)])
    with pytest.raises(Y.InvalidArgument):
        factor.pair_solves([Y.Pair(3, 3)])


@pytest.mark.parametrize("kind", ["cholesky", "lu"])
def test_pair_blocks_give_the_cpu_bits_with_and_without_the_reach(gpu, monkeypatch, kind):
    """A pair block on the wide path sweeps only each tile's reach forward; the rows it skips are
    the CPU's exact +0.0, so the bits are the CPU's either way (YBUS_HIP_REACH=0 sweeps every
    row)."""
    g = O.make_grid(1500, 4242)
    if kind == "cholesky":
        bp, _, _ = g.dc()
        cpu = O.Cholesky(bp)
        n = bp.rows
    else:
        jac, _ = g.jacobian(4242)
        cpu = O.Lu(jac, O.COLAMD)
        n = jac.rows
    rng = np.random.default_rng(3)
    k = 900
    plus = rng.integers(-1, n, size=k)
    minus = rng.integers(-1, n, size=k)
    minus[minus == plus] = -1
    dense = np.zeros((k, n))
    for c in range(k):
        if plus[c] >= 0:
            dense[c, plus[c]] = 1.0
        if minus[c] >= 0:
            dense[c, minus[c]] = -1.0
    want = cpu.solve_many(dense, threads=4)
    pairs = [Y.Pair(None if p < 0 else int(p), None if m < 0 else int(m)) for p, m in zip(plus, minus)]
    for reach in ["1", "0"]:
        monkeypatch.setenv("YBUS_HIP_REACH", reach)
        factor = Y.GpuFactor.from_cholesky(gpu, cpu) if kind == "cholesky" else Y.GpuFactor.from_lu(gpu, cpu)
        factor.set_block_width(k)
        with factor.pair_solves(pairs) as block:
            assert block.solutions().tobytes() == want.tobytes(), reach


def host_maxima(values, weights):
    """Per column (a row of `values`): the largest finite |v| * w, the first row reaching it (None
    when every product is 0) and whether any value isn't finite (pairs.rs:39-50)."""
    out = []
    for column in values:
        row, best, bad = None, 0.0, False
        for i, (v, w) in enumerate(zip(column, weights)):
            if not np.isfinite(v):
                bad = True
                continue
            if abs(v) * w > best:
                row, best = i, abs(v) * w
        out.append(Y.ColumnMax(row, best, bad))
    return out


def test_maxima_split_over_row_chunks_keep_the_first_of_tied_rows(gpu):
    """Blocks and operator products with thousands of rows: the maxima are reduced over runs of
    rows and then combined, and a tie between rows in different runs still goes to the first."""
    n, rows = 1200, 3000
    cpu = O.Cholesky(random_spd(n, 1, 41))
    rng = Lcg(43)
    triplets = []
    for i in range(rows):
        cols = [rng.next() % n for _ in range(3)]
        vals = [rng.value() for _ in range(3)]
        if i in (10, 2500):  # two identical rows, large enough to be every column's maximum
            cols, vals = [5, 6, 7], [400.0, -250.0, 125.0]
        triplets += [(i, c, v) for c, v in zip(cols, vals)]
    m = O.Matrix.from_triplets(rows, n, triplets)
    base = np.array([rng.value() for _ in range(n)])
    weights = np.ones(rows)
    updates = [Y.PairUpdate(p, 0.25, 3.0) for p in pairs(n, 9, 44)]
    zero = [None, 10, None, 7, None, None, 2500, None, 1]
    factor = Y.GpuFactor.from_cholesky(gpu, cpu)
    factor.set_block_width(16)
    op = factor.operator(m)
    op.set_offset(np.zeros(rows))
    op.set_weights(weights)
    block = factor.pair_updates(base, updates)
    theta = block.solutions()
    products = block.product(op, zero)
    maxima = block.product_maxima(op, zero)
    assert maxima == host_maxima(products, weights)
    assert maxima[0].row == 10 and maxima[1].row == 2500 and maxima[6].row == 10
    assert block.maxima() == host_maxima(theta, np.ones(n))
    # the array forms give the same results
    z = np.array([-1 if r is None else r for r in zero])
    rows_a, values_a, bad_a = block.product_maxima_arrays(op, z)
    assert [Y.ColumnMax(None if r < 0 else int(r), float(v), bool(b)) for r, v, b in zip(rows_a, values_a, bad_a)] \
        == maxima
    rows_b, values_b, _ = block.maxima_arrays()
    assert list(values_b) == [mx.value for mx in block.maxima()]
    block.free()
    plus, minus = _pairs_arrays([u.pair for u in updates])
    again = factor.pair_updates_arrays(base, plus, minus, [0.25] * 9, [3.0] * 9)
    assert again.solutions().tobytes() == theta.tobytes()
    again.free()


def _pairs_arrays(ps):
    return (np.array([-1 if p.plus is None else p.plus for p in ps]),
            np.array([-1 if p.minus is None else p.minus for p in ps]))


@pytest.mark.parametrize("kind", ["pairs", "vectors"])
def test_partitioned_sweeps_give_the_same_bits(gpu, monkeypatch, kind):
    """A Cholesky factor's elimination tree cut into independent subtrees, swept side by side and
    then the top of the tree: every row keeps its terms and their order, so the solutions are the
    one-schedule solve's bits, for pair blocks (with and without the reach) and vector blocks."""
    g = O.make_grid(3000, 777)
    bp, _, _ = g.dc()
    cpu = O.Cholesky(bp)
    n = bp.rows
    rng = np.random.default_rng(8)
    k = 300
    if kind == "pairs":
        plus = rng.integers(-1, n, size=k)
        minus = rng.integers(-1, n, size=k)
        minus[minus == plus] = -1
        pairs = [Y.Pair(None if p < 0 else int(p), None if m < 0 else int(m)) for p, m in zip(plus, minus)]
        dense = np.zeros((k, n))
        for c in range(k):
            if plus[c] >= 0:
                dense[c, plus[c]] = 1.0
            if minus[c] >= 0:
                dense[c, minus[c]] = -1.0
    else:
        dense = rng.standard_normal((k, n))
    want = cpu.solve_many(dense, threads=4)
    for parts, reach in [("8", "1"), ("8", "0"), ("33", "1"), ("0", "1")]:
        monkeypatch.setenv("YBUS_HIP_PARTS", parts)
        monkeypatch.setenv("YBUS_HIP_REACH", reach)
        for tile in ["1", "8", "64"]:
            monkeypatch.setenv("YBUS_HIP_TILE", tile)
            factor = Y.GpuFactor.from_cholesky(gpu, cpu)
            factor.set_block_width(k)
            if kind == "pairs":
                with factor.pair_solves(pairs) as block:
                    got = block.solutions()
            else:
                got = factor.solve_many(dense)
            assert got.tobytes() == want.tobytes(), (parts, reach, tile)


@pytest.mark.parametrize("parts", ["8", "32"])
def test_the_top_kernel_takes_its_own_configuration(gpu, monkeypatch, parts):
    """The partitioned solve's top-of-tree kernel on a tile width and workgroup size of its own
    (YBUS_HIP_TOP=tile[:threads], or the autotuner's second stage for blocks of vectors) reads the
    parts' rows in the same block layout, and with the top held in LDS slots (the default where
    they fit; YBUS_HIP_TOP_LDS=0 keeps the block in global memory) every segment still subtracts
    its terms in order and divides as the sweep does: a block of vectors keeps the CPU's bits,
    over split sweeps and unsplit ones, at tile widths whose slots fit LDS and ones that do not."""
    g = O.make_grid(3000, 778)
    bp, _, _ = g.dc()
    cpu = O.Cholesky(bp)
    k = 200
    dense = np.random.default_rng(9).standard_normal((k, bp.rows))
    want = cpu.solve_many(dense, threads=4)
    monkeypatch.setenv("YBUS_HIP_PARTS", parts)
    monkeypatch.setenv("YBUS_HIP_PATH", "wide")
    for split in ["0", "4"]:
        monkeypatch.setenv("YBUS_HIP_SPLIT", split)
        for lds in ["0", "1"]:
            monkeypatch.setenv("YBUS_HIP_TOP_LDS", lds)
            for tile, top in [("8", "1:256"), ("8", "64:1024"), ("1", "16:64"), ("64", "2"), ("4", "4:128")]:
                monkeypatch.setenv("YBUS_HIP_TILE", tile)
                monkeypatch.setenv("YBUS_HIP_TOP", top)
                factor = Y.GpuFactor.from_cholesky(gpu, cpu)
                factor.set_block_width(k)
                case = (split, lds, tile, top)
                assert factor.solve_many(dense).tobytes() == want.tobytes(), case
                assert factor.solve_transpose_many(dense).tobytes() == want.tobytes(), case
        monkeypatch.delenv("YBUS_HIP_TILE")
        monkeypatch.delenv("YBUS_HIP_TOP")
        monkeypatch.delenv("YBUS_HIP_TOP_LDS")
        factor = Y.GpuFactor.from_cholesky(gpu, cpu)
        factor.set_block_width(k)
        tuned = factor.autotune(k)
        assert tuned["top_tile"] in (1, 2, 4, 8, 16, 32, 64) and tuned["top_threads"] >= tuned["top_tile"], tuned
        assert factor.solve_many(dense).tobytes() == want.tobytes(), (split, tuned)


@pytest.mark.parametrize("parts", ["8", "32"])
def test_the_parts_swept_in_lds_slots(gpu, monkeypatch, parts):
    """The partitioned solve's parts with their values in LDS slots (YBUS_HIP_PARTS_LDS=tile[:threads],
    or the autotuner's third stage for blocks of vectors): a wavefront per column of the tile, the
    records loaded rounds ahead, the backward parts reading the top's solutions as halo slots.
    Every segment still subtracts its terms in order and divides as the sweep does, so a block of
    vectors keeps the CPU's bits, over split sweeps and unsplit ones, at every tile width, with
    the last tile ragged (200 columns) or the block narrower than the tile (3 columns), with the
    top in LDS or in global memory."""
    g = O.make_grid(3000, 778)
    bp, _, _ = g.dc()
    cpu = O.Cholesky(bp)
    rng = np.random.default_rng(10)
    blocks = [rng.standard_normal((k, bp.rows)) for k in (200, 3)]
    wants = [cpu.solve_many(d, threads=4) for d in blocks]
    monkeypatch.setenv("YBUS_HIP_PARTS", parts)
    monkeypatch.setenv("YBUS_HIP_PATH", "wide")
    for split in ["0", "4"]:
        monkeypatch.setenv("YBUS_HIP_SPLIT", split)
        for top_lds in ["0", "1"]:
            monkeypatch.setenv("YBUS_HIP_TOP_LDS", top_lds)
            for lds in ["1", "2:128", "4", "8:1024", "16", "0"]:
                monkeypatch.setenv("YBUS_HIP_PARTS_LDS", lds)
                factor = Y.GpuFactor.from_cholesky(gpu, cpu)
                factor.set_block_width(200)
                for dense, want in zip(blocks, wants):
                    case = (split, top_lds, lds, len(dense))
                    assert factor.solve_many(dense).tobytes() == want.tobytes(), case
                    assert factor.solve_transpose_many(dense).tobytes() == want.tobytes(), case
        monkeypatch.delenv("YBUS_HIP_PARTS_LDS")
        monkeypatch.delenv("YBUS_HIP_TOP_LDS")
        factor = Y.GpuFactor.from_cholesky(gpu, cpu)
        factor.set_block_width(200)
        tuned = factor.autotune(200)
        assert tuned["parts_lds_tile"] in (0, 1, 2, 4, 8, 16), tuned
        assert tuned["parts_lds_tile"] == 0 or tuned["parts_lds_threads"] >= 64 * tuned["parts_lds_tile"], tuned
        assert factor.solve_many(blocks[0]).tobytes() == wants[0].tobytes(), (split, tuned)


def test_outage_updates_and_flow_maxima_over_partitioned_sweeps(gpu, monkeypatch):
    """The N-1 screening chain (rank-one pair updates, the operator's product maxima, the angle
    maxima) gives the same bits over a partitioned factor as over one schedule."""
    g = O.make_grid(3000, 515)
    bp, bf, _ = g.dc()
    cpu = O.Cholesky(bp)
    n = bp.rows
    theta = cpu.solve(O.injections(515, 4, 3000, 1)[0])
    plus = np.array([-1 if f == 0 else int(f) - 1 for f in g.branch_from])
    minus = np.array([-1 if t == 0 else int(t) - 1 for t in g.branch_to])
    outages = np.flatnonzero(~g.bridges())[:400]
    x = np.asarray(g.x, dtype=np.float64)
    weights = 1.0 / (0.5 + 1.5 * O.uniforms(515, 5, bf.rows))
    results = []
    for parts in ["0", "6"]:
        monkeypatch.setenv("YBUS_HIP_PARTS", parts)
        factor = Y.GpuFactor.from_cholesky(gpu, cpu)
        factor.set_block_width(len(outages))
        op = factor.operator(bf)
        op.set_offset(np.zeros(bf.rows))
        op.set_weights(weights)
        block = factor.pair_updates_arrays(theta, plus[outages], minus[outages], np.zeros(len(outages)),
                                           x[outages])
        loads = block.product_maxima_arrays(op, outages)
        angles = block.maxima_arrays()
        results.append((block.solutions().tobytes(), [a.tobytes() for a in loads], [a.tobytes() for a in angles]))
        block.free()
    assert results[0] == results[1]
    assert n == bp.rows


def host_outage(cpu, base, group):
    """dc/outage.rs:150-230 over the CPU's solves: Y = A^-1 C, x = base + sum alpha_k Y_k,
    W = D^-1 - C^T Y, z = W^-1 C^T x (a dense solve), and x + sum z_j Y_j."""
    n = len(base)
    ys = [cpu.solve(pair_vector(n, u.pair)) for u in group]
    x = base.copy()
    for u, y in zip(group, ys):
        x = x + u.alpha * y
    m = len(group)
    w = np.zeros((m, m))
    for i, u in enumerate(group):
        for j, y in enumerate(ys):
            w[i, j] = -dot(u.pair, y) + (u.inverse_weight if i == j else 0.0)
    ctx = np.array([dot(u.pair, x) for u in group])
    z = np.linalg.solve(w, ctx)
    for zj, y in zip(z, ys):
        x = x + zj * y
    return x


def test_outage_sets_match_the_host_and_single_branches_match_pair_updates(gpu):
    """Sets of one to five branches: the rank-m update agrees with the host's to rounding (its
    small dense solve may round differently), and a set of one is pair_updates bit for bit."""
    g = O.make_grid(400, 31)
    bp, _, _ = g.dc()
    cpu = O.Cholesky(bp)
    n = bp.rows
    base = cpu.solve(O.injections(31, 4, 400, 1)[0])
    plus = [-1 if f == 0 else int(f) - 1 for f in g.branch_from]
    minus = [-1 if t == 0 else int(t) - 1 for t in g.branch_to]
    x = np.asarray(g.x, dtype=np.float64)
    candidates = np.flatnonzero(~g.bridges())
    rng = np.random.default_rng(12)

    def update(k):
        return Y.PairUpdate(Y.Pair(None if plus[k] < 0 else plus[k], None if minus[k] < 0 else minus[k]), 0.0,
                            float(x[k]))

    sets = [[update(int(k)) for k in rng.choice(candidates, size=m, replace=False)] for m in [1, 2, 3, 5, 1, 4]]
    factor = Y.GpuFactor.from_cholesky(gpu, cpu)
    factor.set_block_width(64)
    block = factor.outage_updates(base, sets)
    got = block.solutions()
    assert not any(block.singular())
    block.free()
    for group, column in zip(sets, got):
        want = host_outage(cpu, base, group)
        assert max_difference(column, want) < 1e-12
    singles = [group[0] for group in sets if len(group) == 1]
    with factor.pair_updates(base, singles) as pu:
        ones = pu.solutions()
    assert got[0].tobytes() == ones[0].tobytes() and got[4].tobytes() == ones[1].tobytes()
    # a set of two branches against removing both from B' and solving
    dense = bp.to_dense()
    group = sets[1]
    for u in group:
        c = pair_vector(n, u.pair)
        dense = dense - np.outer(c, c) / u.inverse_weight
    p = bp.mul_vec(base)
    assert max_difference(got[1], np.linalg.solve(dense, p)) < 1e-9
