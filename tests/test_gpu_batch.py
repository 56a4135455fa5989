"""Batches of factors solved together (new: the reference solves one factor per call). Each
factor's block must give the CPU's bits, whichever LDS class it falls in; the captured graph must
give the same bits on every replay, with other device work between replays; and what a batch
cannot take is refused."""
import numpy as np
import pytest

import oracle as O
import ybus_amd as Y
from tests.fixtures import random_spd, random_square_cuda, right_hand_sides
from tests.gpu_common import gpu  # noqa: F401

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def default_path(monkeypatch):
    monkeypatch.delenv("YBUS_HIP_PATH", raising=False)


def mixed_factors(gpu):
    """(cpu factor, gpu factor) pairs: small and large Cholesky and LU factors, so the batch has
    items in several LDS classes (a 12,000-bus grid's solution alone fills a compute unit's LDS)."""
    out = []
    chol = O.Cholesky(random_spd(300, 3, 2))
    out.append((chol, Y.GpuFactor.from_cholesky(gpu, chol)))
    lu = O.Lu(random_square_cuda(200, 4, 5))
    out.append((lu, Y.GpuFactor.from_lu(gpu, lu)))
    one = O.Lu(O.Matrix.from_triplets(1, 1, [(0, 0, -2.5)], row_major=False))
    out.append((one, Y.GpuFactor.from_lu(gpu, one)))
    for buses in [2000, 12000]:
        bp, _, _ = O.make_grid(buses, 555 + buses).dc()
        c = O.Cholesky(bp)
        out.append((c, Y.GpuFactor.from_cholesky(gpu, c)))
    jac, _ = O.make_grid(1500, 77).jacobian(77)
    j = O.Lu(jac, O.COLAMD)
    out.append((j, Y.GpuFactor.from_lu(gpu, j)))
    return out


def test_each_factor_gets_the_cpu_bits(gpu):
    pairs = mixed_factors(gpu)
    batch = Y.GpuBatch.new(gpu)
    for _, f in pairs:
        batch.add(f)
    assert batch.factors() == len(pairs)
    counts = [5, 17, 3, 0, 9, 12]
    rhs = [right_hand_sides(f.dim(), c, 10 + k) for k, ((_, f), c) in enumerate(zip(pairs, counts))]
    got = batch.solve(rhs)
    for (cpu, f), b, x, c in zip(pairs, rhs, got, counts):
        assert x.shape == (c, f.dim())
        if c:
            assert x.tobytes() == cpu.solve_many(b).tobytes()


def test_graph_replays_give_the_same_bits(gpu):
    pairs = mixed_factors(gpu)[:4]
    batch = Y.GpuBatch.new(gpu)
    for _, f in pairs:
        batch.add(f)
    rhs = [right_hand_sides(f.dim(), 6, 30 + k) for k, (_, f) in enumerate(pairs)]
    batch.stage(rhs)
    batch.solve_staged()
    first = [x.copy() for x in batch.solutions()]
    want = [cpu.solve_many(b) for (cpu, _), b in zip(pairs, rhs)]
    for x, w in zip(first, want):
        assert x.tobytes() == w.tobytes()
    other_cpu, other = pairs[1]
    other_rhs = right_hand_sides(other.dim(), 40, 99)
    for replay in range(4):
        # the program's own device work between replays
        assert other.solve_many(other_rhs).tobytes() == other_cpu.solve_many(other_rhs).tobytes()
        batch.solve_staged()
        for x, w in zip(batch.solutions(), want):
            assert x.tobytes() == w.tobytes(), replay
    assert batch.kernel_stats().launches == 5
    # a new staging (other counts) is captured anew
    rhs2 = [right_hand_sides(f.dim(), 2 + 3 * k, 60 + k) for k, (_, f) in enumerate(pairs)]
    for x, (cpu, _), b in zip(batch.solve(rhs2), pairs, rhs2):
        assert x.tobytes() == cpu.solve_many(b).tobytes()


def test_what_a_batch_cannot_take_is_refused(gpu):
    bp, _, _ = O.make_grid(26000, 3).dc()  # a 25,999-unknown solution does not fit in LDS
    big = Y.GpuFactor.from_cholesky(gpu, O.Cholesky(bp))
    batch = Y.GpuBatch.new(gpu)
    with pytest.raises(Y.InvalidArgument):
        batch.add(big)
    small = Y.GpuFactor.from_cholesky(gpu, O.Cholesky(random_spd(50, 10, 1)))
    batch.add(small)
    with pytest.raises(Y.InvalidArgument):
        batch.stage([])
    with pytest.raises(Y.InvalidArgument):
        batch.solve_staged()  # nothing staged
    assert [x.shape for x in batch.solve([np.zeros((0, 50))])] == [(0, 50)]


def test_fp32_refined_batches_reach_the_fp64_floor_on_every_replay(gpu):
    """The batch's FP32-refined mode refines each column in FP64 against its own factor's matrix:
    the reported ratio is the CPU's relative residual of the returned solution, it reaches 1e-12
    or the FP64 floor the CPU's own solve meets, and every replay of the graph gives the same bits,
    also after FP64 solves of the same batch in between."""
    mats = [random_spd(300, 3, 2)]
    for buses in [2000, 12000]:
        bp, _, _ = O.make_grid(buses, 555 + buses).dc()
        mats.append(bp)
    jac, _ = O.make_grid(1500, 77).jacobian(77)
    batch = Y.GpuBatch.new(gpu)
    cpu = []
    for m in mats:
        c = O.Cholesky(m)
        f = Y.GpuFactor.from_cholesky(gpu, c)
        f.set_matrix(m)
        batch.add(f)
        cpu.append((m, c))
    j = O.Lu(jac, O.COLAMD)
    fj = Y.GpuFactor.from_lu(gpu, j)
    fj.set_matrix(jac)
    batch.add(fj)
    cpu.append((jac, j))
    counts = [4, 7, 3, 9]
    rhs = [right_hand_sides(m.rows, c, 50 + k) for k, ((m, _), c) in enumerate(zip(cpu, counts))]
    batch.stage(rhs)
    batch.solve_staged(Y.FP32_REFINED, 1e-12, 10)
    first = [x.copy() for x in batch.solutions()]
    for (m, c), b, x, (steps, ratio) in zip(cpu, rhs, first, batch.refinement()):
        exact = c.solve_many(b)
        for k in range(len(b)):
            assert ratio[k] == m.relative_residual(x[k], b[k])
            assert ratio[k] <= max(1e-12, 2.0 * m.relative_residual(exact[k], b[k])), (k, ratio[k])
            assert 1 <= steps[k] <= 10
    for _ in range(3):
        batch.solve_staged(Y.FP64)
        for (m, c), b, x in zip(cpu, rhs, batch.solutions()):
            assert x.tobytes() == c.solve_many(b).tobytes()
        batch.solve_staged(Y.FP32_REFINED, 1e-12, 10)
        for x, want in zip(batch.solutions(), first):
            assert x.tobytes() == want.tobytes()
