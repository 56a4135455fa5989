"""The one-schedule wide FP64 solve with each tile's last positions' values in an LDS ring
(YBUS_HIP_WINDOW=1, or the autotuner's stage for it): a term still in the ring is read from LDS, a
term final two levels back is loaded a level ahead, and only the rest go to the block when their
row is swept. Every row still subtracts its terms in order and divides as the level-scheduled
sweep does, so the bits are the CPU's: LU and Cholesky factors, split and unsplit sweeps, blocks
of vectors (ragged tiles, transposed solves) and pair blocks without a reach, at tile widths
whose rings hold from a few hundred to tens of thousands of positions. The allocation-free device
solve takes it only once its encoding is built (it builds nothing), with the same bits either
way."""
import numpy as np
import pytest
import torch

import oracle as O
import ybus_amd as Y
from tests.fixtures import right_hand_sides
from tests.gpu_common import gpu  # noqa: F401

pytestmark = pytest.mark.gpu


@pytest.fixture
def wide(monkeypatch):
    monkeypatch.setenv("YBUS_HIP_PATH", "wide")
    monkeypatch.setenv("YBUS_HIP_PARTS", "0")
    monkeypatch.setenv("YBUS_HIP_WINDOW", "1")
    return monkeypatch


def lu_case():
    jac, _ = O.make_grid(2500, 91).jacobian(91)
    return O.Lu(jac, O.COLAMD), jac.rows


def cholesky_case():
    bp, _, _ = O.make_grid(3000, 92).dc()
    return O.Cholesky(bp), bp.rows


@pytest.mark.parametrize("kind", ["lu", "cholesky"])
def test_blocks_of_vectors_keep_the_cpu_bits(gpu, wide, kind):
    cpu, n = lu_case() if kind == "lu" else cholesky_case()
    b = right_hand_sides(n, 37, 5)
    want = cpu.solve_many(b, threads=4)
    want_t = cpu.solve_transpose_many(b, threads=4) if kind == "lu" else want
    for split in (["0", "4"] if kind == "lu" else ["4"]):
        wide.setenv("YBUS_HIP_SPLIT", split)
        for tile, threads in [("1", "64"), ("2", "512"), ("4", "256"), ("16", "1024"), ("64", "64")]:
            wide.setenv("YBUS_HIP_TILE", tile)
            wide.setenv("YBUS_HIP_THREADS", threads)
            factor = Y.GpuFactor.from_lu(gpu, cpu) if kind == "lu" else Y.GpuFactor.from_cholesky(gpu, cpu)
            factor.set_block_width(37)
            case = (split, tile, threads)
            assert factor.solve_many(b).tobytes() == want.tobytes(), case
            assert factor.solve_transpose_many(b).tobytes() == want_t.tobytes(), case


def test_pair_blocks_without_the_reach_keep_the_cpu_bits(gpu, wide):
    wide.setenv("YBUS_HIP_REACH", "0")
    cpu, n = cholesky_case()
    rng = np.random.default_rng(8)
    k = 300
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
    factor = Y.GpuFactor.from_cholesky(gpu, cpu)
    factor.set_block_width(k)
    with factor.pair_solves(pairs) as block:
        assert block.solutions().tobytes() == want.tobytes()


def test_the_allocation_free_solve_builds_no_encoding(gpu, wide):
    cpu, n = lu_case()
    wide.setenv("YBUS_HIP_TILE", "2")
    factor = Y.GpuFactor.from_lu(gpu, cpu)
    b = torch.tensor(right_hand_sides(n, 20, 9), dtype=torch.float64, device="cuda")
    want = cpu.solve_many(b.cpu().numpy(), threads=4)
    for step in range(2):
        # first with no encoding built (the barrier solve), then after a solve built it
        x = torch.full_like(b, float("nan"))
        torch.cuda.synchronize()
        assert Y.library().ybus_hip_factor_solve_device_noalloc(factor._h, Y.FP64, b.data_ptr(), b.stride(0),
                                                                b.shape[0], x.data_ptr(), x.stride(0)) == Y.OK
        assert x.cpu().numpy().tobytes() == want.tobytes(), step
        factor.solve_device(b)


def test_the_autotuner_times_it(gpu, monkeypatch):
    monkeypatch.setenv("YBUS_HIP_PATH", "wide")
    cpu, n = lu_case()
    b = right_hand_sides(n, 64, 7)
    factor = Y.GpuFactor.from_lu(gpu, cpu)
    factor.set_block_width(64)
    tuned = factor.autotune(64)
    assert isinstance(tuned["window"], bool), tuned
    assert factor.solve_many(b).tobytes() == cpu.solve_many(b, threads=4).tobytes(), tuned
