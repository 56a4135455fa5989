"""The kernel path is chosen per block: the LDS column path for factors whose solution fits in
LDS and blocks of at most two columns per compute unit, the wide tile path otherwise. Both give
the CPU's bits, so a block solved either way is the same block."""
import numpy as np
import pytest

import oracle as O
import ybus_amd as Y
from tests.fixtures import random_spd, right_hand_sides
from tests.gpu_common import gpu  # noqa: F401

pytestmark = pytest.mark.gpu


def test_the_path_follows_the_block_width(gpu, monkeypatch):
    monkeypatch.delenv("YBUS_HIP_PATH", raising=False)
    chol = O.Cholesky(random_spd(200, 3, 2))
    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    assert factor.path(1) == "columns"
    assert factor.path(10**6) == "wide"
    monkeypatch.setenv("YBUS_HIP_PATH", "wide")
    assert Y.GpuFactor.from_cholesky(gpu, chol).path(1) == "wide"
    monkeypatch.setenv("YBUS_HIP_PATH", "columns")
    assert Y.GpuFactor.from_cholesky(gpu, chol).path(10**6) == "columns"


def test_a_block_wide_enough_for_both_paths_gives_the_same_bits(gpu, monkeypatch):
    monkeypatch.delenv("YBUS_HIP_PATH", raising=False)
    g = O.make_grid(1500, 5)
    bp, _, _ = g.dc()
    chol = O.Cholesky(bp)
    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    narrow = right_hand_sides(bp.rows, 40, 8)
    wide = np.concatenate([narrow] * 40)  # 1600 columns: past the column path's default width
    assert factor.path(len(narrow)) == "columns" and factor.path(len(wide)) == "wide"
    cpu = chol.solve_many(narrow)
    assert factor.solve_many(narrow).tobytes() == cpu.tobytes()
    assert factor.solve_many(wide).tobytes() == np.concatenate([cpu] * 40).tobytes()


def test_the_autotuner_keeps_a_configuration_that_gives_the_cpu_bits(gpu, monkeypatch):
    monkeypatch.delenv("YBUS_HIP_PATH", raising=False)
    g = O.make_grid(2000, 11)
    bp, _, _ = g.dc()
    chol = O.Cholesky(bp)
    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    rng = np.random.default_rng(1)
    for columns in [40, 700]:
        for pairs in [True, False]:
            choice = factor.autotune(columns, pairs)
            assert choice["path"] in ("columns", "wide") and choice["milliseconds"] > 0.0
            assert factor.path(columns, pairs) == choice["path"]
        plus = rng.integers(0, bp.rows, size=columns)
        minus = (plus + 1 + rng.integers(0, bp.rows - 1, size=columns)) % bp.rows
        dense = np.zeros((columns, bp.rows))
        dense[np.arange(columns), plus] = 1.0
        dense[np.arange(columns), minus] = -1.0
        pairs = [Y.Pair(int(p), int(m)) for p, m in zip(plus, minus)]
        with factor.pair_solves(pairs) as block:
            assert block.solutions().tobytes() == chol.solve_many(dense).tobytes(), columns
        assert factor.solve_many(dense).tobytes() == chol.solve_many(dense).tobytes()
    with pytest.raises(Y.InvalidArgument):
        factor.autotune(0)
    # an environment override still decides the path
    monkeypatch.setenv("YBUS_HIP_PATH", "wide")
    forced = Y.GpuFactor.from_cholesky(gpu, chol)
    assert forced.autotune(40, pairs=True)["path"] == "wide"
