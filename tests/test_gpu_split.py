"""An LU factor's long rows split into chained segments for the wide solves (split_rows in the host
analysis): each segment starts from the previous one's partial value and subtracts the row's
next terms in the CPU's order, so the split sweeps must give the unsplit sweeps' bits (and the
CPU's) for FP64, transposed and FP32-refined solves, for pair blocks, and after a device refactor
rewrites their values. YBUS_HIP_SPLIT=0 at upload keeps a factor unsplit."""
import numpy as np
import pytest

import oracle as O
import ybus_amd as Y
from tests.fixtures import random_square_cuda, right_hand_sides
from tests.gpu_common import gpu  # noqa: F401

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def wide_path(monkeypatch):
    monkeypatch.setenv("YBUS_HIP_PATH", "wide")
    monkeypatch.delenv("YBUS_HIP_PARTS", raising=False)


def factors(gpu, lu, monkeypatch, cap="4"):
    """The same LU factor uploaded split (segments of at most `cap` terms) and unsplit."""
    monkeypatch.setenv("YBUS_HIP_SPLIT", cap)
    split = Y.GpuFactor.from_lu(gpu, lu)
    monkeypatch.setenv("YBUS_HIP_SPLIT", "0")
    whole = Y.GpuFactor.from_lu(gpu, lu)
    return split, whole


def cases():
    grid = O.make_grid(2500, 77)
    jac, _ = grid.jacobian(77)
    return [("jacobian", jac, O.COLAMD), ("random", random_square_cuda(400, 6, 3), O.AMD)]


@pytest.mark.parametrize("case", cases(), ids=lambda c: c[0])
@pytest.mark.parametrize("cap", ["4", "12"])
def test_split_sweeps_give_the_unsplit_bits(gpu, monkeypatch, case, cap):
    _, a, ordering = case
    lu = O.Lu(a, ordering)
    split, whole = factors(gpu, lu, monkeypatch, cap)
    rhs = right_hand_sides(lu.n, 70, 4)
    cpu = lu.solve_many(rhs)
    for width in (70, 16, 3):
        split.set_block_width(width)
        whole.set_block_width(width)
        assert split.solve_many(rhs).tobytes() == cpu.tobytes(), width
        assert whole.solve_many(rhs).tobytes() == cpu.tobytes(), width
        assert split.solve_transpose_many(rhs).tobytes() == lu.solve_transpose_many(rhs).tobytes(), width
    split.set_block_width(70)
    whole.set_block_width(70)
    # FP32 sweeps: the same float operations in the same order, so the refined iterates agree too
    for f in (split, whole):
        f.set_matrix(a)
    xs, ss, rs = split.solve_many_refined(rhs, 1e-12, 10)
    xw, sw, rw = whole.solve_many_refined(rhs, 1e-12, 10)
    assert xs.tobytes() == xw.tobytes() and np.array_equal(ss, sw) and rs.tobytes() == rw.tobytes()
    # pair blocks, with the reach (unsplit forward sweep) and without it (split solve)
    pairs = [Y.Pair(int(k), int((7 * k + 3) % lu.n) if (7 * k + 3) % lu.n != k else None) for k in range(50)]
    dense = np.zeros((50, lu.n))
    for k, p in enumerate(pairs):
        dense[k, p.plus] = 1.0
        if p.minus is not None:
            dense[k, p.minus] = -1.0
    want = lu.solve_many(dense)
    with split.pair_solves(pairs) as block:
        assert block.solutions().tobytes() == want.tobytes()
    monkeypatch.setenv("YBUS_HIP_REACH", "0")
    monkeypatch.setenv("YBUS_HIP_SPLIT", cap)
    no_reach = Y.GpuFactor.from_lu(gpu, lu)
    with no_reach.pair_solves(pairs) as block:
        assert block.solutions().tobytes() == want.tobytes()


def test_a_refactor_rewrites_the_split_sweeps(gpu, monkeypatch):
    grid = O.make_grid(1200, 20261209)
    a = grid.jacobian(20261209)[0]
    b = grid.jacobian(20261209, 0.002, 4)[0]
    lu = O.Lu(a, O.COLAMD)
    split, _ = factors(gpu, lu, monkeypatch)
    rhs = right_hand_sides(lu.n, 24, 6)
    assert split.refactor(b) == "reused"
    assert lu.refactor(b) == "reused"
    assert split.solve_many(rhs).tobytes() == lu.solve_many(rhs).tobytes()
    assert split.solve_transpose_many(rhs).tobytes() == lu.solve_transpose_many(rhs).tobytes()
    # and values from a host refactor
    c = grid.jacobian(20261209, 0.002, 9)[0]
    assert lu.refactor(c) == "reused"
    split.update_lu_values(lu.l.data, lu.u.data)
    assert split.solve_many(rhs).tobytes() == lu.solve_many(rhs).tobytes()
    split.set_matrix(c)
    x, steps, ratio = split.solve_many_refined(rhs, 1e-12, 10)
    for k in range(len(rhs)):
        assert ratio[k] == c.relative_residual(x[k], rhs[k])
