"""GPU solves against the oracle's CPU solves (ybus_cuda/tests/gpu_factor_test.rs restated):
Cholesky and LU (with pivoting), 2x2 blocks, a single row, block widths narrower and wider than
the number of right-hand sides, repeated runs, empty and mismatched right-hand sides. The FP64
path keeps the CPU's operation order, so beyond the reference's 1e-12 bar its solutions are
checked bit for bit; transposed and FP32-refined solves are checked too."""
import numpy as np
import pytest

import oracle as O
import ybus_amd as Y
from tests.fixtures import (random_spd, random_square_cuda, right_hand_sides, two_by_two_blocks,
                            worst_relative_difference)
from tests.gpu_common import gpu, path  # noqa: F401

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("path")]

BAR = 1e-12


def test_cholesky_solves_match_the_cpu_at_every_block_width(gpu):
    a = random_spd(300, 3, 7)
    cholesky = O.Cholesky(a)
    rhs = right_hand_sides(300, 37, 11)
    cpu = cholesky.solve_many(rhs)
    factor = Y.GpuFactor.from_cholesky(gpu, cholesky)
    for width in [1, 5, 37, 64, 1024, 4096]:
        factor.set_block_width(width)
        solved = factor.solve_many(rhs)
        assert worst_relative_difference(solved, cpu) < BAR, width
        assert solved.tobytes() == cpu.tobytes(), width


def test_lu_solves_match_the_cpu_with_pivoting(gpu):
    a = random_square_cuda(250, 4, 3)
    lu = O.Lu(a)
    assert any(i != p for i, p in enumerate(lu.row_perm)), "the fixture should pivot off the diagonal"
    rhs = right_hand_sides(250, 20, 5)
    cpu = lu.solve_many(rhs)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    factor.set_block_width(8)
    solved = factor.solve_many(rhs)
    assert worst_relative_difference(solved, cpu) < BAR
    assert solved.tobytes() == cpu.tobytes()


def test_two_by_two_blocks_and_a_single_row(gpu):
    lu = O.Lu(two_by_two_blocks(40))
    rhs = right_hand_sides(80, 3, 1)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    solved = factor.solve_many(rhs)
    assert worst_relative_difference(solved, lu.solve_many(rhs)) < BAR
    one = O.Lu(O.Matrix.from_triplets(1, 1, [(0, 0, 3.0)], row_major=False))
    rhs = np.array([[j - 70.5] for j in range(200)])
    factor = Y.GpuFactor.from_lu(gpu, one)
    for width in [1, 64, 65, 1024]:
        factor.set_block_width(width)
        assert factor.solve_many(rhs).tobytes() == one.solve_many(rhs).tobytes()


def test_repeated_solves_give_the_same_bits(gpu):
    cholesky = O.Cholesky(random_spd(500, 2, 13))
    rhs = right_hand_sides(500, 50, 17)
    factor = Y.GpuFactor.from_cholesky(gpu, cholesky)
    factor.set_block_width(16)
    first = factor.solve_many(rhs)
    factor.set_block_width(7)
    factor.set_block_width(16)
    for _ in range(3):
        assert factor.solve_many(rhs).tobytes() == first.tobytes()


def test_empty_and_mismatched_right_hand_sides(gpu):
    cholesky = O.Cholesky(random_spd(10, 20, 1))
    factor = Y.GpuFactor.from_cholesky(gpu, cholesky)
    assert factor.solve_many([]).shape == (0, 10)
    with pytest.raises(Y.ShapeMismatch) as e:
        factor.solve_many([np.zeros(10), np.zeros(9)])
    assert (e.value.expected, e.value.found) == (10, 9)
    assert factor.timings().blocks == 0
    factor.solve_many([np.ones(10)])
    assert factor.timings().blocks == 1
    factor.reset_timings()
    assert factor.timings().blocks == 0


def test_transposed_solves_match_the_cpu_bit_for_bit(gpu):
    a = random_square_cuda(200, 4, 8)
    lu = O.Lu(a)
    rhs = right_hand_sides(200, 33, 2)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    got = factor.solve_transpose_many(rhs)
    assert got.tobytes() == lu.solve_transpose_many(rhs).tobytes()
    at = a.to_dense().T
    assert np.max(np.abs(at @ got.T - rhs.T)) <= 1e-10
    # Cholesky: A^T = A, the same solve
    chol = O.Cholesky(random_spd(120, 5, 4))
    f2 = Y.GpuFactor.from_cholesky(gpu, chol)
    r2 = right_hand_sides(120, 9, 3)
    assert f2.solve_transpose_many(r2).tobytes() == chol.solve_many(r2).tobytes()


@pytest.mark.parametrize("kind", ["cholesky", "lu", "lu_colamd"])
def test_fp32_refined_solves_reach_the_tolerance(gpu, kind):
    if kind == "cholesky":
        a = random_spd(400, 2, 31)
        f = O.Cholesky(a)
        factor = Y.GpuFactor.from_cholesky(gpu, f)
    else:
        g = O.make_grid(300, 41)
        a, _ = g.jacobian(41)
        f = O.Lu(a, O.COLAMD if kind == "lu_colamd" else O.AMD)
        factor = Y.GpuFactor.from_lu(gpu, f)
    factor.set_matrix(a)
    rhs = right_hand_sides(a.rows, 70, 12)
    x, steps, ratio = factor.solve_many_refined(rhs, 1e-12, 10)
    cpu = f.solve_many(rhs)
    for k in range(rhs.shape[0]):
        # the reported ratio is the relative residual of the returned x, by the CPU's own
        # residual (the device sums A x in the same order, so the bits agree)
        assert ratio[k] == a.relative_residual(x[k], rhs[k])
        # 1e-12, unless the column's FP64 floor is above it: ||b - fl(A x)|| / ||b|| cannot
        # drop below the rounding of A x itself, which the CPU's FP64 solve also meets
        floor = a.relative_residual(cpu[k], rhs[k])
        assert ratio[k] <= max(1e-12, 2.0 * floor), (k, ratio[k], floor, steps[k])
        assert 1 <= steps[k] <= 10
    assert worst_relative_difference(x, cpu) < 1e-9


def test_staged_blocks_solve_on_the_device(gpu):
    a = random_spd(256, 3, 5)
    chol = O.Cholesky(a)
    rhs = right_hand_sides(256, 40, 6)
    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    factor.set_matrix(a)
    factor.stage(rhs)
    with factor.solve_staged(Y.FP64) as block:
        assert block.columns() == 40
        assert block.solutions().tobytes() == chol.solve_many(rhs).tobytes()
    with factor.solve_staged(Y.FP32_REFINED, 1e-12, 8) as block:
        steps, ratio = block.refinement()
        assert np.all(ratio <= 1e-12)
        x = block.solutions()
        assert ratio[3] == a.relative_residual(x[3], rhs[3])
    # a block borrows its factor until freed
    block = factor.solve_staged(Y.FP64)
    with pytest.raises(Y.InvalidArgument):
        Y._check(Y.library().ybus_hip_factor_stage(factor._h, rhs, 1))
    block.free()


@pytest.mark.parametrize("buses", [2000, 10000])
def test_synthetic_dc_grids_match_the_cpu_bit_for_bit(gpu, buses):
    g = O.make_grid(buses, 20261203 + buses)
    bp, bf, rows = g.dc()
    chol = O.Cholesky(bp)
    rhs = O.injections(7, 4, buses, 96)
    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    got = factor.solve_many(rhs)
    cpu = chol.solve_many(rhs, threads=4)
    assert got.tobytes() == cpu.tobytes()
    for k in range(0, 96, 17):
        # backward stable: the residual is within a small multiple of the rounding of A x, the
        # floor ||b - fl(A x)|| / ||b|| of any FP64 solution (above 1e-12 for these
        # injections, where |A||x| is far larger than |b|)
        assert bp.relative_residual(got[k], rhs[k]) <= max(1e-12, fp64_floor(bp, got[k], rhs[k]))


def fp64_floor(a, x, b):
    """32 unit roundoffs of || |A| |x| ||inf / ||b||inf."""
    abs_ax = O.Matrix(a.rows, a.cols, a.row_major, a.ptr, a.idx, np.abs(a.data)).mul_vec(np.abs(x))
    return 32 * 2.0**-53 * float(np.max(abs_ax)) / float(np.max(np.abs(b)))


def test_synthetic_jacobian_lu_matches_the_cpu(gpu):
    g = O.make_grid(2500, 77)
    jac, _ = g.jacobian(77)
    lu = O.Lu(jac, O.COLAMD)
    rhs = O.normals(3, 9, 64 * jac.rows).reshape(64, jac.rows)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    assert factor.solve_many(rhs).tobytes() == lu.solve_many(rhs, threads=4).tobytes()
    assert factor.solve_transpose_many(rhs).tobytes() == lu.solve_transpose_many(rhs, threads=4).tobytes()
    factor.set_matrix(jac)
    x, steps, ratio = factor.solve_many_refined(rhs, 1e-12, 10)
    cpu = lu.solve_many(rhs, threads=4)
    floors = np.array([jac.relative_residual(cpu[k], rhs[k]) for k in range(64)])
    assert np.all(ratio <= np.maximum(1e-12, 2.0 * floors)), (ratio.max(), floors.max())


def test_transfers_larger_than_the_staging_chunks(gpu):
    # 1,000 right-hand sides of a 10,000-bus B' are 80 MB each way: the host vectors cross in
    # three 32 MB chunks through the alternating pinned buffers, both ways, bit for bit
    grid = O.make_grid(10000, 20261210)
    bprime, _, _ = grid.dc()
    chol = O.Cholesky(bprime)
    rhs = O.injections(7, 3, grid.buses, 1000)
    assert rhs.nbytes > 2 * 32 * 2**20
    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    cpu = chol.solve_many(rhs, 8)
    assert factor.solve_many(rhs).tobytes() == cpu.tobytes()
    # and a pair block's solutions out of the device
    pairs = [Y.Pair(int(k), int(k) + 1) for k in range(1000)]
    dense = np.zeros((1000, chol.n))
    for k in range(1000):
        dense[k, k], dense[k, k + 1] = 1.0, -1.0
    with factor.pair_solves(pairs) as block:
        assert block.solutions().tobytes() == chol.solve_many(dense, 8).tobytes()


def test_handles_left_open_at_exit_end_the_process_quietly():
    """A process that ends with a context, factor, block, operator and batch still open exits
    with status 0 and prints nothing: the handles are left to the process's exit rather than freed
    from a half-torn-down interpreter."""
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = (
        "import torch, oracle as O, ybus_amd as Y\n"
        "from tests.fixtures import random_spd\n"
        "gpu = Y.Gpu.new()\n"
        "a = random_spd(80, 3, 5)\n"
        "factor = Y.GpuFactor.from_cholesky(gpu, O.Cholesky(a))\n"
        "block = factor.pair_solves([Y.Pair(1, 2), Y.Pair(3, None)])\n"
        "operator = factor.operator(a)\n"
        "batch = Y.GpuBatch.new(gpu)\n"
        "batch.add(factor)\n"
        "print('open', flush=True)\n"
    )
    # the library's own defaults (a path forced by this module's fixture would make the batch refuse)
    env = {k: v for k, v in os.environ.items() if not k.startswith("YBUS_HIP_")}
    out = subprocess.run([sys.executable, "-c", script], cwd=root, env=env, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "open"
    assert "Exception ignored" not in out.stderr and "Traceback" not in out.stderr, out.stderr
