"""The device solve that allocates nothing (ybus_hip_factor_solve_device_noalloc; no counterpart in
the reference): called through the C ABI on device vectors, it must give the bits of
ybus_hip_factor_solve_device on two calls in a row (the second smaller), solve in chunks when the
fixed workspace cannot hold every vector at once, leave the factor's staged block as it was,
refuse what it cannot take, and leave the device's free memory unchanged."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as O
import ybus_amd as Y
from tests.fixtures import random_square_cuda
from tests.gpu_common import gpu  # noqa: F401

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def library_defaults(monkeypatch):
    for name in ("YBUS_HIP_PATH", "YBUS_HIP_PARTS", "YBUS_HIP_TILE", "YBUS_HIP_THREADS"):
        monkeypatch.delenv(name, raising=False)


def noalloc(factor, mode, b, x):
    """The entry point through the C ABI: vectors as rows of 2-D float64 GPU tensors."""
    torch.cuda.synchronize()
    return Y.library().ybus_hip_factor_solve_device_noalloc(factor._h, mode, b.data_ptr(), b.stride(0), b.shape[0],
                                                            x.data_ptr(), x.stride(0))


def device(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def reference(factor, b, mode):
    return factor.solve_device(b, transpose=mode == Y.TRANSPOSE).cpu().numpy()


@pytest.mark.parametrize("buses", [2000, 10000])
def test_two_calls_in_a_row_match_the_allocating_solve(gpu, buses):
    bp, _, _ = O.make_grid(buses, 20261230 + buses).dc()
    chol = O.Cholesky(bp)
    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    first = device(O.injections(3, 4, buses, 300))
    second = device(O.injections(5, 4, buses, 7))
    for b in (first, second):
        x = torch.full_like(b, float("nan"))
        assert noalloc(factor, Y.FP64, b, x) == Y.OK, Y._message()
        got = x.cpu().numpy()
        assert got.tobytes() == reference(factor, b, Y.FP64).tobytes()
        assert got.tobytes() == chol.solve_many(b.cpu().numpy()).tobytes()


def test_transposed_lu_solves_with_strides(gpu):
    a = random_square_cuda(900, 2, 11)
    lu = O.Lu(a)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    for count in (40, 3):
        wide = torch.zeros((count, 1000), dtype=torch.float64, device="cuda")
        rhs = np.random.default_rng(count).standard_normal((count, 900))
        wide[:, :900] = device(rhs)
        b = wide[:, :900]
        out = torch.full((count, 1024), float("nan"), dtype=torch.float64, device="cuda")
        x = out[:, :900]
        for mode, want in ((Y.TRANSPOSE, lu.solve_transpose_many(rhs)), (Y.FP64, lu.solve_many(rhs))):
            assert noalloc(factor, mode, b, x) == Y.OK, Y._message()
            got = x.cpu().numpy()
            assert got.tobytes() == want.tobytes(), (count, mode)
            assert got.tobytes() == reference(factor, b, mode).tobytes()
        assert torch.all(torch.isnan(out[:, 900:]))  # nothing written past each vector


def test_blocks_beyond_the_workspace_are_solved_in_chunks(gpu):
    # 64 vectors of a 70,000-bus B' take about 143 MB of wide-path workspace: more than the
    # 64 MiB array, so the call solves them in chunks, with the same bits
    bp, _, _ = O.make_grid(70000, 20261204).dc()
    chol = O.Cholesky(bp)
    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    factor.set_block_width(64)
    for count in (64, 20):
        b = device(O.injections(count, 4, 70000, count))
        x = torch.empty_like(b)
        assert noalloc(factor, Y.FP64, b, x) == Y.OK, Y._message()
        assert x.cpu().numpy().tobytes() == reference(factor, b, Y.FP64).tobytes()


def test_the_staged_block_is_left_as_it_was(gpu):
    bp, _, _ = O.make_grid(3000, 8).dc()
    chol = O.Cholesky(bp)
    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    staged = O.injections(1, 4, 3000, 50)
    factor.stage(staged)
    b = device(O.injections(2, 4, 3000, 30))
    x = torch.empty_like(b)
    assert noalloc(factor, Y.FP64, b, x) == Y.OK, Y._message()
    with factor.solve_staged(Y.FP64) as block:
        assert block.solutions().tobytes() == chol.solve_many(staged).tobytes()


def test_what_it_cannot_take_is_refused(gpu):
    # a 300,000-unknown chain: one vector's wide-path workspace (b, t and x of 16 interleaved
    # columns, and the vector itself) is about 118 MB, beyond the array
    n = 300000
    t = [(i, i, 4.0) for i in range(n)] + [(i, i + 1, -1.0) for i in range(n - 1)] + \
        [(i + 1, i, -1.0) for i in range(n - 1)]
    chol = O.Cholesky(O.Matrix.from_triplets(n, n, t))
    big = Y.GpuFactor.from_cholesky(gpu, chol)
    b = torch.ones((1, n), dtype=torch.float64, device="cuda")
    x = torch.empty_like(b)
    assert noalloc(big, Y.FP64, b, x) == Y.TOO_LARGE
    assert "no-allocation workspace" in Y._message()
    small = Y.GpuFactor.from_cholesky(gpu, O.Cholesky(O.make_grid(500, 2).dc()[0]))
    b = torch.ones((2, 499), dtype=torch.float64, device="cuda")
    x = torch.empty_like(b)
    assert noalloc(small, Y.FP32_REFINED, b, x) == Y.INVALID_ARGUMENT
    assert Y.library().ybus_hip_factor_solve_device_noalloc(small._h, Y.FP64, b.data_ptr(), 100, 2, x.data_ptr(),
                                                            499) == Y.INVALID_ARGUMENT  # stride below dim()
    assert Y.library().ybus_hip_factor_solve_device_noalloc(None, Y.FP64, b.data_ptr(), 499, 2, x.data_ptr(),
                                                            499) == Y.INVALID_ARGUMENT
    with small.pair_solves([Y.Pair(1, 2)]):
        assert noalloc(small, Y.FP64, b, x) == Y.INVALID_ARGUMENT  # a block borrows the factor


def test_the_device_memory_in_use_does_not_change(gpu):
    bp, _, _ = O.make_grid(8000, 31).dc()
    chol = O.Cholesky(bp)
    warm = Y.GpuFactor.from_cholesky(gpu, chol)
    fresh = Y.GpuFactor.from_cholesky(gpu, chol)  # never solved: no workspace of its own yet
    b = device(O.injections(4, 4, 8000, 200))
    x = torch.empty_like(b)
    assert noalloc(warm, Y.FP64, b, x) == Y.OK  # the runtime's own first-launch set-up
    torch.cuda.synchronize()
    free_before, _ = gpu.memory()
    for count in (200, 50):
        assert Y.library().ybus_hip_factor_solve_device_noalloc(fresh._h, Y.FP64, b.data_ptr(), b.stride(0), count,
                                                                x.data_ptr(), x.stride(0)) == Y.OK
    torch.cuda.synchronize()
    free_after, _ = gpu.memory()
    assert free_after == free_before
    got = x[:50].cpu().numpy()
    assert got.tobytes() == chol.solve_many(b[:50].cpu().numpy()).tobytes()
