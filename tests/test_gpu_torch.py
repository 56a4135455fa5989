"""Vectors already in device memory (PyTorch tensors on the GPU): the device-pointer solve and
adjoint give the host-vector calls' bits, for contiguous and strided vectors, and the autograd
wrapper's gradients are the oracle's adjoint bit for bit."""
import numpy as np
import pytest
import torch

import oracle as O
import ybus_amd as Y
from tests.fixtures import random_spd, random_square_cuda, right_hand_sides
from tests.gpu_common import gpu, path  # noqa: F401

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("path")]


def cuda(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def test_device_solves_match_the_host_solves(gpu):
    a = random_square_cuda(150, 4, 12)
    lu = O.Lu(a)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    rhs = right_hand_sides(150, 23, 4)
    assert factor.solve_device(cuda(rhs)).cpu().numpy().tobytes() == lu.solve_many(rhs).tobytes()
    got = factor.solve_device(cuda(rhs), transpose=True).cpu().numpy()
    assert got.tobytes() == lu.solve_transpose_many(rhs).tobytes()
    # vectors with a stride wider than dim(), in and out
    wide_in = torch.zeros((23, 157), dtype=torch.float64, device="cuda")
    wide_in[:, :150] = cuda(rhs)
    wide_out = torch.full((23, 161), 7.0, dtype=torch.float64, device="cuda")
    factor.solve_device(wide_in[:, :150], wide_out[:, :150])
    assert wide_out[:, :150].cpu().numpy().tobytes() == lu.solve_many(rhs).tobytes()
    assert torch.all(wide_out[:, 150:] == 7.0)
    factor.set_block_width(5)  # several blocks
    assert factor.solve_device(cuda(rhs)).cpu().numpy().tobytes() == lu.solve_many(rhs).tobytes()


def test_device_adjoint_matches_the_host_adjoint(gpu):
    a = random_spd(140, 4, 6)
    chol = O.Cholesky(a)
    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    factor.set_matrix(a)
    x_bar = right_hand_sides(140, 19, 1)
    x = right_hand_sides(140, 19, 2)
    want_b, want_a = factor.adjoint(x_bar, x)
    b_bar, a_bar = factor.adjoint_device(cuda(x_bar), cuda(x))
    assert b_bar.cpu().numpy().tobytes() == want_b.tobytes()
    assert a_bar.cpu().numpy().tobytes() == want_a.tobytes()


def test_autograd_gives_the_oracle_adjoint(gpu):
    from ybus_amd.autograd import solve

    g = O.make_grid(800, 31)
    bp, _, _ = g.dc()
    chol = O.Cholesky(bp)
    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    factor.set_matrix(bp)
    p = O.injections(31, 2, 800, 9)
    w = right_hand_sides(bp.rows, 9, 3)
    a_values = cuda(bp.data).requires_grad_(True)
    b = cuda(p).requires_grad_(True)
    x = solve(factor, a_values, b)
    assert x.detach().cpu().numpy().tobytes() == chol.solve_many(p).tobytes()
    loss = (x * cuda(w)).sum()  # d loss / d x = w exactly
    loss.backward()
    want_b, want_a = O.adjoint(chol, bp, w, chol.solve_many(p))
    assert b.grad.cpu().numpy().tobytes() == want_b.tobytes()
    assert a_values.grad.cpu().numpy().tobytes() == want_a.tobytes()


def test_autograd_in_fp32_refinement_agrees_with_fp64(gpu):
    from ybus_amd.autograd import FP32_REFINED, solve

    g = O.make_grid(1200, 41)
    bp, _, _ = g.dc()
    chol = O.Cholesky(bp)
    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    factor.set_matrix(bp)
    p = O.injections(41, 2, 1200, 7)
    w = cuda(right_hand_sides(bp.rows, 7, 5))
    grads = []
    for mode in [0, FP32_REFINED]:
        a_values = cuda(bp.data).requires_grad_(True)
        b = cuda(p).requires_grad_(True)
        x = solve(factor, a_values, b, mode)
        (x * w).sum().backward()
        grads.append((x.detach(), b.grad, a_values.grad))
    for exact, refined in zip(grads[0], grads[1]):
        assert torch.max(torch.abs(refined - exact)) <= 1e-8 * torch.max(torch.abs(exact))


@pytest.mark.parametrize("mode", [0, 1])
def test_autograd_recomputing_x_gives_the_stored_x_gradients(gpu, mode):
    """The backward that solves for x again gives the bits of the one that kept x, in FP64 (and
    then the oracle's adjoint) and in FP32 refinement alike."""
    from ybus_amd.autograd import solve

    g = O.make_grid(900, 53)
    bp, _, _ = g.dc()
    chol = O.Cholesky(bp)
    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    factor.set_matrix(bp)
    p = O.injections(53, 2, 900, 11)
    w = cuda(right_hand_sides(bp.rows, 11, 6))
    grads = []
    for recompute in [False, True]:
        a_values = cuda(bp.data).requires_grad_(True)
        b = cuda(p).requires_grad_(True)
        x = solve(factor, a_values, b, mode, recompute=recompute)
        (x * w).sum().backward()
        grads.append((x.detach().cpu().numpy(), b.grad.cpu().numpy(), a_values.grad.cpu().numpy()))
    for kept, again in zip(grads[0], grads[1]):
        assert again.tobytes() == kept.tobytes()
    if mode == 0:
        want_b, want_a = O.adjoint(chol, bp, w.cpu().numpy(), chol.solve_many(p))
        assert grads[1][1].tobytes() == want_b.tobytes()
        assert grads[1][2].tobytes() == want_a.tobytes()


def test_autograd_recomputes_from_the_batch_size_up(gpu, monkeypatch):
    """Below the threshold the graph keeps x, so b may change afterwards; from it up the graph
    keeps b, and changing b in place before the backward is refused by autograd's version check
    rather than giving a gradient of the wrong x."""
    import ybus_amd.autograd as AG

    a = random_spd(120, 4, 8)
    factor = Y.GpuFactor.from_cholesky(gpu, O.Cholesky(a))
    factor.set_matrix(a)
    rhs = right_hand_sides(120, 6, 2)
    monkeypatch.setattr(AG, "RECOMPUTE_VALUES", 6 * 120 + 1)
    b = cuda(rhs).requires_grad_(True)
    a_values = cuda(a.data).requires_grad_(True)
    x = AG.solve(factor, a_values, b)
    with torch.no_grad():
        b.add_(1.0)
    x.sum().backward()  # x was kept
    monkeypatch.setattr(AG, "RECOMPUTE_VALUES", 6 * 120)
    b = cuda(rhs).requires_grad_(True)
    x = AG.solve(factor, a_values, b)
    with torch.no_grad():
        b.add_(1.0)
    with pytest.raises(RuntimeError, match="modified by an inplace operation"):
        x.sum().backward()


def test_autograd_refuses_a_factor_changed_before_the_backward(gpu):
    """A refactor between the passes bumps the factor's generation, and the backward raises
    rather than pairing the forward's x with the new factor's transposed solve."""
    from ybus_amd.autograd import solve

    a = random_square_cuda(90, 4, 33)
    lu = O.Lu(a, O.COLAMD)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    factor.set_matrix(a)
    start = factor.generation
    b = cuda(right_hand_sides(90, 3, 4)).requires_grad_(True)
    a_values = cuda(a.data).requires_grad_(True)
    x = solve(factor, a_values, b)
    assert factor.refactor(a.to_csc()) == "reused"
    assert factor.generation == start + 1
    with pytest.raises(RuntimeError, match="changed between the forward"):
        x.sum().backward()
