"""The autograd wrapper's host-side rules, checked on the CPU with a stand-in factor (a dense
solve in torch) so that no device is needed: the backward raises when the factor's values
changed after the forward, keeps b for large batches (an in-place change to b then raises), and
gives the dense adjoint either way."""
import pytest
import torch

from ybus_amd import autograd


class DenseFactor:
    """solve_device / adjoint_device over a dense matrix, with a generation counter like
    GpuFactor's."""

    def __init__(self, a):
        self.a = a
        self.generation = 0
        self.solves = 0

    def change(self, a):
        self.a = a
        self.generation += 1

    def solve_device(self, b, mode=0, tolerance=1e-12, max_steps=10):
        self.solves += 1
        return torch.linalg.solve(self.a, b.T).T

    def adjoint_device(self, x_bar, x, mode=0, tolerance=1e-12, max_steps=10):
        b_bar = torch.linalg.solve(self.a.T, x_bar.T).T
        a_bar = -(b_bar.T @ x)
        return b_bar, a_bar


def system(seed=3, n=6, count=4):
    g = torch.Generator().manual_seed(seed)
    a = torch.rand((n, n), generator=g, dtype=torch.float64) + n * torch.eye(n, dtype=torch.float64)
    b = torch.rand((count, n), generator=g, dtype=torch.float64)
    return a, b


@pytest.mark.parametrize("recompute", [False, True])
def test_gradients_match_the_dense_adjoint(recompute):
    a, b = system()
    f = DenseFactor(a)
    a_values = a.clone().requires_grad_(True)
    b = b.clone().requires_grad_(True)
    x = autograd.solve(f, a_values, b, recompute=recompute)
    x.square().sum().backward()
    want_b = torch.linalg.solve(a.T, (2 * x.detach()).T).T
    assert torch.allclose(b.grad, want_b)
    assert torch.allclose(a_values.grad, -(want_b.T @ x.detach()))
    assert f.solves == (2 if recompute else 1)


def test_a_changed_factor_makes_the_backward_raise():
    a, b = system()
    f = DenseFactor(a)
    a_values = a.clone().requires_grad_(True)
    x = autograd.solve(f, a_values, b.clone().requires_grad_(True))
    f.change(a + torch.eye(a.shape[0], dtype=torch.float64))
    with pytest.raises(RuntimeError, match="changed between the forward"):
        x.sum().backward()


def test_an_in_place_change_to_b_raises_only_where_b_is_kept():
    a, b = system()
    for recompute, raises in ((True, True), (False, False)):
        f = DenseFactor(a)
        a_values = a.clone().requires_grad_(True)
        leaf = b.clone().requires_grad_(True)
        rhs = leaf * 1.0  # a non-leaf the caller may change in place
        x = autograd.solve(f, a_values, rhs, recompute=recompute)
        rhs.mul_(2.0)
        if raises:
            with pytest.raises(RuntimeError):
                x.sum().backward()
        else:
            x.sum().backward()


def test_the_default_recomputes_from_the_size_threshold(monkeypatch):
    a, b = system(count=4)
    monkeypatch.setattr(autograd, "RECOMPUTE_VALUES", 8)
    f = DenseFactor(a)
    x = autograd.solve(f, a.clone().requires_grad_(True), b.clone().requires_grad_(True))
    x.sum().backward()
    assert f.solves == 2  # 24 values >= 8: the backward solved again
