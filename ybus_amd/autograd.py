"""A sparse solve that PyTorch can differentiate, on the GPU.

``solve(factor, a_values, b)`` returns x with A x = b for each row of ``b`` (a (count, dim)
float64 tensor on the GPU), by the factor's device solve. Its backward pass is the factor's
adjoint: ``b.grad`` gets b_bar = A^-T x_bar and ``a_values.grad`` gets -(sum over the rows of
b_bar x^T) on A's stored pattern, both computed on the GPU without leaving device memory.
``a_values`` is the tensor of A's stored values in the order given to ``factor.set_matrix``; the
solve does not read it (the factor was made from those values on the host), it only carries A's
gradient. Results are bit for bit the CPU's transposed solve and outer-product sums.

PyTorch ships its own HIP runtime: import torch before the first call that loads
libybus_hip.so (creating a ybus_amd.Gpu does), so that the process has one HIP runtime.
"""
from __future__ import annotations

import torch


FP64, FP32_REFINED = 0, 1


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, factor, a_values, b, mode, tolerance, max_steps):
        x = factor.solve_device(b.contiguous(), mode=mode, tolerance=tolerance, max_steps=max_steps)
        ctx.factor = factor
        ctx.settings = (mode, tolerance, max_steps)
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, x_bar):
        (x,) = ctx.saved_tensors
        mode, tolerance, max_steps = ctx.settings
        b_bar, a_bar = ctx.factor.adjoint_device(x_bar, x, mode, tolerance, max_steps)
        return None, a_bar, b_bar, None, None, None


def solve(factor, a_values: torch.Tensor, b: torch.Tensor, mode: int = FP64, tolerance: float = 1e-12,
          max_steps: int = 10) -> torch.Tensor:
    """x = A^-1 b row by row, differentiable in A's stored values and in b (see the module). With
    mode=FP32_REFINED both the solve and the backward's transposed solve are FP32 sweeps refined in
    FP64 to `tolerance` (refine_normwise's rule); FP64 gives the CPU's bits."""
    return _Solve.apply(factor, a_values, b, mode, tolerance, max_steps)
