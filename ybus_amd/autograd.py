"""A sparse solve that PyTorch can differentiate, on the GPU.

``solve(factor, a_values, b)`` returns x with A x = b for each row of ``b`` (a (count, dim)
float64 tensor on the GPU), by the factor's device solve. Its backward pass is the factor's
adjoint: ``b.grad`` gets b_bar = A^-T x_bar and ``a_values.grad`` gets -(sum over the rows of
b_bar x^T) on A's stored pattern, both computed on the GPU without leaving device memory.
``a_values`` is the tensor of A's stored values in the order given to ``factor.set_matrix``; the
solve does not read it (the factor was made from those values on the host), it only carries A's
gradient. Results are bit for bit the CPU's transposed solve and outer-product sums.

The backward needs x. For a large batch the graph keeps b (which the caller holds anyway) and
the backward solves for x again instead of keeping a second (count, dim) tensor alive between
the passes: ``recompute=None`` does so from RECOMPUTE_VALUES values of x up (128 MiB of FP64),
True and False force either way. Every sum in the device solves runs in a fixed order, so the
second solve gives the forward's bits and the gradients are the same either way.

Two rules follow. Where the graph keeps b (recompute), an in-place change to b between the
forward and the backward makes the backward raise (autograd's version check); where it keeps x,
b is not needed again and may change. And the factor must not change between the passes: a
refactor, update_lu_values or set_matrix in between would give a backward over other values
than the forward's, so the backward raises instead (the factor's ``generation``).

PyTorch ships its own HIP runtime: import torch before the first call that loads
libybus_hip.so (creating a ybus_amd.Gpu does), so that the process has one HIP runtime.
"""
from __future__ import annotations

import torch


FP64, FP32_REFINED = 0, 1
RECOMPUTE_VALUES = 1 << 24


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, factor, a_values, b, mode, tolerance, max_steps, recompute):
        x = factor.solve_device(b.contiguous(), mode=mode, tolerance=tolerance, max_steps=max_steps)
        if recompute is None:
            recompute = x.numel() >= RECOMPUTE_VALUES
        ctx.factor = factor
        ctx.generation = getattr(factor, "generation", 0)
        ctx.settings = (mode, tolerance, max_steps)
        ctx.recompute = recompute
        # b is the caller's own tensor (an in-place change to it after this call makes the
        # backward raise, by autograd's version check); x would be one more tensor of its size
        ctx.save_for_backward(b if recompute else x)
        return x

    @staticmethod
    def backward(ctx, x_bar):
        if getattr(ctx.factor, "generation", 0) != ctx.generation:
            raise RuntimeError("the factor's values changed between the forward solve and its backward "
                               "(a refactor, update_lu_values or set_matrix in between)")
        (kept,) = ctx.saved_tensors
        mode, tolerance, max_steps = ctx.settings
        if ctx.recompute:
            x = ctx.factor.solve_device(kept.contiguous(), mode=mode, tolerance=tolerance,
                                        max_steps=max_steps)
        else:
            x = kept
        b_bar, a_bar = ctx.factor.adjoint_device(x_bar, x, mode, tolerance, max_steps)
        return None, a_bar, b_bar, None, None, None, None


def solve(factor, a_values: torch.Tensor, b: torch.Tensor, mode: int = FP64, tolerance: float = 1e-12,
          max_steps: int = 10, recompute: bool | None = None) -> torch.Tensor:
    """x = A^-1 b row by row, differentiable in A's stored values and in b (see the module). With
    mode=FP32_REFINED both the solve and the backward's transposed solve are FP32 sweeps refined in
    FP64 to `tolerance` (refine_normwise's rule); FP64 gives the CPU's bits. `recompute` chooses
    whether the backward solves for x again rather than keeping it (None: from RECOMPUTE_VALUES
    values of x up); then b must not be changed in place before the backward. The factor must not
    change between the passes (the backward raises if it did)."""
    return _Solve.apply(factor, a_values, b, mode, tolerance, max_steps, recompute)
