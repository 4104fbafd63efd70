"""torch.autograd wrapper of the continuous constrained-CCC score and its hand-written adjoint.

`ccc_score(plan, d_parts, descs, pidx, seeds, params)` returns the score S (one per row of `params`, the six
pose parameters rx, ry, rz in degrees and tx, ty, tz in voxels around each seed) as a differentiable torch
tensor; its backward pass uses the gradient the library's adjoint kernel computed with the forward pass
(ettk_ccc_grad_batch), so `S.sum().backward()` fills `params.grad` without any autograd tape over the
FFTs or the interpolator.
"""
from __future__ import annotations

import numpy as np
import torch


class _CCCScore(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, plan, d_parts, descs, pidx, seeds):
        x = params.detach().cpu().numpy().astype(np.float64)
        val, grad = plan.ccc_grad(d_parts, descs, pidx, seeds, x)
        ctx.save_for_backward(torch.from_numpy(grad).to(params.dtype))
        return torch.from_numpy(val).to(params.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out[:, None] * g, None, None, None, None, None


def ccc_score(plan, d_parts, descs, pidx, seeds, params: torch.Tensor) -> torch.Tensor:
    return _CCCScore.apply(params, plan, d_parts, descs, pidx, seeds)
