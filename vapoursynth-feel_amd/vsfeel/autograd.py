"""Differentiable BM3Dv2 (basic hard-threshold pass, radius 0) for PyTorch.

With the block-match groups and the threshold masks held fixed, the filter
is out = A x / D: A is a fixed symmetric linear map (per group: gather the
8 blocks, 3-D DCT, keep mask, inverse DCT, weight, scatter back) and D the
per-pixel weight sums, which depend on the masks only.  The gradient is
therefore d(loss)/dx = A (g / D), computed by the same kernels with the
saved masks applied instead of a threshold.  Groups and masks are piecewise
constant in x; their derivative is zero almost everywhere.

``recompute=True`` saves only x and rebuilds groups, masks, weights and D in
the backward pass (for batches whose saved state would not fit).
"""

from __future__ import annotations

import torch

from .torch_bm3d import Instance, _instance

__all__ = ["bm3d_differentiable"]


class _Bm3dFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, inst: Instance, recompute: bool):
        out, saved = inst.forward_saved(x)
        ctx.inst = inst
        ctx.recompute = recompute
        if recompute:
            ctx.save_for_backward(x)
        else:
            ctx.save_for_backward(*saved)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        inst = ctx.inst
        if ctx.recompute:
            (x,) = ctx.saved_tensors
            _, saved = inst.forward_saved(x)
        else:
            saved = ctx.saved_tensors
        return inst.backward(grad_out, saved), None, None


def bm3d_differentiable(x: torch.Tensor, *, sigma=3.0, block_step=8, bm_range=9, recompute=False,
                        fp_mode=1) -> torch.Tensor:
    """BM3Dv2 basic pass (radius 0) with a backward pass; x is (N, H, W) float32."""
    squeeze = x.dim() == 2
    if squeeze:
        x = x.unsqueeze(0)
    x = x.contiguous()
    kw = dict(sigma=sigma, block_step=block_step, bm_range=bm_range, radius=0, ps_num=2, ps_range=4,
              extractor_exp=0, fp_mode=fp_mode)
    inst = _instance(x, False, kw)
    out = _Bm3dFunction.apply(x, inst, recompute)
    return out.squeeze(0) if squeeze else out
