"""torch.autograd.Function over the adjoint DP (pg_count_grad).

The colorful count of each coloring is a polynomial in the edge weights
(every colorful map contributes the product of the weights of the entries it
uses).  `colorful_count(weights, ...)` returns the raw counts of `ncolorings`
colorings as a differentiable tensor; its backward is the library's adjoint
DP, so d count / d weight[e] is the (weighted) number of colorful maps that
use entry e.  PyTorch only carries the device buffers: forward and backward
both run in libpgbsc.so.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Context, Plan


class ColorfulCount(torch.autograd.Function):
    @staticmethod
    def forward(ctx, weights: torch.Tensor, pgctx: Context, plan: Plan, ncolorings: int, seed: int,
                precision: str):
        if not weights.is_cuda or weights.dtype != torch.float32 or weights.dim() != 1:
            raise ValueError("weights must be a 1-D float32 CUDA tensor with one value per CSC entry")
        if weights.numel() != pgctx.nnz:
            raise ValueError(f"weights has {weights.numel()} values, the graph {pgctx.nnz} entries")
        w = weights.detach().contiguous()
        torch.cuda.current_stream().synchronize()  # the library runs on its own stream
        pgctx.set_weights(device_ptr=w.data_ptr())
        raw = pgctx.count(plan, ncolorings, seed=seed, precision=precision)
        ctx.pg = (pgctx, plan, ncolorings, seed, precision, w)
        return torch.as_tensor(raw, dtype=torch.float64, device=weights.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        pgctx, plan, ncolorings, seed, precision, w = ctx.pg
        cot = grad_out.detach().to("cpu", torch.float64).numpy()
        dtype = torch.float32 if precision == "fp32" else torch.float64
        grad = torch.empty(pgctx.nnz, dtype=dtype, device=w.device)
        torch.cuda.current_stream().synchronize()
        pgctx.set_weights(device_ptr=w.data_ptr())
        pgctx.count_grad_device(plan, ncolorings, grad.data_ptr(), seed=seed, precision=precision,
                                cotangent=np.ascontiguousarray(cot))
        return grad.to(torch.float32), None, None, None, None, None


def colorful_count(weights: torch.Tensor, pgctx: Context, plan: Plan, ncolorings: int, seed: int = 0,
                   precision: str = "fp32") -> torch.Tensor:
    """Differentiable raw colorful counts ([ncolorings] float64) of `plan` on
    `pgctx`'s graph with entry weights `weights` ([nnz] float32, CUDA)."""
    return ColorfulCount.apply(weights, pgctx, plan, ncolorings, seed, precision)
