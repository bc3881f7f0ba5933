"""Dispersion and k-linearisation autofocus on the GPU.

The image entropy of a frame is a differentiable function of
theta = (d2, d3, c0, c1, c2, c3): the dispersion phase
phi[j] = d2 (j-N/2)^2 + d3 (j-N/2)^3 and the resampling polynomial
r[j] = c0 + c1 u + c2 u^2 + c3 u^3, u = j/(N-1) (the polynomial of the
reference's generateResamplingCurve, processorconfiguration.cpp:438-459).
Forward and backward are hand-written HIP kernels behind the C ABI
(ope_hip_autofocus_loss / ope_hip_autofocus_grad); ``ChainEntropy`` exposes
them to torch.autograd, and ``autofocus`` runs Adam on them.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _capi


class ChainEntropy(torch.autograd.Function):
    """loss = entropy(frames processed with theta); theta is a float64 CPU
    tensor of 6 parameters, frames a device tensor [A][N] in the context's
    input type."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, context: _capi.Context, frames: torch.Tensor):
        th = theta.detach().to("cpu", torch.float64).numpy()
        stream = torch.cuda.current_stream(frames.device).cuda_stream
        loss, s, _ = context.autofocus_loss(frames.data_ptr(), frames.shape[0], th, stream)
        ctx.context, ctx.frames, ctx.th, ctx.s, ctx.loss = context, frames, th, s, loss
        return theta.new_tensor(loss)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        stream = torch.cuda.current_stream(ctx.frames.device).cuda_stream
        g = ctx.context.autofocus_grad(ctx.frames.data_ptr(), ctx.frames.shape[0], ctx.th, ctx.s, ctx.loss, stream)
        return torch.as_tensor(g, dtype=grad_out.dtype) * grad_out.cpu(), None, None


def entropy(theta: torch.Tensor, context: _capi.Context, frames: torch.Tensor) -> torch.Tensor:
    return ChainEntropy.apply(theta, context, frames)


def autofocus(context: _capi.Context, frames: torch.Tensor, theta0, steps: int = 50, lr: float = 0.02,
              scale=None, free=(True,) * 6):
    """Adam on theta / scale (each parameter in units of its scale); entries
    with free[i] False stay fixed.  Returns (theta, history of losses)."""
    theta0 = np.asarray(theta0, dtype=np.float64)
    scale = np.where(theta0 != 0, np.abs(theta0), 1.0) if scale is None else np.asarray(scale, np.float64)
    sc = torch.as_tensor(scale, dtype=torch.float64)
    mask = torch.as_tensor(np.asarray(free, dtype=np.float64))
    z = torch.as_tensor(theta0 / scale, dtype=torch.float64).requires_grad_(True)
    opt = torch.optim.Adam([z], lr=lr)
    history = []
    for _ in range(steps):
        opt.zero_grad()
        loss = entropy(z * sc, context, frames)
        loss.backward()
        z.grad.mul_(mask)
        opt.step()
        history.append(float(loss))
    theta = (z.detach() * sc).numpy()
    history.append(float(entropy(torch.as_tensor(theta), context, frames)))
    return theta, history
