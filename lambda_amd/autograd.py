"""torch.autograd binding of the path-replay adjoint, for checking gradients
with torch's tools.  Both directions run on the GPU through the C ABI
(lam_render forward, lam_adjoint backward); tensors cross as host buffers.

    dp = DiffParams(scene, d=[...], albedo=[...])
    theta = torch.tensor(dp.values(scene), requires_grad=True)
    rgb = render(ctx, scene, dp, params(...), theta)      # (h, w, 3)
    ((rgb - target) ** 2).sum().backward()                 # theta.grad

The forward uploads the scene with θ applied (DiffParams.apply); the backward
re-uploads it if another scene was uploaded in between.
"""
from __future__ import annotations

import numpy as np
import torch


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(fctx, theta, ctx, scene, dp, p):
        variant = dp.apply(scene, theta.detach().to("cpu", torch.float32).numpy())
        ctx.upload(variant)
        rgb = ctx.render(p)
        fctx.lam = (ctx, variant, dp, p)
        return torch.from_numpy(rgb).to(theta.device)

    @staticmethod
    def backward(fctx, grad_rgb):
        ctx, variant, dp, p = fctx.lam
        if ctx.scene is not variant:
            ctx.upload(variant)
        g = np.ascontiguousarray(grad_rgb.detach().to("cpu", torch.float32).numpy())
        grad = ctx.adjoint(p, dp, g)
        return torch.from_numpy(grad).to(grad_rgb.device), None, None, None, None


def render(ctx, scene, dp, p, theta):
    """Linear-RGB image (h, w, 3) of `scene` with parameters θ, differentiable
    in θ (a float32 tensor of dp.n_params entries)."""
    return _Render.apply(theta, ctx, scene, dp, p)
