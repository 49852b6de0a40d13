"""Parameter fitting support: the GFN-FF gradient as a differentiable function
of the per-element parameter scales theta (repulsion charge, repulsion
exponent, D4 sqrt(Z r4r2)), for force-matching fits with torch.

The backward pass is the engine's adjoint kernel (gfnff_param_adjoint): for an
upstream gradient w = dLoss/dg it returns dLoss/dtheta = sum_i w_i . dg_i/dtheta
without materialising dg/dtheta; pair terms are recomputed, not stored.
"""
from __future__ import annotations

import numpy as np
import torch

from .gfnff import GFNFFMethod


class GradientOfTheta(torch.autograd.Function):
    """g = GFNFF gradient (natoms x 3, Eh/Bohr) at parameter scales theta.

    One method object per conformer; the backward pass uses the state of that
    method's last evaluation, so call backward before evaluating the same
    method again (the usual order of a loss.backward())."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, method: GFNFFMethod) -> torch.Tensor:
        method.setParamScales(theta.detach().cpu().numpy())
        method.calculateEnergy(True)
        if method.hasError():
            raise RuntimeError(method.getErrorMessage())
        ctx.method = method
        return torch.from_numpy(method.getGradient().copy()).to(theta.dtype)

    @staticmethod
    def backward(ctx, grad_g: torch.Tensor):
        d = ctx.method.paramAdjoint(grad_g.detach().cpu().numpy().astype(np.float64))
        return torch.from_numpy(d).to(grad_g.dtype), None


def force_matching_loss(theta: torch.Tensor, methods, g_refs) -> torch.Tensor:
    """sum over conformers of |g(theta) - g_ref|^2 (= |F - F_ref|^2)."""
    loss = torch.zeros((), dtype=theta.dtype)
    for m, gr in zip(methods, g_refs):
        g = GradientOfTheta.apply(theta, m)
        loss = loss + ((g - torch.as_tensor(gr, dtype=theta.dtype)) ** 2).sum()
    return loss
