"""torch.autograd binding of the render loss and its hand-written adjoint (tdi_render_loss_grad).

    loss = render_loss(batch, theta, dverts)      # [n] float64, one loss per frame
    loss.sum().backward()                         # theta.grad [n, K], dverts.grad [n, V, 3]

The loss of a frame is the sum over its visible pixels of |g^_r - g^_s|^2 (normalised gradients
of the blurred render and of the blurred observation) plus the squared reprojection error of the
51 landmarks; dverts are per-vertex offsets of the model-space vertices.  The forward runs the
library's render; the backward recomputes the render (rasterisation included) from the saved
inputs and runs the adjoint kernels, holding the visibility of that render fixed.  Gradients are
returned for the pose, expression, identity and light unknowns of theta (albedo unknowns get 0)
and for every vertex offset.  There is no CPU path: the library must be built and a GPU present.
"""
from __future__ import annotations

import importlib

import torch

tdi = importlib.import_module("3di_amd.tdi")


class RenderLoss(torch.autograd.Function):
    """Per-frame render loss of a tdi.Batch; backward = the library's adjoint."""

    @staticmethod
    def forward(ctx, theta, dverts, batch):
        loss, _, _ = batch.render_loss(theta.detach(), None if dverts is None else dverts.detach(), grad=False)
        ctx.batch = batch
        ctx.has_dverts = dverts is not None
        if dverts is not None:
            ctx.save_for_backward(theta.detach(), dverts.detach())
        else:
            ctx.save_for_backward(theta.detach())
        return loss

    @staticmethod
    def backward(ctx, gloss):
        saved = ctx.saved_tensors
        theta = saved[0]
        dverts = saved[1] if ctx.has_dverts else None
        _, gth, gv = ctx.batch.render_loss(theta, dverts, gloss=gloss.detach(), grad=True)
        gth = gth.to(theta.dtype) if ctx.needs_input_grad[0] else None
        gv = gv if (ctx.has_dverts and ctx.needs_input_grad[1]) else None
        return gth, gv, None


def render_loss(batch, theta, dverts=None):
    """Per-frame loss [n] (float64) of theta [n, K] (and vertex offsets [n, V, 3]), differentiable."""
    return RenderLoss.apply(theta, dverts, batch)
