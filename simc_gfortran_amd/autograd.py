"""torch.autograd binding of the adjoint form (simc_adjoint_run).

    problem = OpticsProblem(ctx, arm, events, mask, p_spec=..., m2=..., seed=...)
    coef = torch.tensor(ctx.recon_coefficients(arm), device="cuda", requires_grad=True)
    offsets = torch.zeros(6, dtype=torch.float64, requires_grad=True)
    loss = optics_loss(problem, coef, offsets)
    loss.backward()          # coef.grad, offsets.grad from the adjoint kernel

The loss is the mean over the masked events of the squared reconstruction
residuals (delta, x', y', y_tar); offsets = (ox, oy, ox', oy', s_p, dth).
Forward and backward are one call of the adjoint kernel: the forward pass
computes the gradients and backward scales them.
"""
from __future__ import annotations

import torch


class OpticsProblem:
    def __init__(self, ctx, arm, events, mask, **run_kw):
        self.ctx = ctx
        self.arm = arm
        self.events = events
        self.mask = mask
        self.kw = run_kw
        self.last = None

    def evaluate(self, coef, offsets):
        off = [float(v) for v in offsets.detach().cpu().tolist()]
        g, out9 = self.ctx.adjoint(self.arm, self.events, self.mask, coef.detach().contiguous(), off, **self.kw)
        self.last = out9
        goff = out9[1:7].to(offsets.device)
        return out9[0].clone(), g.reshape(coef.shape), goff


class OpticsLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coef, offsets, problem):
        loss, gcoef, goff = problem.evaluate(coef, offsets)
        ctx.save_for_backward(gcoef, goff)
        return loss

    @staticmethod
    def backward(ctx, grad):
        gcoef, goff = ctx.saved_tensors
        return grad * gcoef, (grad.to(goff.device) * goff), None


def optics_loss(problem: OpticsProblem, coef: torch.Tensor, offsets: torch.Tensor) -> torch.Tensor:
    return OpticsLoss.apply(coef, offsets, problem)
