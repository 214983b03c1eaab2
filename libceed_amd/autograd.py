"""PyTorch autograd for the variable-coefficient diffusion operator.

``kappa_diffusion(u, kappa, op)`` returns v = A(kappa) u with
A(kappa) u = E^T G^T (kappa S) G E u (S = w det J J^-1 J^-T at each
quadrature point, kappa given per quadrature point, layout [elem][Q^3]).
Its backward runs the fused adjoint kernel once: for an upstream gradient
lambda = dL/dv it returns dL/du = A(kappa)^T lambda and
dL/dkappa[e, q] = grad(lambda)^T S grad(u) at each quadrature point.
"""

from __future__ import annotations

import torch

from . import ops


class KappaDiffusion(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u: torch.Tensor, kappa: torch.Tensor, op: ops.HexOperator) -> torch.Tensor:
        if op.qf != ops.QF_DIFFUSION_KAPPA or op.dtype != torch.float64:
            raise ValueError("KappaDiffusion needs an FP64 QF_DIFFUSION_KAPPA operator")
        u, kappa = u.contiguous(), kappa.contiguous()
        op.set_kappa(kappa)
        v = torch.empty_like(u)
        op.apply(u, v)
        ctx.op = op
        ctx.save_for_backward(u, kappa)
        return v

    @staticmethod
    def backward(ctx, lam: torch.Tensor):
        u, kappa = ctx.saved_tensors
        op = ctx.op
        op.set_kappa(kappa)  # the plan may have been re-used with another kappa since forward
        gu = torch.empty_like(u)
        gk = torch.empty_like(kappa)
        op.apply_adjoint(u, lam.contiguous(), gu, gk)
        return gu, gk, None


def kappa_diffusion(u: torch.Tensor, kappa: torch.Tensor, op: ops.HexOperator) -> torch.Tensor:
    return KappaDiffusion.apply(u, kappa, op)
