"""torch.autograd.Function over the C-ABI adjoint of an application's explicit step.

`explicit_steps(phi0, L, kappa, ctx, n_steps, checkpoint_every, precision)` returns
phi_N = F^n(phi0) computed by ppf_adjoint_forward; its backward calls ppf_adjoint_vjp
with lam = dJ/dphi_N and returns dJ/dphi0, dJ/dL, dJ/dkappa.  F is the Allen-Cahn step of
an Allen-Cahn application (L the mobility) or the Cahn-Hilliard step of a Cahn-Hilliard
application (phi the concentration, L the mobility M).  Tensors are FP64 device
tensors in the stored layout (periodic images omitted).  Test harness only.
"""

from __future__ import annotations

import torch


class ExplicitSteps(torch.autograd.Function):
    @staticmethod
    def forward(fctx, phi0, L, kappa, ctx, n_steps, checkpoint_every, precision):
        phi0 = phi0.detach().contiguous()
        out = torch.empty_like(phi0)
        torch.cuda.synchronize()
        ctx.adjoint_forward(phi0.data_ptr(), float(L), float(kappa), n_steps, precision,
                            out.data_ptr())
        fctx.save_for_backward(phi0, L.detach(), kappa.detach())
        fctx.args = (ctx, n_steps, checkpoint_every, precision)
        return out

    @staticmethod
    def backward(fctx, grad_out):
        phi0, L, kappa = fctx.saved_tensors
        ctx, n_steps, every, precision = fctx.args
        lam = grad_out.detach().contiguous()
        dphi0 = torch.empty_like(phi0)
        torch.cuda.synchronize()
        dL, dk = ctx.adjoint_vjp(phi0.data_ptr(), lam.data_ptr(), float(L), float(kappa),
                                 n_steps, every, precision, dphi0.data_ptr())
        return (dphi0, torch.tensor(dL, dtype=L.dtype, device=L.device),
                torch.tensor(dk, dtype=kappa.dtype, device=kappa.device), None, None, None, None)


def explicit_steps(phi0, L, kappa, ctx, n_steps, checkpoint_every, precision):
    return ExplicitSteps.apply(phi0, L, kappa, ctx, n_steps, checkpoint_every, precision)


def lumped_mass(ctx):
    m = torch.empty(ctx.info.n_nodes, dtype=torch.float64, device="cuda")
    ctx.lumped_mass(m.data_ptr())
    torch.cuda.synchronize()
    return m
