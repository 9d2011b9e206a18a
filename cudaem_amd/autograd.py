"""Differentiable multislice: torch.autograd.Function over the HIP forward and
adjoint kernels (em_multislice_forward / em_multislice_adjoint).

    Psi = multislice(plan, t, psi0)          # t: [S][ny][nx] complex64
    Psi = multislice_potential(plan, V, psi0)  # V: [S][ny][nx] float32, t = exp(iV)

Psi are the exit waves in Fourier space (unnormalised) of the incident waves
psi0 [B][ny][nx] (real space).  Gradients follow torch's convention for
complex tensors (dL/dRe + i dL/dIm).  `every` stores the wave entering every
`every`-th slice and recomputes the others during the backward pass.
"""
import torch

from .engine import Plan


class _Multislice(torch.autograd.Function):
    @staticmethod
    def forward(ctx, trans, psi0, plan, every, is_potential):
        S = trans.shape[0]
        B = psi0.shape[0]
        psi = psi0.detach().contiguous().clone()
        ck = torch.empty((Plan.n_checkpoints(S, every), B, plan.ny, plan.nx), dtype=torch.complex64,
                         device=psi.device)
        tr = trans.detach().contiguous()
        if is_potential:
            plan.multislice_forward(psi, V=tr, ckpt=ck, every=every)
        else:
            plan.multislice_forward(psi, t=tr, ckpt=ck, every=every)
        ctx.save_for_backward(tr, ck)
        ctx.plan, ctx.every, ctx.is_potential = plan, every, is_potential
        return psi

    @staticmethod
    def backward(ctx, gPsi):
        tr, ck = ctx.saved_tensors
        plan = ctx.plan
        g = gPsi.contiguous().to(torch.complex64).clone()
        gpsi0 = torch.empty_like(g) if ctx.needs_input_grad[1] else None
        if ctx.is_potential:
            gV = torch.empty_like(tr)
            plan.multislice_adjoint(ck, g, V=tr, gV=gV, gpsi0=gpsi0, every=ctx.every)
            gtrans = gV
        else:
            gtrans = plan.multislice_adjoint(ck, g, t=tr, gpsi0=gpsi0, every=ctx.every)
        return (gtrans if ctx.needs_input_grad[0] else None), gpsi0, None, None, None


def multislice(plan, t, psi0, every=1):
    """Exit waves of psi0 through the complex transmissions t (differentiable)."""
    return _Multislice.apply(t, psi0, plan, int(every), False)


def multislice_potential(plan, V, psi0, every=1):
    """Exit waves of psi0 through t = exp(i V) with V sigma-scaled (differentiable in V)."""
    return _Multislice.apply(V, psi0, plan, int(every), True)
