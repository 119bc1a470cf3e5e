"""torch.autograd wrapper of the device Taylor adjoint (a harness for gradient checks).

``taylor_overlap_objective(v, eps, ...)`` propagates psi_0 through nsteps Taylor steps of
H(v) + eps_s (pol . r) on the device and returns J = sum_j |<phi_j|psi_j(T)>|^2 as a
differentiable torch scalar; its backward calls the hand-written adjoint through the C ABI
(octhip_adjoint_backward).  The objective itself (overlaps, chi_N = <phi|psi> phi) is
evaluated with torch on the device: it is the harness's choice of J, as the target
functional is in the reference's optimal control (src/opt_control/target.F90).
"""
from __future__ import annotations

import torch

from . import api


def complex_rows(b: api.Batch, rows: int) -> torch.Tensor:
    """(rows, pack) complex128 view of a complex packed batch."""
    return torch.view_as_complex(b.ff[:rows].reshape(rows, b.pack_size, 2))


def overlap_objective(phi: api.Batch, psi: api.Batch, rows: int, vol: float):
    """J = sum_j |<phi_j|psi_j>|^2 and chi_N = <phi_j|psi_j> phi_j as a new batch."""
    a, b = complex_rows(phi, rows), complex_rows(psi, rows)
    c = vol * torch.sum(torch.conj(a) * b, dim=0)
    J = torch.sum(torch.abs(c) ** 2)
    chi = api.Batch(psi.nst, psi.nrows, True, psi.pack_size)
    chi.ff[:rows] = torch.view_as_real(a * c[None, :]).reshape(rows, -1)
    return J, chi


class _TaylorOverlap(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v, eps, adj, psi0, phi, vol):
        adj.hm.set_potential(v.detach().cpu().numpy())
        psi = psi0.clone()
        adj.forward(psi, eps.detach().cpu().numpy())
        J, chi = overlap_objective(phi, psi, adj.hm.np, vol)
        ctx.adj, ctx.chi = adj, chi
        ctx.v_device, ctx.eps_device = v.device, eps.device
        return J.detach()

    @staticmethod
    def backward(ctx, gJ):
        deps, dv = ctx.adj.backward(ctx.chi)
        g = gJ.to(torch.float64)
        return (g.to(ctx.v_device) * dv.to(ctx.v_device),
                g.to(ctx.eps_device) * torch.from_numpy(deps).to(ctx.eps_device), None, None, None, None)


def taylor_overlap_objective(v: torch.Tensor, eps: torch.Tensor, adj: api.Adjoint, psi0: api.Batch,
                             phi: api.Batch, vol: float) -> torch.Tensor:
    """Differentiable J(v, eps); psi0 is left unchanged."""
    return _TaylorOverlap.apply(v, eps, adj, psi0, phi, vol)


__all__ = ["taylor_overlap_objective", "overlap_objective", "complex_rows"]
