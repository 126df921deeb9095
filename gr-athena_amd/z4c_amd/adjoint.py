"""torch.autograd bindings of the hand-written adjoint kernels.

``Z4cRHSFunction`` and ``Z4cStepFunction`` wrap the batched Z4cRHS
(src/z4c/calculate_z4c_rhs.cpp:29-916) and one fused bt_rk4 step
(``Z4c.rk4_step_fused``) as differentiable functions of the state and of
the gauge parameters shift_eta, damp_kappa1 and damp_kappa2
(src/z4c/z4c.hpp:374-430).  Forward passes call the C-ABI forward entry
points; backward passes call z4c_rhs_vjp / z4c_rk4_step_vjp, so gradients
come from the gfx950 adjoint kernels (the reference has no derivative code;
these are the sensitivities to the initial data and the gauge parameters).
"""
from __future__ import annotations

import torch

from ._capi import PBAR, Params

_PNAMES = tuple(PBAR)


def _params(z, eta, kappa1, kappa2):
    d = z.params.as_dict()
    for name, v in zip(_PNAMES, (eta, kappa1, kappa2)):
        d[name] = float(v)
    return Params(**d)


def _scalar(x, ref):
    return torch.as_tensor(x, dtype=torch.float64, device=ref.device)


class Z4cRHSFunction(torch.autograd.Function):
    """rhs = Z4cRHS(u; shift_eta, damp_kappa1, damp_kappa2) over every block
    of ``z`` (a ``Z4c``); u and rhs are flat 22-field registers."""

    @staticmethod
    def forward(ctx, u, eta, kappa1, kappa2, z):
        p = _params(z, eta, kappa1, kappa2)
        u = u.detach().contiguous()
        rhs = z.zeros()
        saved = z.params
        z.params = p
        try:
            z.Z4cRHS(u, None, rhs)
        finally:
            z.params = saved
        ctx.z, ctx.p = z, p
        ctx.save_for_backward(u)
        return rhs

    @staticmethod
    def backward(ctx, g):
        (u,) = ctx.saved_tensors
        ubar, pbar = ctx.z.rhs_vjp(u, g.contiguous(), 1, params=ctx.p)
        return ubar, pbar[0, 0], pbar[0, 1], pbar[0, 2], None


class Z4cStepFunction(torch.autograd.Function):
    """u_{n+1} = one fused bt_rk4 step of u_n (ghosts held at u_n,
    AlgConstr after every stage) with time step dt."""

    @staticmethod
    def forward(ctx, u, eta, kappa1, kappa2, z, dt):
        p = _params(z, eta, kappa1, kappa2)
        u = u.detach().contiguous()
        out = u.clone()
        saved = z.params
        z.params = p
        try:
            z.rk4_step_fused(dt, out, z.make_step_work())
        finally:
            z.params = saved
        ctx.z, ctx.p, ctx.dt = z, p, dt
        ctx.save_for_backward(u)
        return out

    @staticmethod
    def backward(ctx, g):
        (u,) = ctx.saved_tensors
        ubar, pbar = ctx.z.rk4_step_vjp(ctx.dt, u, g.contiguous(), 1, params=ctx.p)
        return ubar, pbar[0, 0], pbar[0, 1], pbar[0, 2], None, None


def rhs(z, u, eta=None, kappa1=None, kappa2=None):
    """Differentiable Z4cRHS; parameters default to z.params."""
    eta = _scalar(z.params.shift_eta if eta is None else eta, u)
    kappa1 = _scalar(z.params.damp_kappa1 if kappa1 is None else kappa1, u)
    kappa2 = _scalar(z.params.damp_kappa2 if kappa2 is None else kappa2, u)
    return Z4cRHSFunction.apply(u, eta, kappa1, kappa2, z)


def rk4_step(z, u, dt, eta=None, kappa1=None, kappa2=None):
    """Differentiable fused bt_rk4 step; parameters default to z.params."""
    eta = _scalar(z.params.shift_eta if eta is None else eta, u)
    kappa1 = _scalar(z.params.damp_kappa1 if kappa1 is None else kappa1, u)
    kappa2 = _scalar(z.params.damp_kappa2 if kappa2 is None else kappa2, u)
    return Z4cStepFunction.apply(u, eta, kappa1, kappa2, z, dt)
