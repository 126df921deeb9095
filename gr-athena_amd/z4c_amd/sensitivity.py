"""Adjoint sensitivity of the Hamiltonian-constraint norm after an evolution.

``hamiltonian_sensitivity`` evolves a batched state with the fused bt_rk4
step, evaluates J = sum dx dy dz Ht^2 (z4c_hamiltonian_norm: the
Hamiltonian constraint formed inside the right-hand side,
src/z4c/calculate_z4c_rhs.cpp:606-624, over the history output's cell loop)
and propagates dJ/du back through the last ``window`` steps with the step
VJP (z4c_rk4_step_vjp, or z4c_rk4_step_ghosts_vjp for steps that run the
ghost update after every stage).  The result is the sensitivity of J to the gauge
parameters shift_eta, damp_kappa1, damp_kappa2 through those steps (the
state at the start of the window held fixed) and to that state.

The states at the start of every step of the window are kept in device
memory (one 22-field register each); each step's VJP recomputes its stages
from them (stage checkpointing inside the step VJP).
"""
from __future__ import annotations

import torch

from ._capi import PBAR


def hamiltonian_sensitivity(z, u0, dt, nsteps, window, graph=True, exchange=False,
                            physical=False):
    """Returns a dict: J, dJ/d(shift_eta, damp_kappa1, damp_kappa2) through
    the last ``window`` of ``nsteps`` steps ("pbar"), dJ/du at step
    nsteps - window ("ubar", flat register), the final state ("u") and the
    state at the start of the window ("u_window").

    Without ``exchange`` the steps hold every block's ghosts at u_n
    (z4c_rk4_step, independent blocks) and the adjoint is z4c_rk4_step_vjp.
    With ``exchange`` (``z.set_exchange`` called first) every stage runs the
    ghost update, across levels on a block tree, plus the outflow fill with
    ``physical``, as GR_Z4c orders it (z4c_rk4_step_ghosts), and the adjoint
    is z4c_rk4_step_ghosts_vjp, which carries the cotangent through the
    update's transpose after every stage."""
    if not 0 < window <= nsteps:
        raise ValueError("need 0 < window <= nsteps")
    if physical and not exchange:
        raise ValueError("the outflow fill runs with the exchange here")
    work = z.make_step_work()
    u = u0.clone()

    def step(x):
        z.rk4_step_fused(dt, x, work, graph=graph, exchange=exchange, physical=physical)

    for _ in range(nsteps - window):
        step(u)
    states = []
    for _ in range(window):
        states.append(u.clone())
        step(u)
    J = z.hamiltonian_norm(u)
    ubar = z.hamiltonian_norm_vjp(u)
    pbar = torch.zeros(len(PBAR), dtype=torch.float64, device=u.device)
    for s in reversed(states):
        if exchange:
            ubar, pb = z.rk4_step_ghosts_vjp(dt, s, ubar, 1, physical=physical)
        else:
            ubar, pb = z.rk4_step_vjp(dt, s, ubar, 1)
        pbar += pb[0]
    return {"J": J, "pbar": dict(zip(PBAR, pbar.tolist())), "ubar": ubar, "u": u,
            "u_window": states[0]}
