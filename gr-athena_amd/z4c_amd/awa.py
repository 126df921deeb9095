"""Apples-with-apples gauge-wave evolution on the device (the periodic
single-block configuration).

One vertex-centred MeshBlock covers the periodic cube [-L/2, L/2]^3 with
``n`` cells per side.  The initial data is the 1-D gauge wave
(ADMGaugeWave1, src/z4c/awa_z4c.cpp:298-351):

    ds^2 = -H dt^2 + H dx^2 + dy^2 + dz^2,  H = 1 - A sin(2 pi (x - t)/d).

The evolution is bt_rk4 in the GR_Z4c stage order with the periodic
self-exchange after every stage (src/task_list/gr/task_list_gr_z4c.cpp:45-89).
At the history cadence it runs Z4cToADM + ADMConstraints and the Z4c columns
of the .hst output (src/outputs/history.cpp:178-205, header :416-430).

Gauges:
* "harmonic": lapse_oplog = 0, lapse_harmonic = 1, no shift driver -- the
  gauge in which the gauge wave is an exact solution;
* "puncture": the values of inputs/z4c/awa/z4c_AwA_gauge_wave_1s.inp
  (1+log slicing, Gamma-driver shift, eta = 2).
Both use diss = 0.02, chi_div_floor = 1e-5, damp_kappa1 = 0.02 from that
input.
"""
from __future__ import annotations

import math

import numpy as np

from ._capi import Params
from .initial_data import gauge_wave_adm
from .outputs import write_hst
from .z4c import MeshBlockGeom, Z4c

GAUGES = {
    "harmonic": dict(lapse_oplog=0.0, lapse_harmonicf=1.0, lapse_harmonic=1.0,
                     lapse_advect=1.0, lapse_K=0.0, shift_Gamma=0.0, shift_eta=0.0,
                     shift_advect=0.0, shift_alpha2Gamma=0.0, shift_H=0.0),
    "puncture": dict(lapse_oplog=2.0, lapse_harmonicf=1.0, lapse_harmonic=0.0,
                     lapse_advect=1.0, lapse_K=0.0, shift_Gamma=1.0, shift_eta=2.0,
                     shift_advect=1.0, shift_alpha2Gamma=0.0, shift_H=0.0),
}
COMMON = dict(diss=0.02, chi_div_floor=1e-5, damp_kappa1=0.02, damp_kappa2=0.0)


def gauge_wave_setup(n=64, length=1.0, amplitude=0.01, d=1.0, gauge="harmonic",
                     device=0):
    """Block, parameters, device context and initial Z4c state: ADMToZ4c,
    the periodic exchange, AlgConstr on the ghosts and the derivative
    arrays (Mesh::FinalizeZ4cADMGhosts, src/mesh/mesh_communicate.cpp:100-168)."""
    import torch
    h = length / n
    geom = MeshBlockGeom.cube(n, h, -0.5 * length)
    params = Params(**GAUGES[gauge], **COMMON)
    z = Z4c([geom], params, device=device)
    adm = z.upload([gauge_wave_adm(geom, A=amplitude, d=d)])
    u = z.zeros()
    z.ADMToZ4c(adm, u)
    z.periodic_fill(u)
    nn = geom.nn
    z.AlgConstr(u, rng=(0, nn[0] - 1, 0, nn[1] - 1, 0, nn[2] - 1), skip_physical=True)
    z.PrepareZ4cDerivatives(u)
    torch.cuda.synchronize()
    return geom, params, z, u


def constraint_history(z, u):
    """Z4cToADM + ADMConstraints + the .hst constraint columns of ``u``."""
    from .z4c import Z4C_NADM, Z4C_NCON
    adm = z.zeros(Z4C_NADM)
    con = z.zeros(Z4C_NCON)
    z.Z4cToADM(u, adm)
    z.ADMConstraints(con, adm, None, u)
    return z.history(u, con)


def run_gauge_wave(n=64, t_end=10.0, cfl=0.25, amplitude=0.01, d=1.0, gauge="harmonic",
                   hst_every=None, hst_path=None, progress=None):
    """Evolve to ``t_end``; returns (times, rows, final state on the host).
    ``hst_every`` (steps) defaults to the step count of t = 0.1; rows hold the
    eight Z4c history columns.  ``hst_path`` writes the rows as a .hst file."""
    import torch
    geom, params, z, u = gauge_wave_setup(n, 1.0, amplitude, d, gauge)
    dt = z.NewBlockTimeStep(cfl)
    nsteps = int(round(t_end / dt))
    if not math.isclose(nsteps * dt, t_end, rel_tol=1e-12):
        raise ValueError("t_end must be a whole number of steps of cfl*dx")
    hst_every = hst_every or max(1, int(round(0.1 / dt)))
    work = z.make_step_work()
    times, rows = [0.0], [constraint_history(z, u)]
    for s in range(1, nsteps + 1):
        # the fused step with the periodic fill after every stage (the
        # stage sequence's order and sums; ADMConstraints prepares its own
        # derivatives)
        z.rk4_step_fused(dt, u, work, periodic=True)
        if s % hst_every == 0 or s == nsteps:
            times.append(s * dt)
            rows.append(constraint_history(z, u))
            if progress:
                progress(s, nsteps, rows[-1])
    torch.cuda.synchronize()
    if hst_path:
        write_hst(hst_path, times, rows, dt)
    return np.array(times), np.array(rows), z.download(u, 22)[0], geom


def exact_lapse(geom, t, amplitude=0.01, d=1.0):
    """alpha = sqrt(H(x - t)) of the exact gauge-wave solution on the block."""
    x, _, _ = geom.coords()
    H = 1.0 - amplitude * np.sin(2.0 * np.pi * (x - t) / d)
    return np.sqrt(H)
