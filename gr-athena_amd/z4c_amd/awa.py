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


class HistorySchedule:
    """The main loop's time stepping around the history output, as the
    reference runs it with task_triggers/adjust_mesh_dt = true (its default,
    src/main.cpp:114-125):

    * the Z4c_ADM_constraints/hst and hst triggers (dt = the hst output's
      dt, force_first_iter; PopulateTrigger, src/main_triggers.hpp:565-589)
      shorten the step that would pass their next time so that it lands on
      it (Trigger::reduce_mesh_dt, :80-130, via AdjustFromAny_mesh_dt);
    * after the step: the constraints are evaluated when the trigger is
      satisfied (Z4c_DerivedQuantities, src/main.hpp:970), the triggers
      advance (Trigger::update, :59-77), time += dt, and the next dt is the
      block dt, or min(2 dt, block dt) when no trigger shortened the step
      (Mesh::NewTimeStepBegin, src/mesh/mesh.cpp:1822-1850);
    * a history row is written when time >= next_time (next_time starts at
      the start time and grows by the output dt, src/outputs/outputs.cpp:
      208, 1405-1412, src/outputs/history.cpp:451), inside the loop only
      while time < tlim, and once more after it (the final outputs,
      src/main.cpp:356-361);
    * dt is clamped to tlim - time (Mesh::NewTimeStepFinish,
      src/mesh/mesh.cpp:1908-1909).

    ``iterations()`` yields one dict per loop iteration: ``dt`` of the step,
    ``constraints`` (evaluate them on the new state), ``row`` (write a row
    after the step) with its ``time`` and ``row_dt`` (the mesh dt when the
    row is written, the .hst dt column)."""

    def __init__(self, block_dt, dt_hst, tlim, time=0.0):
        self.block_dt = float(block_dt)
        self.tlim = float(tlim)
        self.time = float(time)
        self.dt = self.block_dt
        if self.time < self.tlim and self.tlim - self.time < self.dt:
            self.dt = self.tlim - self.time
        # trigger state (both triggers share dt and force_first_iter)
        self.tdt = float(dt_hst)
        nd = int(self.time / self.tdt) if self.tdt > 0 else 0
        self.t_last = nd * self.tdt
        self.t_next = min(self.time + self.dt, (nd + 1) * self.tdt)
        self.force_first = True
        self.next_time = self.time  # the hst output's next_time

    def _satisfied(self):
        if self.tdt <= 0:
            return False
        if self.force_first:
            return True
        t_end = self.time + self.dt
        if self.t_next > self.tlim:  # the last point(s): trigger dt -> mesh dt
            self.tdt = self.dt
            return True
        return (t_end >= self.t_next) or (self.time >= self.tlim)

    def _reduce(self):
        t_end = self.time + self.dt
        with np.errstate(divide="ignore", invalid="ignore"):
            diff_tol = abs(1.0 - np.float64(self.t_next) / np.float64(self.time)) > 1e-13
        if t_end > self.t_next and self.time < self.t_next and diff_tol:
            self.dt = self.t_next - self.time
            self.t_next = self.time + self.dt
            self.t_last = self.t_next - self.tdt
            return True
        return False

    def _update(self):
        self.t_last = self.t_next
        if self.force_first:
            self.t_next = (int(self.t_last / self.tdt) + 1) * self.tdt
        else:
            self.t_next = self.t_last + self.tdt
        self.force_first = False

    def initial_row(self):
        """The row of the initial data (the outputs made before the loop)."""
        t = self.time
        self.next_time += self.tdt
        return t, self.dt

    def iterations(self):
        while self.time < self.tlim:
            adjusted = self._satisfied() and self._reduce()
            dt = self.dt
            con = self._satisfied()
            if self._satisfied():
                self._update()
            self.time += dt
            self.dt = self.block_dt if adjusted else min(2.0 * self.dt, self.block_dt)
            row = False
            if self.time < self.tlim and self.time >= self.next_time:
                row = True
                self.next_time += self.tdt
            if self.time < self.tlim and self.tlim - self.time < self.dt:
                self.dt = self.tlim - self.time
            if self.time >= self.tlim:
                row = True  # the final outputs
            yield {"dt": dt, "constraints": con, "row": row, "time": self.time,
                   "row_dt": self.dt}


def run_gauge_wave(n=64, t_end=10.0, cfl=0.25, amplitude=0.01, d=1.0, gauge="harmonic",
                   dt_hst=0.1, hst_path=None, progress=None):
    """Evolve to ``t_end`` with the reference's output-driven step control
    (HistorySchedule); returns (times, dts, rows, final state on the host,
    geometry, steps taken).  Rows hold the eight Z4c history columns at the
    initial data and at every output time; ``dts`` is each row's dt column.
    ``hst_path`` writes them as a .hst file."""
    import torch
    geom, params, z, u = gauge_wave_setup(n, 1.0, amplitude, d, gauge)
    sched = HistorySchedule(z.NewBlockTimeStep(cfl), dt_hst, t_end)
    work = z.make_step_work()
    t0, dt0 = sched.initial_row()
    times, dts, rows = [t0], [dt0], [constraint_history(z, u)]
    nsteps = 0
    for it in sched.iterations():
        # the fused step with the periodic fill after every stage (the
        # stage sequence's order and sums; ADMConstraints prepares its own
        # derivatives)
        z.rk4_step_fused(it["dt"], u, work, periodic=True)
        nsteps += 1
        if it["row"]:
            if not it["constraints"]:
                raise RuntimeError("history row without a constraint evaluation")
            times.append(it["time"])
            dts.append(it["row_dt"])
            rows.append(constraint_history(z, u))
            if progress:
                progress(nsteps, it["time"], rows[-1])
    torch.cuda.synchronize()
    if hst_path:
        write_hst(hst_path, times, rows, dts, nbtotal=1)
    return (np.array(times), np.array(dts), np.array(rows), z.download(u, 22)[0], geom,
            nsteps)


def exact_lapse(geom, t, amplitude=0.01, d=1.0):
    """alpha = sqrt(H(x - t)) of the exact gauge-wave solution on the block."""
    x, _, _ = geom.coords()
    H = 1.0 - amplitude * np.sin(2.0 * np.pi * (x - t) / d)
    return np.sqrt(H)
