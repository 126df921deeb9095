"""Writes tests/golden/gauge_wave_history.npz: the oracle's run of the
periodic 1-D gauge wave (ADMGaugeWave1, src/z4c/awa_z4c.cpp:298-351; A = 0.01,
d = 1) on [-0.5, 0.5] in x, harmonic gauge, bt_rk4 at CFL 0.25 to t = 10,
history rows every 0.1 with the reference's output-driven step control
(z4c_amd.awa.HistorySchedule).

The wave depends on x only and every axis is periodic, so a block of n x 8 x
8 cells at the same spacing 1/n evolves exactly the x-profile of the n^3
block (y and z differences of y/z-constant fields are exactly zero); its
history sums cover 64/n^2 of the cells, so the rows are scaled by (n/8)^2
to the n^3 domain.

Stored (n = 64, and n = 32 for the convergence bound):
  times_64, dts_64, rows_64  history rows of the 64^3 run (8 Z4c columns)
  line_64                    the 22 fields on the x line j = k = 4 at t = 10
  alpha_err_64, alpha_err_32 max |alpha - sqrt(H(x - t))| over interior x at
                             t = 10 (exact wave)
  steps_64                   steps taken

Run: python tests/golden/make_gauge_wave_history.py  (about two minutes).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "gr-athena_amd"))

import oracle as orc  # noqa: E402
from z4c_amd.awa import COMMON, GAUGES, HistorySchedule  # noqa: E402
from z4c_amd.initial_data import gauge_wave_adm  # noqa: E402
from z4c_amd.z4c import MeshBlockGeom  # noqa: E402

OUT = os.path.join(HERE, "gauge_wave_history.npz")
T_END, DT_HST, CFL, AMP = 10.0, 0.1, 0.25, 0.01
NT = 8  # transverse cells


def run(n, t_end=T_END, dt_hst=DT_HST):
    """Oracle run on n x 8 x 8 cells at spacing 1/n; returns the schedule's
    rows (scaled to n^3 cells), the final state, the geometry and the step
    count."""
    h = 1.0 / n
    geom = MeshBlockGeom((n, NT, NT), (h, h, h), (-0.5, -0.5 * NT * h, -0.5 * NT * h))
    o = orc.Params(**GAUGES["harmonic"], **COMMON)
    g = orc.Geom.block(geom.nx, geom.dx, geom.xmin)
    u = orc.adm_to_z4c(o, g, np.ascontiguousarray(gauge_wave_adm(geom, A=AMP, d=1.0)))
    orc.periodic_fill(g, u)
    orc.algconstr_range(o, g, u, 0, g.nn1 - 1, 0, g.nn2 - 1, 0, g.nn3 - 1, 1)
    scale = (n / NT) ** 2

    def hist(u):
        adm = orc.z4c_to_adm(o, g, u)
        con = orc.adm_constraints(o, g, u, adm)
        return np.asarray(orc.history(o, g, u, con)) * scale

    sched = HistorySchedule(orc.new_dt(g, CFL), dt_hst, t_end)
    t0, dt0 = sched.initial_row()
    times, dts, rows = [t0], [dt0], [hist(u)]
    steps = 0
    for it in sched.iterations():
        u, _ = orc.rk4_step(o, g, u, it["dt"], between_stages=lambda c: orc.periodic_fill(g, c))
        steps += 1
        if it["row"]:
            times.append(it["time"])
            dts.append(it["row_dt"])
            rows.append(hist(u))
    return np.array(times), np.array(dts), np.array(rows), u, geom, steps


def alpha_error(u, geom, t):
    x, _, _ = geom.coords()
    exact = np.sqrt(1.0 - AMP * np.sin(2.0 * np.pi * (x - t)))
    s = slice(4, 4 + geom.nx[0] + 1)
    return float(np.max(np.abs(u[18, 4, 4, s] - exact[s])))


def main():
    t0 = time.time()
    times, dts, rows, u, geom, steps = run(64)
    print(f"n=64: {steps} steps, {len(rows)} rows, {time.time() - t0:.1f} s", flush=True)
    t1 = time.time()
    _, _, _, u32, geom32, _ = run(32)
    print(f"n=32: {time.time() - t1:.1f} s", flush=True)
    np.savez(OUT, times_64=times, dts_64=dts, rows_64=rows, line_64=u[:, 4, 4, :].copy(),
             alpha_err_64=alpha_error(u, geom, T_END), alpha_err_32=alpha_error(u32, geom32, T_END),
             steps_64=steps)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
