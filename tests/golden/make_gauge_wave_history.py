"""Writes tests/golden/gauge_wave_history.npz: the oracle's run of the
periodic 1-D gauge wave (ADMGaugeWave1, src/z4c/awa_z4c.cpp:298-351; A = 0.01,
d = 1) on [-0.5, 0.5] in x, harmonic gauge, bt_rk4 at CFL 0.25 to t = 10,
history rows every 0.1 with the reference's output-driven step control
(z4c_amd.awa.HistorySchedule).

The run is the full n^3 block.  A thinner periodic slab (n x 8 x 8 cells)
is not equivalent: the initial AlgConstr skips the ghosts (skip_physical), so
every block face carries a ~3e-13 jump in g~ between ghost and interior
vertices, and its contribution to the constraint norms scales with the
face-to-volume ratio (H-norm2 at t = 0 is 4.1e-18 on 64 x 8 x 8 cells and
6.7e-19 on 64^3).

Stored (n = 64, and n = 32 for the convergence bound):
  times_64, dts_64, rows_64  history rows of the 64^3 run (8 Z4c columns)
  line_64                    the 22 fields on the x line j = k = 4 at t = 10
  alpha_err_64, alpha_err_32 max |alpha - sqrt(H(x - t))| over interior x at
                             t = 10 (exact wave)
  steps_64                   steps taken

Each stage right-hand side is the oracle's Z4cRHS evaluated on 8 overlapping
sub-blocks of (n/2)^3 cells in parallel threads and assembled: the RHS at a
vertex reads the state within 4 vertices of it only, so the assembly equals
the whole-block RHS bit for bit (checked first, on a 16^3 block, together
with the whole step against oracle.rk4_step).

Run: python tests/golden/make_gauge_wave_history.py  (8 threads; about half an
hour).
"""
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

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


POOL = ThreadPoolExecutor(8)


def rhs_split(o, geom, u):
    """Z4cRHS of the whole block from its 8 octant sub-blocks (each with its
    own 4 ghost layers cut from u), run in parallel threads."""
    nx = np.array(geom.nx)
    half = nx // 2
    ng = orc.NGHOST
    out = np.zeros_like(u)
    jobs = []
    for c in range(2):
        for b in range(2):
            for a in range(2):
                lo = np.array([a, b, c]) * half  # first owned vertex, in cells
                sub_nx = np.where(np.array([a, b, c]) == 0, half, nx - half)
                sl = tuple(slice(int(lo[d]), int(lo[d] + sub_nx[d] + 1 + 2 * ng)) for d in (2, 1, 0))
                us = np.ascontiguousarray(u[(slice(None),) + sl])
                gs = orc.Geom.block(sub_nx, geom.dx, np.array(geom.xmin) + lo * np.array(geom.dx))
                inner = tuple(slice(int(lo[d]) + ng, int(lo[d] + sub_nx[d]) + ng + 1) for d in (2, 1, 0))
                jobs.append((POOL.submit(orc.rhs, o, gs, us), inner))
    for fut, inner in jobs:
        r = fut.result()
        ri = tuple(slice(ng, ng + (s.stop - s.start)) for s in inner)
        out[(slice(None),) + inner] = r[(slice(None),) + ri]
    return out


def rk4_step_split(o, g, geom, u, dt, between_stages):
    """oracle.rk4_step with rhs_split for the stage right-hand sides."""
    un = u.copy()
    ks = []
    cur = u
    for s in range(4):
        ks.append(rhs_split(o, geom, cur))
        ws = orc.BT_RK4_A[s + 1][:len(ks)] if s < 3 else orc.BT_RK4_B
        cur = orc.bt_sum(g, un, ks, ws, dt)
        between_stages(cur)
        orc.algconstr(o, g, cur)
    return cur


def check_split(n=16):
    """rhs_split and rk4_step_split equal the oracle's whole-block forms bit
    for bit on an n^3 block of the gauge wave."""
    h = 1.0 / n
    geom = MeshBlockGeom.cube(n, h, -0.5)
    o = orc.Params(**GAUGES["puncture"], **COMMON)
    g = orc.Geom.block(geom.nx, geom.dx, geom.xmin)
    u = orc.adm_to_z4c(o, g, np.ascontiguousarray(gauge_wave_adm(geom, A=AMP, d=1.0)))
    orc.periodic_fill(g, u)
    orc.algconstr(o, g, u)
    u += 1e-3 * np.random.default_rng(0).standard_normal(u.shape)
    orc.periodic_fill(g, u)
    inner = (slice(None),) + (slice(orc.NGHOST, orc.NGHOST + n + 1),) * 3
    assert np.array_equal(rhs_split(o, geom, u)[inner], orc.rhs(o, g, u)[inner])
    fill = lambda c: orc.periodic_fill(g, c)  # noqa: E731
    a = rk4_step_split(o, g, geom, u.copy(), 0.25 * h, fill)
    b, _ = orc.rk4_step(o, g, u.copy(), 0.25 * h, between_stages=fill)
    assert np.array_equal(a, b)
    print("split RHS and step equal the whole-block oracle bit for bit", flush=True)


def run(n, t_end=T_END, dt_hst=DT_HST):
    """Oracle run on the n^3 block at spacing 1/n; returns the schedule's
    rows, the final state, the geometry and the step count."""
    h = 1.0 / n
    geom = MeshBlockGeom.cube(n, h, -0.5)
    o = orc.Params(**GAUGES["harmonic"], **COMMON)
    g = orc.Geom.block(geom.nx, geom.dx, geom.xmin)
    u = orc.adm_to_z4c(o, g, np.ascontiguousarray(gauge_wave_adm(geom, A=AMP, d=1.0)))
    orc.periodic_fill(g, u)
    orc.algconstr_range(o, g, u, 0, g.nn1 - 1, 0, g.nn2 - 1, 0, g.nn3 - 1, 1)

    def hist(u):
        adm = orc.z4c_to_adm(o, g, u)
        con = orc.adm_constraints(o, g, u, adm)
        return np.asarray(orc.history(o, g, u, con))

    sched = HistorySchedule(orc.new_dt(g, CFL), dt_hst, t_end)
    t0, dt0 = sched.initial_row()
    times, dts, rows = [t0], [dt0], [hist(u)]
    steps = 0
    for it in sched.iterations():
        u = rk4_step_split(o, g, geom, u, it["dt"], lambda c: orc.periodic_fill(g, c))
        steps += 1
        if steps % 200 == 0:
            print(f"n={n}: step {steps}, t = {it['time']:.4f}", flush=True)
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
    check_split()
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
