#!/usr/bin/env python3
"""Throughput benchmark of the batched FP64 Z4c RK4 step on one MI355X.

Workload (the batched uneven-block configuration): 2048 vertex-centred
MeshBlocks from a seeded octree refined toward two punctures (levels 0-7,
spacing 2^-level), interior sizes 70% 16^3 / 20% 24^3 / 10% 32^3 cells,
Brill-Lindquist data (m1 = m2 = 0.5, separation 10, pre-collapsed lapse,
zero shift, K = 0), moving-puncture gauge.  A step is one classical RK4 step
(bt_rk4) on every block: per stage Z4cRHS (+KO) with SumBT and AlgConstr
fused into its assembly kernel, one launch per kernel per stage for all
blocks.  Ghost vertices stay at u_n through the step (independent blocks,
no exchange).  The reference prepares derivative arrays every stage because
its RHS reads them; here the RHS forms its own, and the arrays are prepared
only where the constraints are evaluated (not part of the step).

value = owned point-updates per second: sum over blocks of nx^3 (the
reference's zone count, src/main.hpp:729-731) per RK4 step, divided by the
wall time of the timed steps.  Inputs are resident in HBM.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gr-athena_amd"))

import numpy as np  # noqa: E402

PEAK_FP64 = 78.6e12   # FP64 vector, FMA counted as two operations
PEAK_FP64_NOFMA = 39.3e12  # the same issue rate when every operation is a separate add or mul
PEAK_HBM = 8.0e12
# kernel of each timed pass, and the committed counter summary holding the
# measured HBM bytes per launch of each kernel (FETCH_SIZE / WRITE_SIZE
# passes, calibrated against a copy kernel of known size)
PASS_KERNEL = {"stencil": "k_stencil", "assembly": "k_rk4_stage",
               "ghosts": "k_rk4_ghosts_inplace", "prepare": "k_prep"}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "z4c_rk4_step_2048blocks_pmc.md")

# FLOPs of one block of n^3 cells, counted exactly by the FLOP-counting build
# of the oracle (oracle/flops/fdouble.hpp: additions, subtractions,
# multiplications, divisions and libm calls of the restated algorithm; the
# perturbed flat state, moving-puncture gauge; tests/test_oracle_flops.py
# recounts them and pins the stencil share against the operator tally):
#   (stencil share of one RHS: every Dx, Dxx, Lx, Diss of PrepareZ4cDerivatives
#    and Z4cRHS; the rest of one RHS; one whole bt_rk4 step: 4 RHS, SumBT,
#    AlgConstr over the padded block)
FLOPS_BLOCK = {
    16: (22187346, 14724327, 157209199),
    24: (69053250, 46828191, 485519215),
    32: (157149234, 107703255, 1101589615),
}
# The oracle's C restatement runs this many times the reference's own CPU
# rate on the same host and workload (oracle/calibrate_cpu.py: 2.09 and 1.99
# over two runs in the build container, 8 threads, against the reference's
# 2.47e5 point-steps/s); the reference cannot run on the GPU box, so its
# rate there is estimated as the oracle's divided by this factor.
ORACLE_OVER_REFERENCE = 2.04

# Algorithmic HBM bytes per owned point-step (classical RK4 with the minimal
# fused 4-register scheme: 17 register passes of 22 FP64 values) and per RHS
# (read 22, write 22).
B_STEP = 17 * 22 * 8
B_RHS = 2 * 22 * 8


def pass_flops(name, blocks):
    """Algorithmic FLOPs of one launch of a pass over ``blocks``: the
    stencil pass carries the stencil share of one RHS; the stage kernel
    (curvature, assembly, stage update, AlgConstr) a quarter of the rest of
    the step; the other passes carry data movement only."""
    tot = 0
    for g in blocks:
        st, _, step = FLOPS_BLOCK[g.nx[0]]
        if name == "stencil":
            tot += st
        elif name == "assembly":
            tot += (step - 4 * st) / 4.0
    return float(tot)


def step_flops(blocks):
    return float(sum(FLOPS_BLOCK[g.nx[0]][2] for g in blocks))


def measured_traffic(kernel):
    """HBM bytes per pass launch of ``kernel`` from the committed counter
    summary (the block between 'traffic-bytes-per-launch' markers: kernel,
    read bytes, write bytes per dispatch), or None.  A pass that dispatches
    one kernel instance per workgroup configuration (the stencil pass) lists
    one line per instance; the pass's traffic is their sum."""
    try:
        text = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    if "traffic-bytes-per-launch" not in text:
        return None
    body = text.split("traffic-bytes-per-launch", 1)[1].split("-->", 1)[0]
    total, found = 0.0, False
    for line in body.splitlines():
        f = line.split()
        if len(f) == 3 and f[0] == kernel:
            total += float(f[1]) + float(f[2])
            found = True
    return total if found else None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nblocks", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-perblock", action="store_true", help="skip the per-block launch comparison")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default stencil plan (no autotuner trial launches, e.g. under a profiler)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost leg")
    ap.add_argument("--no-sensitivity", action="store_true",
                    help="skip the 100-step evolution + 10-step adjoint sensitivity leg")
    ap.add_argument("--no-exchange", action="store_true",
                    help="skip the step-with-exchange leg on a uniform periodic block grid")
    ap.add_argument("--no-multilevel", action="store_true",
                    help="skip the multi-level evolution leg (balanced octree, ghost update across "
                         "levels after every stage)")
    ap.add_argument("--multilevel-only", action="store_true",
                    help="run only the multi-level evolution leg and print its object")
    ap.add_argument("--ml-steps", type=int, default=100,
                    help="steps of the multi-level evolution leg")
    ap.add_argument("--sensitivity-only", action="store_true",
                    help="run only the sensitivity leg and print its object (for profiling)")
    ap.add_argument("--sens-steps", type=int, default=100,
                    help="steps of --sensitivity-only's evolution")
    ap.add_argument("--sens-window", type=int, default=10,
                    help="adjoint window of --sensitivity-only")
    ap.add_argument("--sens-no-fd", action="store_true",
                    help="skip --sensitivity-only's difference check (short profiled runs)")
    ap.add_argument("--sensitivity-tree", choices=("independent", "balanced"),
                    default="independent",
                    help="block set of --sensitivity-only: independent 16^3 blocks with ghosts "
                         "held at u_n, or the balanced tree with the ghost update")
    ap.add_argument("--calib-copy", action="store_true",
                    help="end with one device copy of the state (known bytes) for counter calibration")
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_workload(nblocks):
    from z4c_amd import Params
    from z4c_amd.initial_data import brill_lindquist_adm, octree_blocks, owned_points
    blocks = octree_blocks(nblocks=nblocks, seed=0)
    params = Params(lapse_oplog=2.0, lapse_harmonicf=1.0, lapse_advect=1.0, shift_Gamma=1.0,
                    shift_advect=1.0, shift_eta=2.0, diss=0.02, damp_kappa1=0.02,
                    chi_div_floor=1e-5)
    return blocks, params, owned_points(blocks), brill_lindquist_adm


def cpu_baseline(blocks, params, adm_fn, budget_s):
    """The CPU oracle (single-threaded C per block, one block per host
    thread) on a sample of the same workload: every 32nd block, one bt_rk4
    step each, cycled until about ``budget_s`` seconds of wall time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
    op = orc.Params(**params.as_dict())
    stride = max(1, len(blocks) // 64)
    sample = blocks[::stride]
    nthreads = max(1, min(16, os.cpu_count() or 1))
    states = []
    for g in sample:
        og = orc.Geom.block(g.nx, g.dx, g.xmin)
        states.append((og, orc.adm_to_z4c(op, og, np.ascontiguousarray(adm_fn(g))),
                       0.25 * min(g.dx), g.nx[0] * g.nx[1] * g.nx[2]))

    def one(i):
        og, u, dt, pts = states[i]
        orc.rk4_step(op, og, u.copy(), dt)
        return pts

    done_pts, steps, nxt = 0, 0, 0
    t_start = time.perf_counter()
    with ThreadPoolExecutor(nthreads) as ex:
        pending = set()
        while True:
            while len(pending) < nthreads and time.perf_counter() - t_start < budget_s:
                pending.add(ex.submit(one, nxt % len(states)))
                nxt += 1
            if not pending:
                break
            finished, pending = wait(pending, return_when=FIRST_COMPLETED)
            for f in finished:
                done_pts += f.result()
                steps += 1
    t_wall = time.perf_counter() - t_start
    value = done_pts / t_wall
    return {"value": value, "unit": "point-updates/s", "cores": nthreads,
            "kind": "port",
            "sample": f"{steps} block-steps over {len(states)} blocks (every {stride}th of the "
                      f"workload, one bt_rk4 step each), oracle C code, {nthreads} host "
                      f"threads, {t_wall:.1f} s",
            "calibration": {
                "oracle_over_reference": ORACLE_OVER_REFERENCE,
                "reference_equivalent_value": value / ORACLE_OVER_REFERENCE,
                "source": "oracle/calibrate_cpu.py in the build container: the oracle on the "
                          "64^3 gauge wave as 8 blocks of 32^3, 8 threads, against the "
                          "reference built from its sources on the same host and workload "
                          "(2.47e5 point-steps/s, 8 OpenMP threads)"}}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the step runs on one GPU; --gpus must be 1 (the "
                                   "block list would shard by blocks across GPUs; not run)"}))
        sys.exit(2)
    import torch
    from z4c_amd.z4c import Z4c

    blocks, params, owned, adm_fn = make_workload(args.nblocks)
    if args.sensitivity_only:
        print(json.dumps({"sensitivity": sensitivity_run(
            adm_fn, params, args.sens_steps, args.sens_window, tree=args.sensitivity_tree,
            fd_check=not args.sens_no_fd)}), flush=True)
        return
    if args.multilevel_only:
        print(json.dumps({"multilevel_evolution": multilevel_run(adm_fn, params, args.ml_steps)}),
              flush=True)
        return
    interior_pts = int(sum((g.nx[0] + 1) * (g.nx[1] + 1) * (g.nx[2] + 1) for g in blocks))
    log(f"{len(blocks)} blocks, owned points {owned}, interior points {interior_pts}")
    z = Z4c(blocks, params)
    log(f"padded points {z.total_points}")
    # initial data: Brill-Lindquist ADM -> Z4c on the device
    adm = z.empty(17)
    for b, g in enumerate(blocks):
        z.block_view(adm, b, 17).copy_(torch.from_numpy(adm_fn(g)))
    u = z.zeros()
    z.ADMToZ4c(adm, u)
    del adm
    work = z.make_step_work()
    dt = min(z.NewBlockTimeStep(0.25), 1.0)
    torch.cuda.synchronize()
    log("initial data ready")
    t_tune = time.perf_counter()
    plan = z.stencil_plan() if args.no_tune else z.autotune(u, reps=3)
    log(f"stencil workgroup sizes by block size ("
        + ("default plan" if args.no_tune else f"autotuned in {time.perf_counter() - t_tune:.2f} s")
        + "): " + ", ".join(f"{'x'.join(map(str, k))}: {v}" for k, v in sorted(plan.items())))

    def step():
        # the fused step replayed from its HIP graph (captured on first use)
        z.rk4_step_fused(dt, u, work, graph=True)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    T = time.perf_counter() - t0
    rate = owned * args.steps / T
    if not torch.isfinite(u).all():
        raise RuntimeError("non-finite state after the timed steps")
    log(f"{args.steps} steps (HIP graph replay) in {T:.3f} s -> {rate:.4e} point-updates/s")
    # per-pass kernel times: the same steps issued eagerly with HIP events
    # around every pass launch (a graph cannot hold the timer's events)
    z.set_timing(True)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        z.rk4_step_fused(dt, u, work)
    torch.cuda.synchronize()
    T_eager = time.perf_counter() - t1
    z.set_timing(False)
    passes = z.get_timing()
    log(f"{args.steps} eager steps with pass timers in {T_eager:.3f} s")
    pinfo = {}
    step_f = step_flops(blocks)
    for name, (ms, n) in passes.items():
        if n == 0:
            continue
        avg = ms / n * 1e-3
        e = {"launches": n, "avg_ms": avg * 1e3}
        fl = pass_flops(name, blocks)
        # algorithmic bytes of the pass: the stencil pass reads the 22
        # fields once per owned point; the stage kernel moves the rest of a
        # stage's share of the 2992 B per owned point-step
        bts = {"stencil": 22 * 8.0 * owned, "assembly": (B_STEP / 4.0 - 22 * 8.0) * owned}.get(name)
        if fl > 0:
            e["alg_flops"] = fl
            e["tflops"] = fl / avg / 1e12
            e["frac_fp64"] = fl / avg / PEAK_FP64
        if bts:
            e["alg_bytes"] = bts
            e["gbs"] = bts / avg / 1e9
            e["frac_hbm"] = bts / avg / PEAK_HBM
        tr = measured_traffic(PASS_KERNEL.get(name, ""))
        if tr:
            e["measured_traffic_bytes"] = tr
            e["measured_gbs"] = tr / avg / 1e9
        pinfo[name] = e
        log(f"  {name:10s} {n:4d} launches, avg {avg*1e3:8.3f} ms"
            + (f", {e['tflops']:6.2f} TF/s ({e['frac_fp64']*100:5.1f}% of FP64)" if "tflops" in e else "")
            + (f", measured {e['measured_gbs']:7.1f} GB/s" if "measured_gbs" in e else ""))

    out = {
        "metric": "FP64 Z4c point-updates/s, batched vs per-block launch; adjoint/forward cost",
        "value": rate, "unit": "point-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": T / args.steps * 1e3, "higher_is_better": True,
        "eager_ms_per_step": T_eager / args.steps * 1e3,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batched uneven AMR blocks: 2048 VC MeshBlocks (70% 16^3, "
                               "20% 24^3, 10% 32^3), levels 0-7, Brill-Lindquist binary, "
                               "bt_rk4 step replayed from a HIP graph, one launch per kernel "
                               "per stage",
                   "nblocks": len(blocks), "owned_points": owned,
                   "interior_points": interior_pts, "padded_points": z.total_points,
                   "parallelism": "single GPU",
                   "stencil_threads": {"x".join(map(str, k)): v for k, v in sorted(plan.items())}},
    }
    # roofline of the dominant kernel (largest total time): algorithmic FLOPs
    # (counted, FLOPS_BLOCK) and bytes per launch over its event-timed
    # average duration; the binding one is the larger fraction
    dom = max((k for k in pinfo if "tflops" in pinfo[k]),
              key=lambda k: pinfo[k]["avg_ms"] * pinfo[k]["launches"])
    d = pinfo[dom]
    fr_f, fr_b = d["frac_fp64"], d.get("frac_hbm", 0.0)
    ms_step = T / args.steps * 1e3
    step = {"F_flops_per_owned_point_step": step_f / owned, "B_bytes_per_owned_point_step": B_STEP,
            "tflops": step_f / (ms_step * 1e-3) / 1e12,
            "frac_fp64": step_f / (ms_step * 1e-3) / PEAK_FP64,
            "frac_hbm": B_STEP * rate / PEAK_HBM,
            "frac_fp64_nofma_ceiling": step_f / (ms_step * 1e-3) / PEAK_FP64_NOFMA}
    step["achieved_frac"] = max(step["frac_fp64"], step["frac_hbm"])
    out["roofline"] = {"bound": "fp64" if fr_f >= fr_b else "hbm",
                       "achieved": d["tflops"] if fr_f >= fr_b else d["gbs"],
                       "peak": PEAK_FP64 / 1e12 if fr_f >= fr_b else PEAK_HBM / 1e9,
                       "unit": "TFLOP/s" if fr_f >= fr_b else "GB/s",
                       "frac": max(fr_f, fr_b),
                       "traffic": d.get("measured_traffic_bytes"),
                       "kernel": PASS_KERNEL[dom], "pass": dom, "avg_ms": d["avg_ms"],
                       "alg_flops_per_launch": d["alg_flops"],
                       "alg_bytes_per_launch": d.get("alg_bytes"),
                       "frac_hbm_alg": fr_b,
                       "frac_fp64_nofma_ceiling": d["tflops"] * 1e12 / PEAK_FP64_NOFMA,
                       "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT),
                       "step": step}
    out["passes"] = pinfo

    if not args.no_adjoint:
        out["diagnostics"] = diagnostics_cost(z, u)
    if args.calib_copy:
        tmp = torch.empty_like(u)
        tmp.copy_(u)
        torch.cuda.synchronize()
        out["calib_copy_bytes"] = 2 * u.numel() * 8
        del tmp
    if not args.no_perblock:
        out["per_block_launch"] = per_block(z, blocks, params, u, dt, owned)
        out["batched_speedup_vs_per_block"] = rate / out["per_block_launch"]["value"]
    if not args.no_adjoint:
        out["adjoint"] = adjoint_cost(params)
    if not args.no_sensitivity:
        z.close()
        del u, work
        torch.cuda.empty_cache()
        out["sensitivity"] = sensitivity_run(adm_fn, params)
        torch.cuda.empty_cache()
        out["sensitivity_balanced_tree"] = sensitivity_run(adm_fn, params, tree="balanced")
    if not args.no_exchange:
        out["exchange_step"] = exchange_step_cost(params)
    if not args.no_multilevel:
        out["multilevel_evolution"] = multilevel_run(adm_fn, params, args.ml_steps)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(blocks, params, adm_fn, args.cpu_seconds)
    print(json.dumps(out), flush=True)


def diagnostics_cost(z, u, reps=3):
    """Output-cadence kernels on the same 2048-block state: Z4cToADM and the
    Weyl scalar psi4 (with its PrepareZ4cDerivatives), ADMConstraints (with
    its own prepare), timed with events on the launch stream."""
    import torch
    adm = z.zeros(17)
    con = z.zeros(7)
    w = z.empty(2)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    res = {"z4c_to_adm_ms": timed(lambda: z.Z4cToADM(u, adm)),
           "weyl_ms": timed(lambda: z.weyl(u, adm, w)),
           "adm_constraints_ms": timed(lambda: z.ADMConstraints(con, adm, None, u))}
    if not np.isfinite(z.download(w, 2)[0][:, 4:-4, 4:-4, 4:-4]).all():
        raise RuntimeError("non-finite psi4 on the workload")
    log("diagnostics: " + ", ".join(f"{k} {v:.2f}" for k, v in res.items()))
    del adm, con, w
    return res


def exchange_step_cost(params, nb=8, n=16, steps=3):
    """A step with the ghost update after every stage, on a uniform periodic
    domain of nb^3 blocks of n^3 cells (the gauge wave, A = 0.01): the fused
    step (z4c_rk4_step_ghosts, exchange) against the GR_Z4c stage sequence
    (four z4c_rk4_stage calls, exchange, with the derivative preparation of
    the task list), both timed over ``steps`` steps after one warmup."""
    import torch
    from z4c_amd.initial_data import gauge_wave_adm
    from z4c_amd.z4c import MeshBlockGeom, Z4c
    h = 1.0 / (nb * n)
    gs = [MeshBlockGeom.cube(n, h, (-0.5 + a * n * h, -0.5 + b * n * h, -0.5 + c * n * h))
          for c in range(nb) for b in range(nb) for a in range(nb)]
    z = Z4c(gs, params)
    z.set_exchange([-0.5] * 3, [0.5] * 3, (1, 1, 1))
    adm = z.zeros(17)
    for q, g in enumerate(gs):
        z.block_view(adm, q, 17).copy_(torch.from_numpy(gauge_wave_adm(g, A=0.01, d=1.0)))
    u = z.zeros()
    z.ADMToZ4c(adm, u)
    del adm
    z.exchange(u)
    dt = 0.25 * h
    work = z.make_step_work()
    un, k = z.make_rk_registers()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps * 1e3

    t_fused = timed(lambda: z.rk4_step_fused(dt, u, work, exchange=True))
    t_stages = timed(lambda: z.rk4_step(dt, u, un, k, exchange=True))
    # the VJP of the same step (one cotangent), the exchange's transpose
    # after every stage
    w = torch.randn_like(u)
    t_vjp = timed(lambda: z.rk4_step_ghosts_vjp(dt, u, w, 1))
    owned = len(gs) * n ** 3
    res = {"workload": f"gauge wave on a periodic [-0.5, 0.5]^3 split into {nb}^3 blocks of "
                       f"{n}^3 cells, exchange after every stage",
           "fused_ms_per_step": t_fused, "stage_sequence_ms_per_step": t_stages,
           "fused_point_updates_per_s": owned / t_fused * 1e3,
           "step_vjp_ms": t_vjp, "step_vjp_over_step": t_vjp / t_fused}
    log(f"exchange step: fused {t_fused:.2f} ms, stage sequence {t_stages:.2f} ms per step; "
        f"its VJP {t_vjp:.2f} ms ({t_vjp / t_fused:.1f}x)")
    z.close()
    return res


def multilevel_run(adm_fn, params, nsteps=100, nblocks=2048):
    """The Brill-Lindquist binary on a 2:1-balanced block tree of 16^3-cell
    blocks over the open box [-32, 32]^3 (levels 0-7, refined toward the
    punctures, no leaf dropped), moving-puncture gauge, CFL 0.25, outflow
    physical fill on the box faces: ``nsteps`` bt_rk4 steps replayed from a
    HIP graph with the GR_Z4c ghost update after every stage (restriction,
    the same-level / coarser / finer gather with the node multiplicity, the
    coarse level's physical fill, prolongation, the fine physical fill,
    AlgConstr of the whole block).  Also times the same step with ghosts
    held at u_n on the same blocks, for the cost of the update."""
    import torch
    from z4c_amd.initial_data import balanced_octree_blocks, owned_points
    from z4c_amd.z4c import Z4c
    blocks, lo, hi = balanced_octree_blocks(nblocks=nblocks)
    z = Z4c(blocks, params)
    t0 = time.perf_counter()
    z.set_exchange(lo, hi, (False, False, False))
    t_plan = time.perf_counter() - t0
    adm = z.empty(17)
    for b, g in enumerate(blocks):
        z.block_view(adm, b, 17).copy_(torch.from_numpy(adm_fn(g)))
    u = z.zeros()
    z.ADMToZ4c(adm, u)
    del adm
    # the ghosts of the initial data (Mesh::FinalizeZ4cADMGhosts order)
    z.exchange(u)
    z.physical_bc(u)
    z.AlgConstr(u)
    dt = min(g.dx[0] for g in blocks) * 0.25
    work = z.make_step_work()
    z.autotune(u, reps=2)
    z.rk4_step_fused(dt, u, work, graph=True, exchange=True, physical=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        z.rk4_step_fused(dt, u, work, graph=True, exchange=True, physical=True)
    torch.cuda.synchronize()
    T = time.perf_counter() - t0
    if not torch.isfinite(u).all():
        raise RuntimeError("non-finite state in the multi-level evolution")
    J = z.hamiltonian_norm(u)
    x = u.clone()
    z.rk4_step_fused(dt, x, work, graph=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(10):
        z.rk4_step_fused(dt, x, work, graph=True)
    torch.cuda.synchronize()
    T_frozen = (time.perf_counter() - t1) / 10
    owned = owned_points(blocks)
    levels = sorted({g.level for g in blocks})
    z.close()
    res = {"workload": f"Brill-Lindquist binary, {len(blocks)} blocks of 16^3 cells on a 2:1-"
                       f"balanced tree, levels {levels[0]}-{levels[-1]}, open box [-32, 32]^3 with "
                       f"outflow faces, moving-puncture gauge, CFL 0.25, {nsteps} steps (HIP "
                       f"graph), ghost update across levels after every stage",
           "nblocks": len(blocks), "owned_points": owned, "dt": dt, "steps": nsteps,
           "plan_build_s": t_plan, "ms_per_step": T / nsteps * 1e3,
           "point_updates_per_s": owned * nsteps / T,
           "frozen_ghost_ms_per_step": T_frozen * 1e3,
           "ghost_update_ms_per_step": T / nsteps * 1e3 - T_frozen * 1e3,
           "J_hamiltonian_norm": J}
    log(f"multi-level evolution: {len(blocks)} blocks, {nsteps} steps in {T:.2f} s "
        f"({T / nsteps * 1e3:.2f} ms per step, frozen ghosts {T_frozen * 1e3:.2f} ms), "
        f"plan {t_plan:.2f} s, J = {J:.6e}")
    return res


def sensitivity_run(adm_fn, params, nsteps=100, window=10, tree="independent", fd_check=True):
    """Brill-Lindquist binary, moving-puncture gauge (1+log, Gamma-driver,
    eta = 2, kappa1 = 0.02), CFL 0.25: ``nsteps`` fused steps replayed from
    a HIP graph, then J = sum dx dy dz Ht^2 and the adjoint through the last
    ``window`` steps (dJ/d eta, kappa1, kappa2 with the window's start state
    held fixed).  Also times single forward and adjoint steps and checks
    dJ/d eta and dJ/d kappa1 against central differences over the window
    (steps 1e-3 and 1e-4 of the parameter scale).

    tree "independent": 2048 blocks of 16^3 cells at levels 0-7 (octree
    refined toward the punctures, leaves dropped to reach the count), every
    block's ghosts held at u_n.  tree "balanced": the 2:1-balanced tree of
    2150 blocks of 16^3 cells (levels 0-7, open box [-32, 32]^3 with outflow
    faces) with the ghost update across levels and the outflow fill after
    every stage, forward and adjoint (z4c_rk4_step_ghosts_vjp)."""
    import torch
    from z4c_amd import Params
    from z4c_amd.initial_data import balanced_octree_blocks, octree_blocks, owned_points
    from z4c_amd.sensitivity import hamiltonian_sensitivity
    from z4c_amd.z4c import Z4c
    ghosts = tree == "balanced"
    if ghosts:
        blocks, lo, hi = balanced_octree_blocks(nblocks=2048)
    else:
        blocks = octree_blocks(nblocks=2048, seed=0, sizes=(16,), mix=(1.0,))
    z = Z4c(blocks, params)
    if ghosts:
        z.set_exchange(lo, hi, (False, False, False))
    adm = z.empty(17)
    for b, g in enumerate(blocks):
        z.block_view(adm, b, 17).copy_(torch.from_numpy(adm_fn(g)))
    u0 = z.zeros()
    z.ADMToZ4c(adm, u0)
    del adm
    if ghosts:
        # the ghosts of the initial data (Mesh::FinalizeZ4cADMGhosts order)
        z.exchange(u0)
        z.physical_bc(u0)
        z.AlgConstr(u0)
    dt = min(z.NewBlockTimeStep(0.25), 1.0)
    z.autotune(u0, reps=2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = hamiltonian_sensitivity(z, u0, dt, nsteps, window, exchange=ghosts, physical=ghosts)
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t0
    # single-step costs on the window's start state
    w = res["u_window"]
    work = z.make_step_work()
    x = w.clone()
    z.rk4_step_fused(dt, x, work, graph=True, exchange=ghosts, physical=ghosts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        z.rk4_step_fused(dt, x, work, graph=True, exchange=ghosts, physical=ghosts)
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / 5
    del x
    t0 = time.perf_counter()
    if ghosts:
        z.rk4_step_ghosts_vjp(dt, w, res["ubar"], 1, physical=True)
    else:
        z.rk4_step_vjp(dt, w, res["ubar"], 1)
    torch.cuda.synchronize()
    t_vjp = time.perf_counter() - t0
    free, total = torch.cuda.mem_get_info()

    def J_of(**kw):
        q = Params(**dict(params.as_dict(), **kw))
        saved = z.params
        z.params = q
        y = w.clone()
        for _ in range(window):
            z.rk4_step_fused(dt, y, work, exchange=ghosts, physical=ghosts)
        J = z.hamiltonian_norm(y)
        z.params = saved
        return J

    # Central differences at two small steps (agreement at both shows that
    # the comparison is limited neither by truncation, O(h^2), nor by
    # rounding, O(1/h)), and the 5-point difference at half the parameter's
    # scale: over a window of t ~ 0.02 J is close to a quartic in the
    # parameter, and where dJ/d kappa1 is ~1e-7 of J (Theta ~ 0) only a
    # wide step lifts the difference of J above its rounding.
    fd = {}
    for name in ("shift_eta", "damp_kappa1") if fd_check else ():
        x0 = getattr(params, name)
        an = res["pbar"][name]
        fd[name] = {"adjoint": an}
        for tag, rel_h in (("h1e-3", 1e-3), ("h1e-4", 1e-4)):
            h = rel_h * max(abs(x0), 1e-2)
            d = (J_of(**{name: x0 + h}) - J_of(**{name: x0 - h})) / (2 * h)
            fd[name][tag] = {"central_difference": d,
                             "rel_diff": abs(an - d) / max(abs(d), 1e-300)}
        h = 0.5 * max(abs(x0), 1e-2)
        d = (-J_of(**{name: x0 + 2 * h}) + 8 * J_of(**{name: x0 + h})
             - 8 * J_of(**{name: x0 - h}) + J_of(**{name: x0 - 2 * h})) / (12 * h)
        fd[name]["five_point_h0.5"] = {"difference": d,
                                       "rel_diff": abs(an - d) / max(abs(d), 1e-300)}
        fd[name]["rel_diff"] = fd[name]["five_point_h0.5"]["rel_diff"]
    z.close()
    form = ("2:1-balanced tree, open box [-32, 32]^3 with outflow faces, ghost update across "
            "levels and outflow fill after every stage (forward and adjoint)" if ghosts else
            "independent blocks (ghosts held at u_n)")
    out = {
        "workload": f"Brill-Lindquist binary, {len(blocks)} blocks of 16^3 cells, levels 0-7, "
                    f"{form}, moving-puncture gauge, CFL 0.25, {nsteps} steps (HIP graph), "
                    f"adjoint through the last {window}",
        "tree": tree, "nblocks": len(blocks),
        "device_memory_in_use_gb": (total - free) / 1e9,
        "owned_points": owned_points(blocks), "dt": dt, "J": res["J"], "dJ": res["pbar"],
        "seconds_total": t_total, "forward_step_ms": t_step * 1e3, "adjoint_step_ms": t_vjp * 1e3,
        "adjoint_step_over_forward_step": t_vjp / t_step, "fd_check": fd,
    }
    log(f"sensitivity ({tree} tree): {nsteps} steps + {window}-step adjoint in {t_total:.2f} s, "
        f"J = {res['J']:.6e}, dJ/deta = {res['pbar']['shift_eta']:.6e}, "
        f"dJ/dkappa1 = {res['pbar']['damp_kappa1']:.6e}; step {t_step*1e3:.2f} ms, "
        f"adjoint step {t_vjp*1e3:.2f} ms; FD rel diff "
        + ", ".join(f"{k} {v['h1e-3']['rel_diff']:.2e} (h 1e-3), {v['h1e-4']['rel_diff']:.2e} "
                    f"(h 1e-4), {v['rel_diff']:.2e} (5-point, h 0.5)" for k, v in fd.items()))
    return out


def adjoint_cost(params, nw=64, reps=5):
    """Adjoint cost on one MeshBlock of 32^3 cells (33^3 interior vertices,
    41^3 padded) holding the perturbed flat state (8 Gaussian bumps per
    variable, amplitudes 1e-6..1e-3, AlgConstr applied) with ``nw`` seeded
    Gaussian cotangents over all 22 variables, timed with events on the
    launch stream (average of ``reps`` calls after one warmup call):
      * one forward Z4cRHS of the block, and the same RHS batched over nw
        copies of the block (one launch per kernel);
      * z4c_rhs_vjp with the nw cotangents in one call;
      * one fused bt_rk4 step and z4c_rk4_step_vjp with one cotangent."""
    import torch
    from z4c_amd.initial_data import perturbed_flat
    from z4c_amd.z4c import MeshBlockGeom, Z4c
    g = MeshBlockGeom.cube(32, 1.0 / 32, (-0.5, -0.5, -0.5))
    z = Z4c([g], params)
    u = z.upload([perturbed_flat(g, seed=0)])
    z.AlgConstr(u)
    gen = torch.Generator(device=u.device).manual_seed(1)
    w = torch.randn(nw * u.numel(), dtype=torch.float64, device=u.device, generator=gen)
    rhs = z.zeros()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    t_rhs = timed(lambda: z.Z4cRHS(u, None, rhs))
    t_vjp = timed(lambda: z.rhs_vjp(u, w, nw))
    zb = Z4c([g] * nw, params)
    ub = u.repeat(nw)
    rb = zb.zeros()
    t_rhs_b = timed(lambda: zb.Z4cRHS(ub, None, rb))
    zb.close()
    del ub, rb
    dt = 0.25 / 32
    work = z.make_step_work()
    us = u.clone()
    t_step = timed(lambda: z.rk4_step_fused(dt, us, work))
    t_step_vjp = timed(lambda: z.rk4_step_vjp(dt, u, w[:u.numel()], 1))
    z.close()
    res = {
        "workload": f"one 32^3-cell block (33^3 interior vertices), perturbed flat state, "
                    f"{nw} Gaussian cotangents over 22 variables",
        "rhs_ms": t_rhs, "rhs_vjp_ms": t_vjp, "nw": nw,
        "rhs_vjp_ms_per_cotangent": t_vjp / nw,
        "vjp_per_cotangent_over_rhs": t_vjp / nw / t_rhs,
        "rhs_batched_ms": t_rhs_b,
        "vjp_over_batched_rhs": t_vjp / t_rhs_b,
        "step_ms": t_step, "step_vjp_ms": t_step_vjp,
        "step_vjp_over_step": t_step_vjp / t_step,
    }
    log(f"adjoint: rhs {t_rhs:.3f} ms, rhs_vjp x{nw} {t_vjp:.3f} ms "
        f"({res['vjp_per_cotangent_over_rhs']:.2f}x rhs per cotangent, "
        f"{res['vjp_over_batched_rhs']:.2f}x the batched rhs), step {t_step:.3f} ms, "
        f"step_vjp {t_step_vjp:.3f} ms ({res['step_vjp_over_step']:.2f}x)")
    return res


def per_block(z, blocks, params, u, dt, owned):
    """The same fused RK4 step issued block by block (one launch per kernel
    per stage per block, a single-block context each, over views of the
    batched state, one shared work buffer): the per-MeshBlock loop of the
    reference's task list on the same GPU."""
    import torch
    from z4c_amd.z4c import Z4c
    ctxs = [Z4c([g], params) for g in blocks]
    views = [u[z.offsets[b] * 22:z.offsets[b] * 22 + g.npts * 22] for b, g in enumerate(blocks)]
    work = torch.empty(3 * 22 * max(g.npts for g in blocks), dtype=torch.float64, device=u.device)

    def step():
        for c, v in zip(ctxs, views):
            c.rk4_step_fused(dt, v, work[:3 * 22 * c.total_points])
    step()
    torch.cuda.synchronize()
    n = 2
    t0 = time.perf_counter()
    for _ in range(n):
        step()
    torch.cuda.synchronize()
    T = time.perf_counter() - t0
    for c in ctxs:
        c.close()
    return {"value": owned * n / T, "unit": "point-updates/s", "steps": n,
            "ms_per_step": T / n * 1e3,
            "form": "per-block launches: 2048 single-block contexts, one fused rk4 step "
                    "(14 kernel launches) per block per step"}


if __name__ == "__main__":
    main()
