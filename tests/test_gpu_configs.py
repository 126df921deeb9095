"""The benchmark configurations at their stated sizes, on the GPU.

* 64^3 periodic gauge wave, harmonic gauge, bt_rk4 at CFL 0.25 to t = 10
  with the reference's output-driven step control: the .hst Z4c columns
  every 0.1 and the final fields against the oracle's run
  (tests/golden/gauge_wave_history.npz, written by
  tests/golden/make_gauge_wave_history.py), and the final lapse against
  the exact wave (awa_z4c.cpp:298-351) within the error bound the oracle's
  32^3 run gives under 4th-order convergence.
* One 32^3-cell block of perturbed flat space: one bt_rk4 step (the fused
  step and the GR_Z4c stage sequence) against the oracle.
* The same block: the RHS VJP with 64 Gaussian cotangents in one launch
  (each equal to its single-cotangent call bit for bit, two against
  torch.autograd of the restatement, all 64 through the dot-product
  identity <w, J v> = <J^T w, v> with J v by central differences,
  eps = 1e-6), gauge-parameter gradients (eta in [1, 2], kappa1 in
  [0, 0.05]) against central differences, and the step VJP through the
  dot-product identity.
* 2048 blocks of 16^3 cells, Brill-Lindquist binary, moving-puncture gauge:
  12 steps, then the adjoint through the last 6; dJ/d eta and dJ/d kappa1
  against central differences of J over the window.
"""
import os

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from z4c_amd import Params, awa  # noqa: E402
from z4c_amd._capi import PBAR  # noqa: E402
from z4c_amd.initial_data import (brill_lindquist_adm, octree_blocks,  # noqa: E402
                                  perturbed_flat)
from z4c_amd.z4c import MeshBlockGeom, Z4c  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MP_GAUGE = dict(lapse_oplog=2.0, lapse_harmonicf=1.0, lapse_harmonic=0.0, lapse_advect=1.0,
                shift_Gamma=1.0, shift_advect=1.0, shift_eta=2.0, diss=0.02,
                damp_kappa1=0.02, damp_kappa2=0.0, chi_div_floor=1e-5)


def _rel(got, ref, floor=1e-300):
    num = np.abs(got - ref).reshape(got.shape[0], -1).max(axis=1)
    den = np.maximum(np.abs(ref).reshape(ref.shape[0], -1).max(axis=1), floor)
    return num / den


# ---------------------------------------------------------------------------
# 64^3 gauge wave to t = 10
# ---------------------------------------------------------------------------
def test_gauge_wave_64_to_t10_history_and_fields():
    fx = np.load(os.path.join(GOLDEN, "gauge_wave_history.npz"))
    times, dts, rows, u, geom, steps = awa.run_gauge_wave(n=64, t_end=10.0, dt_hst=0.1)
    print(f"gauge wave 64^3: {steps} steps, {len(rows)} history rows", flush=True)
    # the step control is host logic: identical schedule
    assert steps == int(fx["steps_64"])
    np.testing.assert_array_equal(times, fx["times_64"])
    np.testing.assert_array_equal(dts, fx["dts_64"])
    exp = fx["rows_64"]
    # H-norm2 etc. integrate squares of constraints that are cancellation-
    # dominated remainders of O(1) terms: relative 1e-5 (the histories'
    # tolerance), above an absolute floor for the rows at truncation level
    # near t = 0 (a pointwise difference eps = 1e-12 in a constraint moves
    # a column by at most 2 eps sqrt(col V) + eps^2 V, V = 1).
    eps = 1e-12
    floor = 2 * eps * np.sqrt(np.maximum(rows, exp)) + eps * eps
    err = np.abs(rows - exp)
    ratio = err / (1e-5 * np.abs(exp) + floor)
    worst = np.max(ratio)
    r, col = np.unravel_index(np.argmax(ratio), ratio.shape)
    print(f"history: max |gpu - oracle| / (1e-5 |oracle| + floor) = {worst} at t = {times[r]:.4g}, "
          f"column {col}: gpu {rows[r, col]:.6e}, oracle {exp[r, col]:.6e}, floor {floor[r, col]:.3e}")
    assert np.all(err <= 1e-5 * np.abs(exp) + floor)
    # the H-norm2 history is a real signal: it grows from ~1e-18 to ~1e-10
    assert exp[-1, 0] > 1e6 * exp[0, 0]
    # final fields: independent of y and z to the rounding the VC periodic
    # fill leaves (it averages a vertex's periodic images, 2 on faces, 4 on
    # edges, 8 at corners, so the box's edge lines differ from the faces by
    # an ulp from t = 0 on, in the oracle as here; measured 5.0e-10 of the
    # field's scale at t = 10), and the x line against the oracle's
    line = u[:, 4, 4, :]
    yz = np.abs(u[:, 4:69, 4:69, :] - line[:, None, None, :]).max(axis=(1, 2, 3))
    yz_rel = yz / np.maximum(np.abs(line).max(axis=1), 1e-2)
    print("final fields: max per-variable y/z spread", yz_rel.max())
    assert np.all(yz_rel <= 1e-9), yz_rel
    r = _rel(line, fx["line_64"], floor=1e-2)
    print("final fields: max per-variable relative diff", r.max())
    assert np.all(r <= 1e-9), r
    # exact wave: error within the 32^3 oracle run's error / 2^4 (RK4 at
    # fixed CFL converges at 4th order), with 25% headroom
    x, _, _ = geom.coords()
    s = slice(4, 69)
    a_err = np.max(np.abs(line[18, s] - np.sqrt(1.0 - 0.01 * np.sin(2 * np.pi * (x[s] - 10.0)))))
    bound = 1.25 * float(fx["alpha_err_32"]) / 16.0
    print(f"|alpha - exact| = {a_err:.3e}, oracle 64: {float(fx['alpha_err_64']):.3e}, "
          f"bound {bound:.3e}")
    assert a_err <= bound


# ---------------------------------------------------------------------------
# 32^3 block: one RK4 step
# ---------------------------------------------------------------------------
def _block32(p, seed=0):
    g = MeshBlockGeom.cube(32, 1.0 / 32, (-0.5, -0.5, -0.5))
    og = orc.Geom.block(g.nx, g.dx, g.xmin)
    u = orc.algconstr(orc.Params(**p.as_dict()), og, np.ascontiguousarray(perturbed_flat(g, seed)))
    return g, og, u


def test_rk4_step_32cube_against_oracle():
    """bt_rk4 on the 32^3 block, moving-puncture gauge, CFL 0.25.

    Stage by stage (the GR_Z4c stage sequence): the oracle re-forms each
    stage from the GPU's own stage input, so every stage right-hand side is
    held to the RHS bar (1e-12 of the variable's largest value; measured
    4e-16) and every stage state (SumBT + AlgConstr) to the same bar
    (measured: bit-equal).  The fused step equals the stage sequence to
    1e-13 (measured: bit-equal).  The whole step against the oracle's
    independent step is held to 1e-11: the stage right-hand sides differ at
    rounding level, which now and then flips the last bit of an O(1) metric
    component of a stage state (2e-16); the next stage's second-derivative
    stencils (sum |w| ~ 6/h^2 per axis, h = 1/32) lift one such flip to
    ~4e-12 in R_ij, i.e. ~3e-14 in A~ after dt = 1/128, which is ~2e-11 of
    A~'s largest value (1.3e-3).  Measured: 1.6e-12 (A~_zz)."""
    p = Params(**MP_GAUGE)
    op = orc.Params(**p.as_dict())
    g, og, u = _block32(p)
    dt = 0.25 / 32
    z = Z4c([g], p)
    inner = (slice(None), slice(4, 37), slice(4, 37), slice(4, 37))
    # the GR_Z4c stage sequence, checked stage by stage
    us = z.upload([u])
    un, k = z.make_rk_registers()
    un.copy_(us)
    un_h = u.copy()
    ks_h = []
    for s in range(1, 5):
        u_in = z.download(us, 22)[0]
        z.rk4_stage(s, dt, us, un, k)
        torch.cuda.synchronize()
        k_got = z.download(k[s - 1], 22)[0]
        k_exp = orc.rhs(op, og, u_in)
        rk = _rel(k_got[inner], k_exp[inner])
        print(f"stage {s}: RHS max per-variable relative diff {rk.max():.3e}")
        assert np.all(rk <= 1e-12), (s, rk)
        ks_h.append(k_got)
        ws = orc.BT_RK4_A[s][:s] if s < 4 else orc.BT_RK4_B
        u_exp = orc.algconstr(op, og, orc.bt_sum(og, un_h, ks_h, ws, dt))
        ru = _rel(z.download(us, 22)[0], u_exp)
        print(f"stage {s}: state max per-variable relative diff {ru.max():.3e}")
        assert np.all(ru <= 1e-12), (s, ru)
    stages = z.download(us, 22)[0]
    # the fused step (the benchmarked form)
    ud = z.upload([u])
    z.rk4_step_fused(dt, ud, z.make_step_work())
    torch.cuda.synchronize()
    fused = z.download(ud, 22)[0]
    rf = _rel(fused, stages)
    print(f"fused vs stage sequence: max per-variable relative diff {rf.max():.3e}")
    assert np.all(rf <= 1e-13), rf
    exp, _ = orc.rk4_step(op, og, u.copy(), dt)
    for name, got in (("fused", fused), ("stages", stages)):
        r = _rel(got, exp)
        print(f"{name} vs oracle step: max per-variable relative diff {r.max():.3e}")
        assert np.all(r <= 1e-11), (name, r)
    z.close()


# ---------------------------------------------------------------------------
# 32^3 block: adjoints with 64 cotangents
# ---------------------------------------------------------------------------
ADJ = dict(MP_GAUGE, shift_eta=1.5, damp_kappa1=0.03)


def test_rhs_vjp_64_cotangents_32cube():
    import z4c_torch as zt
    p = Params(**ADJ)
    g, og, u = _block32(p, seed=2)
    nw = 64
    z = Z4c([g], p)
    ud = z.upload([u])
    reg = 22 * z.total_points
    gen = torch.Generator(device=ud.device).manual_seed(7)
    w = torch.randn(nw * reg, dtype=torch.float64, device=ud.device, generator=gen)
    ubar, pbar = z.rhs_vjp(ud, w, nw)
    torch.cuda.synchronize()
    # each cotangent of the batched launch equals its own single call
    for c in (0, 1, 31, 63):
        u1, p1 = z.rhs_vjp(ud, w[c * reg:(c + 1) * reg].clone(), 1)
        torch.cuda.synchronize()
        assert torch.equal(u1, ubar[c * reg:(c + 1) * reg]), c
        assert torch.equal(p1[0], pbar[c]), c
    # two cotangents against torch.autograd of the restatement
    d = p.as_dict()
    for c in (0, 63):
        wc = w[c * reg:(c + 1) * reg].cpu().numpy().reshape(u.shape)
        ut = torch.from_numpy(u).requires_grad_(True)
        pt = {k: torch.tensor(d[k], dtype=torch.float64, requires_grad=True) for k in PBAR}
        s = (zt.rhs(ut, og, dict(d, **pt)) * torch.from_numpy(wc)).sum()
        gr = torch.autograd.grad(s, [ut] + [pt[k] for k in PBAR])
        got = ubar[c * reg:(c + 1) * reg].cpu().numpy().reshape(u.shape)
        r = _rel(got, gr[0].numpy())
        print(f"cotangent {c}: vs autograd {r.max():.3e}")
        assert np.all(r <= 1e-10), r
        pref = np.array([x.item() for x in gr[1:]])
        assert np.allclose(pbar[c].cpu().numpy(), pref, rtol=1e-10, atol=1e-12 * np.abs(pref).max())
    # dot-product identity for all 64: <w_c, J v> = <ubar_c, v>, J v by
    # central differences of the GPU right-hand side (eps = 1e-6).  Only
    # interior rhs values exist; w is read at the interior.
    v = torch.randn(reg, dtype=torch.float64, device=ud.device, generator=gen)
    eps = 1e-6
    rp, rm = z.zeros(), z.zeros()
    z.Z4cRHS(ud + eps * v, None, rp)
    z.Z4cRHS(ud - eps * v, None, rm)
    Jv = ((rp - rm) / (2 * eps)).view(22, *g.shape)
    mask = torch.zeros_like(Jv)
    mask[:, 4:-4, 4:-4, 4:-4] = 1.0
    W = w.view(nw, 22, *g.shape)
    lhs = (W * (Jv * mask)[None]).sum(dim=(1, 2, 3, 4))
    rhs_ = (ubar.view(nw, reg) * v[None]).sum(dim=1)
    rel = ((lhs - rhs_).abs() / rhs_.abs()).cpu().numpy()
    print("dot-product identity, 64 cotangents: max rel", rel.max())
    assert np.all(rel <= 1e-6), rel
    # parameter gradients: <w_c, d rhs / d eta> and d kappa1 by central differences
    for k, name in enumerate(PBAR[:2]):
        x0 = getattr(p, name)
        h = 1e-6 * max(abs(x0), 1.0)
        outs = []
        for sgn in (1, -1):
            q = Params(**dict(p.as_dict(), **{name: x0 + sgn * h}))
            zq = Z4c([g], q)
            r_ = zq.zeros()
            zq.Z4cRHS(ud, None, r_)
            outs.append(r_)
            zq.close()
        dr = ((outs[0] - outs[1]) / (2 * h)).view(22, *g.shape) * mask
        fd = (W * dr[None]).sum(dim=(1, 2, 3, 4)).cpu().numpy()
        an = pbar[:, k].cpu().numpy()
        relp = np.abs(an - fd) / np.abs(fd)
        print(f"d/d{name}: max rel {relp.max():.3e}")
        assert np.all(relp <= 1e-6), (name, relp)
    z.close()


def test_rk4_step_vjp_32cube_dot_product():
    """Step VJP with stage checkpointing on the 32^3 block: <w, S'(u) v> by
    central differences of the fused step equals <S'(u)^T w, v> for two
    cotangents.  eps_floor is set far below zero: AlgConstr's
    regularize_nth_root switches to a constant branch at det g~ - 1 <
    eps_floor, which every projected state sits on, so finite differences
    would straddle that switch."""
    p = Params(**dict(ADJ, eps_floor=-1e-3))
    g, og, u = _block32(p, seed=3)
    z = Z4c([g], p)
    ud = z.upload([u])
    reg = 22 * z.total_points
    gen = torch.Generator(device=ud.device).manual_seed(11)
    nw = 2
    w = torch.randn(nw * reg, dtype=torch.float64, device=ud.device, generator=gen)
    v = torch.randn(reg, dtype=torch.float64, device=ud.device, generator=gen)
    dt = 0.25 / 32
    ubar, _ = z.rk4_step_vjp(dt, ud, w, nw)
    work = z.make_step_work()
    eps = 1e-6
    up, um = ud + eps * v, ud - eps * v
    z.rk4_step_fused(dt, up, work)
    z.rk4_step_fused(dt, um, work)
    Sv = (up - um) / (2 * eps)
    lhs = (w.view(nw, reg) * Sv[None]).sum(dim=1)
    rhs_ = (ubar.view(nw, reg) * v[None]).sum(dim=1)
    rel = ((lhs - rhs_).abs() / rhs_.abs()).cpu().numpy()
    print("step VJP dot-product identity: rel", rel)
    assert np.all(rel <= 1e-6), rel
    z.close()


# ---------------------------------------------------------------------------
# 2048 blocks of 16^3: shortened sensitivity run
# ---------------------------------------------------------------------------
def test_sensitivity_2048_blocks_fd():
    """eps_floor is set far below zero, as in the other finite-difference
    checks: with the default, AlgConstr's regularize_nth_root takes its
    constant branch wherever det g~ - 1 < eps_floor, so J is only piecewise
    smooth in the parameters."""
    from z4c_amd.sensitivity import hamiltonian_sensitivity
    params = Params(**dict(MP_GAUGE, eps_floor=-1e-3))
    blocks = octree_blocks(nblocks=2048, seed=0, sizes=(16,), mix=(1.0,))
    z = Z4c(blocks, params)
    adm = z.empty(17)
    for b, g in enumerate(blocks):
        z.block_view(adm, b, 17).copy_(torch.from_numpy(brill_lindquist_adm(g)))
    u0 = z.zeros()
    z.ADMToZ4c(adm, u0)
    del adm
    dt = min(z.NewBlockTimeStep(0.25), 1.0)
    nsteps, window = 12, 6
    res = hamiltonian_sensitivity(z, u0, dt, nsteps, window)
    print(f"J = {res['J']:.6e}, dJ = {res['pbar']}", flush=True)
    w0 = res["u_window"]
    work = z.make_step_work()

    def J_of(**kw):
        saved = z.params
        z.params = Params(**dict(params.as_dict(), **kw))
        y = w0.clone()
        for _ in range(window):
            z.rk4_step_fused(dt, y, work)
        J = z.hamiltonian_norm(y)
        z.params = saved
        return J

    # 5-point central differences at h = 5% of the parameter's scale: over a
    # short window dJ is ~1e-7 of J, so at small h the difference is
    # dominated by rounding (3-point differences at h = 1e-3 of the scale
    # over a 2-step window agreed to 1.1e-5 for kappa1 and 2.3e-7 for eta);
    # the stencil's truncation error is O(h^4)
    for name in ("shift_eta", "damp_kappa1"):
        x0 = getattr(params, name)
        h = 5e-2 * max(abs(x0), 1e-2)
        Jp2, Jp1 = J_of(**{name: x0 + 2 * h}), J_of(**{name: x0 + h})
        Jm1, Jm2 = J_of(**{name: x0 - h}), J_of(**{name: x0 - 2 * h})
        fd = (-Jp2 + 8.0 * Jp1 - 8.0 * Jm1 + Jm2) / (12.0 * h)
        an = res["pbar"][name]
        rel = abs(an - fd) / abs(fd)
        print(f"dJ/d{name}: adjoint {an:.9e}, central difference {fd:.9e}, rel {rel:.2e}")
        assert rel <= 1e-6
    z.close()


def test_sensitivity_balanced_tree_with_ghost_update_fd():
    """The sensitivity on 2150 blocks of 16^3 cells on a 2:1-balanced tree
    (levels 0-7, open box with outflow faces) with the ghost update across
    levels and the outflow fill after every stage, forward and adjoint
    (z4c_rk4_step_ghosts_vjp): dJ/d eta, dJ/d kappa1 and dJ/du along a
    direction at interior vertices against 5-point central differences
    over the window.  eps_floor as in the independent-block check.  The
    initial data get a constant shift and a Theta bump: from zero shift on
    the conformally flat slice (Gamma~ = 0) eta acts only at second order,
    and from Theta = 0 kappa1 barely acts, so over a short window dJ/d eta
    and dJ/d kappa1 are ~1e-12 and ~1e-8 of J, below what differences of J
    resolve."""
    from z4c_amd.initial_data import balanced_octree_blocks
    from z4c_amd.sensitivity import hamiltonian_sensitivity
    params = Params(**dict(MP_GAUGE, eps_floor=-1e-3))
    blocks, lo, hi = balanced_octree_blocks(nblocks=2048)
    z = Z4c(blocks, params)
    z.set_exchange(lo, hi, (False, False, False))
    adm = z.empty(17)
    for b, g in enumerate(blocks):
        z.block_view(adm, b, 17).copy_(torch.from_numpy(brill_lindquist_adm(g)))
    u0 = z.zeros()
    z.ADMToZ4c(adm, u0)
    del adm
    for b in range(len(blocks)):
        v = z.block_view(u0, b, 22)
        for a, sh in enumerate((0.05, -0.03, 0.02)):
            v[19 + a] += sh
        x = blocks[b].coords()
        r2 = (x[0][None, None, :] ** 2 + x[1][None, :, None] ** 2 + x[2][:, None, None] ** 2)
        v[14] += torch.from_numpy(0.05 * np.exp(-r2 / 128.0)).to(z.device)
    z.exchange(u0)
    z.physical_bc(u0)
    z.AlgConstr(u0)
    dt = 0.25 * min(g.dx[0] for g in blocks)
    nsteps, window = 12, 6
    res = hamiltonian_sensitivity(z, u0, dt, nsteps, window, exchange=True, physical=True)
    print(f"{len(blocks)} blocks: J = {res['J']:.6e}, dJ = {res['pbar']}", flush=True)
    w0 = res["u_window"]
    work = z.make_step_work()

    def J_of(u=None, **kw):
        saved = z.params
        z.params = Params(**dict(params.as_dict(), **kw))
        y = (w0 if u is None else u).clone()
        for _ in range(window):
            z.rk4_step_fused(dt, y, work, exchange=True, physical=True)
        J = z.hamiltonian_norm(y)
        z.params = saved
        return J

    # 5-point differences at half the parameter's scale: over a window of
    # t ~ 0.05 J is close to a quartic in eta and kappa1 (the stencil's
    # truncation error is O(h^4) times powers of t), while a small h would
    # leave differences of J at the level of its rounding
    for name in ("shift_eta", "damp_kappa1"):
        x0 = getattr(params, name)
        h = 0.5 * max(abs(x0), 1e-2)
        Jp2, Jp1 = J_of(**{name: x0 + 2 * h}), J_of(**{name: x0 + h})
        Jm1, Jm2 = J_of(**{name: x0 - h}), J_of(**{name: x0 - 2 * h})
        fd = (-Jp2 + 8.0 * Jp1 - 8.0 * Jm1 + Jm2) / (12.0 * h)
        an = res["pbar"][name]
        rel = abs(an - fd) / abs(fd)
        print(f"dJ/d{name}: adjoint {an:.9e}, central difference {fd:.9e}, rel {rel:.2e}",
              flush=True)
        assert rel <= 1e-6
    # a direction at interior vertices, scaled per field to the state
    gen = torch.Generator(device=z.device).manual_seed(3)
    v = torch.randn(w0.shape, generator=gen, device=z.device, dtype=torch.float64)
    for b in range(len(blocks)):
        vb = z.block_view(v, b, 22)
        m = torch.zeros_like(vb[0], dtype=torch.bool)
        m[4:-4, 4:-4, 4:-4] = True
        vb.mul_(m)
        vb.mul_(z.block_view(w0, b, 22).abs().amax(dim=(1, 2, 3), keepdim=True) + 1e-3)
    eps = 1e-6
    fd = (-J_of(w0 + 2 * eps * v) + 8.0 * J_of(w0 + eps * v) - 8.0 * J_of(w0 - eps * v)
          + J_of(w0 - 2 * eps * v)) / (12.0 * eps)
    an = float((res["ubar"] * v).sum())
    rel = abs(an - fd) / abs(fd)
    print(f"dJ/du . v: adjoint {an:.9e}, central difference {fd:.9e}, rel {rel:.2e}", flush=True)
    assert rel <= 1e-6
    z.close()


# ---------------------------------------------------------------------------
# Brill-Lindquist on a balanced block tree: ghost update across levels
# ---------------------------------------------------------------------------
def test_multilevel_octree_step_graph_equals_stage_sequence():
    """2150 blocks of 16^3 cells on a 2:1-balanced tree (levels 0-7, open box
    with outflow faces), the Brill-Lindquist binary in the moving-puncture
    gauge: one bt_rk4 step with the ghost update across levels and the
    physical fill after every stage, replayed from a HIP graph (twice, from
    the same state), equals the GR_Z4c stage sequence (z4c_rk4_stage per
    stage: RHS, SumBT, exchange + physical fill, AlgConstr, prepare) bit for
    bit; the state stays finite and every ghost is written."""
    from z4c_amd.initial_data import balanced_octree_blocks
    blocks, lo, hi = balanced_octree_blocks(nblocks=2048)
    print(f"balanced octree: {len(blocks)} blocks, levels "
          f"{sorted({g.level for g in blocks})}", flush=True)
    params = Params(**MP_GAUGE)
    z = Z4c(blocks, params)
    z.set_exchange(lo, hi, (False, False, False))
    adm = z.empty(17)
    for b, g in enumerate(blocks):
        z.block_view(adm, b, 17).copy_(torch.from_numpy(brill_lindquist_adm(g)))
    u0 = z.zeros()
    z.ADMToZ4c(adm, u0)
    del adm
    z.exchange(u0)
    z.physical_bc(u0)
    z.AlgConstr(u0)
    dt = 0.25 * min(g.dx[0] for g in blocks)
    ug = u0.clone()
    work = z.make_step_work()
    z.rk4_step_fused(dt, ug, work, graph=True, exchange=True, physical=True)
    torch.cuda.synchronize()
    ug2 = u0.clone()
    z.rk4_step_fused(dt, ug2, work, graph=True, exchange=True, physical=True)
    us = u0.clone()
    un, k = z.make_rk_registers()
    z.rk4_step(dt, us, un, k, exchange=True, physical=True)
    torch.cuda.synchronize()
    assert torch.isfinite(ug).all()
    assert torch.equal(ug, ug2)
    assert torch.equal(ug, us), float((ug - us).abs().max())
    z.close()
