"""Ghost update across refinement levels on the device (z4c_set_exchange on
blocks of several spacings, then z4c_exchange: restriction, the gather of
same-level, coarser and finer contributions with the node multiplicity,
prolongation) against the oracle's restatement of the reference's
SEND/RECV/SETB/PROLONG sequence (oracle/ml_exchange.py), and one bt_rk4
step of a two-level periodic mesh with that update after every stage
against the oracle's stage sequence."""
import numpy as np
import pytest

import ml_exchange as mlx
import oracle as orc
from ml_meshes import open_faces, refined_blocks

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from z4c_amd import Params, Z4cError  # noqa: E402
from z4c_amd.initial_data import gauge_wave_adm  # noqa: E402
from z4c_amd.z4c import MeshBlockGeom, Z4c  # noqa: E402

CASES = {
    "periodic_corner": dict(nrb=4, lo=-2.0, hi=2.0, periodic=(True,) * 3, levels=2,
                            refine=lambda c, L: bool(np.all(c < -1.0))),
    "open_centre": dict(nrb=4, lo=-2.0, hi=2.0, periodic=(False,) * 3, levels=2,
                        refine=lambda c, L: bool(np.all(np.abs(c) < 1.0))),
    "open_3level": dict(nrb=4, lo=-2.0, hi=2.0, periodic=(False,) * 3, levels=3,
                        refine=lambda c, L: bool(np.all(np.abs(c) < (1.0 if L == 0 else 0.5)))),
    "open_corner": dict(nrb=4, lo=-2.0, hi=2.0, periodic=(False,) * 3, levels=2,
                        refine=lambda c, L: bool(np.all(c < -1.0))),
    "periodic_2level": dict(nrb=2, lo=-0.5, hi=0.5, periodic=(True,) * 3, levels=2,
                            refine=lambda c, L: bool(np.all(c < 0.0))),
}


def _blocks(case, nx=8):
    c = CASES[case]
    spec = refined_blocks(c["nrb"], nx, [c["lo"]] * 3, [c["hi"]] * 3, c["refine"], c["levels"],
                          c["periodic"])
    geoms = [MeshBlockGeom(tuple(n), tuple(dx), tuple(x0)) for n, dx, x0 in spec]
    mesh = mlx.Mesh([mlx.Block(*s) for s in spec], [c["lo"]] * 3, [c["hi"]] * 3, c["periodic"])
    return c, geoms, mesh


@pytest.mark.parametrize("case", ["periodic_corner", "open_centre", "open_3level", "open_corner"])
def test_ml_exchange_matches_oracle(case):
    """All fields of all blocks bit-equal to the oracle after one update and
    after a second; on open domains every face on the domain boundary
    carries the outflow fill, so blocks that prolongate there first fill
    their coarse ghosts (open_corner)."""
    c, geoms, mesh = _blocks(case)
    faces = [open_faces(mesh, q) for q in range(len(geoms))]
    for g, f in zip(geoms, faces):
        g.outflow_faces = sum(1 << i for i in range(6) if f[i])
    print(f"{case}: {len(geoms)} blocks, levels {sorted(set(mesh.level))}", flush=True)
    rng = np.random.default_rng(5)
    host = [np.ascontiguousarray(rng.standard_normal((22,) + g.shape)) for g in geoms]
    z = Z4c(geoms, Params())
    z.set_exchange([c["lo"]] * 3, [c["hi"]] * 3, c["periodic"])
    u = z.upload(host)
    z.exchange(u)
    torch.cuda.synchronize()
    got = z.download(u, 22)
    exp = [h.copy() for h in host]
    mesh.ghost_fill(exp, faces=faces)
    for b in range(len(geoms)):
        if not np.array_equal(got[b], exp[b]):
            d = np.argwhere(got[b] != exp[b])
            raise AssertionError(f"block {b} (level {mesh.level[b]}): {len(d)} vertices differ, "
                                 f"first {d[:4].tolist()}, "
                                 f"max {np.abs(got[b] - exp[b]).max():.3e}")
    # a second update from the updated state gives the oracle's again
    z.exchange(u)
    torch.cuda.synchronize()
    got = z.download(u, 22)
    mesh.ghost_fill(exp, faces=faces)
    for b in range(len(geoms)):
        assert np.array_equal(got[b], exp[b]), b
    z.close()


def test_ml_exchange_three_fields_12cubed():
    """The update of a 3-field buffer (the gather's any-field path) on 3
    levels of 12^3 blocks, twice, bit-equal to the oracle."""
    c, geoms, mesh = _blocks("open_3level", nx=12)
    faces = [open_faces(mesh, q) for q in range(len(geoms))]
    for g, f in zip(geoms, faces):
        g.outflow_faces = sum(1 << i for i in range(6) if f[i])
    rng = np.random.default_rng(9)
    host = [np.ascontiguousarray(rng.standard_normal((3,) + g.shape)) for g in geoms]
    z = Z4c(geoms, Params())
    z.set_exchange([c["lo"]] * 3, [c["hi"]] * 3, c["periodic"])
    u = z.upload(host, 3)
    exp = [h.copy() for h in host]
    for _ in range(2):
        z.exchange(u, 3)
        torch.cuda.synchronize()
        got = z.download(u, 3)
        mesh.ghost_fill(exp, faces=faces)
        for b in range(len(geoms)):
            assert np.array_equal(got[b], exp[b]), b
    z.close()


def test_ml_exchange_rejects_unbalanced_and_unfilled_open_faces():
    # a level-2 block next to a level-0 block
    gs = [MeshBlockGeom.cube(8, 1.0 / 8, (0.0, 0.0, 0.0)),
          MeshBlockGeom.cube(8, 1.0 / 32, (1.0, 0.0, 0.0))]
    z = Z4c(gs, Params())
    with pytest.raises(Z4cError):
        z.set_exchange([0.0] * 3, [2.0] * 3, (True,) * 3)
    z.close()
    # refine a block at the open boundary: its children prolongate there,
    # and no face carries a physical fill
    spec = refined_blocks(4, 8, [-2.0] * 3, [2.0] * 3, lambda c_, L: bool(np.all(c_ < -1.0)), 2,
                          (False,) * 3)
    z = Z4c([MeshBlockGeom(tuple(n), tuple(dx), tuple(x0)) for n, dx, x0 in spec], Params())
    with pytest.raises(Z4cError, match="open face"):
        z.set_exchange([-2.0] * 3, [2.0] * 3, (False,) * 3)
    z.close()


# ---------------------------------------------------------------------------
# one bt_rk4 step on a two-level periodic mesh
# ---------------------------------------------------------------------------
MP_GAUGE = dict(lapse_oplog=2.0, lapse_harmonicf=1.0, lapse_harmonic=0.0, lapse_advect=1.0,
                shift_Gamma=1.0, shift_advect=1.0, shift_eta=2.0, diss=0.02,
                damp_kappa1=0.02, chi_div_floor=1e-5)


def _initial_state(mesh, geoms, o):
    """Gauge wave plus small periodic perturbations of every variable, the
    ghost update across levels, then AlgConstr (host, oracle)."""
    us = []
    for g in geoms:
        og = orc.Geom.block(g.nx, g.dx, g.xmin)
        u = orc.adm_to_z4c(o, og, np.ascontiguousarray(gauge_wave_adm(g, A=0.01, d=1.0)))
        x, y, zc = np.meshgrid(*g.coords()[::-1], indexing="ij")[::-1]
        for v in range(22):
            u[v] += 1e-4 * np.sin(2 * np.pi * (x + 2 * y + 3 * zc) + 0.37 * v)
        us.append(np.ascontiguousarray(u))
    mesh.ghost_fill(us)
    for g, u in zip(geoms, us):
        orc.algconstr(o, orc.Geom.block(g.nx, g.dx, g.xmin), u)
    return us


def _oracle_step(mesh, geoms, o, us, dt):
    """GR_Z4c stage sequence over all blocks: RHS, SumBT, the ghost update
    across levels, AlgConstr."""
    ogs = [orc.Geom.block(g.nx, g.dx, g.xmin) for g in geoms]
    un = [u.copy() for u in us]
    ks = [[] for _ in us]
    cur = us
    for s in range(4):
        for b, og in enumerate(ogs):
            ks[b].append(orc.rhs(o, og, cur[b]))
        ws = orc.BT_RK4_A[s + 1][:s + 1] if s < 3 else orc.BT_RK4_B
        cur = [orc.bt_sum(og, un[b], ks[b], ws, dt) for b, og in enumerate(ogs)]
        mesh.ghost_fill(cur)
        for b, og in enumerate(ogs):
            orc.algconstr(o, og, cur[b])
    return cur


def test_ml_rk4_step_against_oracle():
    """Stage by stage: the oracle re-forms each GR_Z4c stage (RHS, SumBT, the
    ghost update across levels, AlgConstr) from the GPU's own stage input,
    and every stage right-hand side and stage state must agree to 1e-12 of
    the variable's largest value.  The fused step with the exchange equals
    the stage sequence to 1e-13, and the whole step the oracle's independent
    step to 1e-11 (stage right-hand sides that differ at rounding level flip
    last bits of O(1) metric components, which the next stage's second
    derivatives lift; see test_gpu_configs.test_rk4_step_32cube_against_oracle)."""
    c, geoms, mesh = _blocks("periodic_2level")
    p = Params(**MP_GAUGE)
    o = orc.Params(**p.as_dict())
    ogs = [orc.Geom.block(g.nx, g.dx, g.xmin) for g in geoms]
    print(f"{len(geoms)} blocks, levels {sorted(set(mesh.level))}", flush=True)
    us = _initial_state(mesh, geoms, o)
    dt = 0.25 * min(g.dx[0] for g in geoms)
    z = Z4c(geoms, p)
    z.set_exchange([c["lo"]] * 3, [c["hi"]] * 3, c["periodic"])

    def rel(got, exp):
        scale = np.maximum(np.max([np.abs(e).reshape(22, -1).max(1) for e in exp], axis=0),
                           1e-300)
        err = np.max([np.abs(g_ - e).reshape(22, -1).max(1) for g_, e in zip(got, exp)], axis=0)
        return err / scale

    ua = z.upload(us)
    un, k = z.make_rk_registers()
    un.copy_(ua)
    ks = [[] for _ in geoms]
    for s_ in range(1, 5):
        u_in = z.download(ua, 22)
        z.rk4_stage(s_, dt, ua, un, k, exchange=True)
        torch.cuda.synchronize()
        k_got = z.download(k[s_ - 1], 22)
        k_exp = [orc.rhs(o, og, x) for og, x in zip(ogs, u_in)]
        rk = rel(k_got, k_exp)
        print(f"stage {s_}: RHS max per-variable relative diff {rk.max():.3e}", flush=True)
        assert np.all(rk <= 1e-12), (s_, rk)
        ws = orc.BT_RK4_A[s_][:s_] if s_ < 4 else orc.BT_RK4_B
        for b in range(len(geoms)):
            ks[b].append(k_got[b])
        nxt = [orc.bt_sum(og, us[b], ks[b], ws, dt) for b, og in enumerate(ogs)]
        mesh.ghost_fill(nxt)
        for og, x in zip(ogs, nxt):
            orc.algconstr(o, og, x)
        ru = rel(z.download(ua, 22), nxt)
        print(f"stage {s_}: state max per-variable relative diff {ru.max():.3e}", flush=True)
        assert np.all(ru <= 1e-12), (s_, ru)
    stages = z.download(ua, 22)
    ub = z.upload(us)
    z.rk4_step_fused(dt, ub, z.make_step_work(), exchange=True)
    torch.cuda.synchronize()
    fused = z.download(ub, 22)
    rf = rel(fused, stages)
    print(f"fused vs stage sequence: max per-variable relative diff {rf.max():.3e}", flush=True)
    assert np.all(rf <= 1e-13), rf
    exp = _oracle_step(mesh, geoms, o, [u.copy() for u in us], dt)
    r = rel(stages, exp)
    print(f"stage sequence vs oracle step: max per-variable relative diff {r.max():.3e}",
          flush=True)
    assert np.all(r <= 1e-11), r
    z.close()


@pytest.mark.parametrize("kind", ["exchange_two_levels", "periodic_fill"])
def test_graph_replayed_step_with_ghost_update(kind):
    """z4c_rk4_step_ghosts_graph equals z4c_rk4_step_ghosts bit for bit on
    four replays in a row, with eager work and synchronisation between
    them: the two-level exchange (restriction, gather, prolongation) and the
    single-block periodic fill."""
    p = Params(**MP_GAUGE)
    o = orc.Params(**p.as_dict())
    if kind == "exchange_two_levels":
        c, geoms, mesh = _blocks("periodic_2level")
        us = _initial_state(mesh, geoms, o)
        flags = dict(exchange=True)
    else:
        geoms = [MeshBlockGeom.cube(16, 1.0 / 16, -0.5)]
        og = orc.Geom.block(geoms[0].nx, geoms[0].dx, geoms[0].xmin)
        u0 = orc.adm_to_z4c(o, og, np.ascontiguousarray(gauge_wave_adm(geoms[0], A=0.01, d=1.0)))
        orc.periodic_fill(og, u0)
        us = [orc.algconstr(o, og, u0)]
        flags = dict(periodic=True)
    dt = 0.25 * min(g.dx[0] for g in geoms)
    z = Z4c(geoms, p)
    if kind == "exchange_two_levels":
        z.set_exchange([c["lo"]] * 3, [c["hi"]] * 3, c["periodic"])
    ug, ue = z.upload(us), z.upload(us)
    wg, we = z.make_step_work(), z.make_step_work()
    for rep in range(4):
        z.rk4_step_fused(dt, ug, wg, graph=True, **flags)
        z.rk4_step_fused(dt, ue, we, **flags)
        torch.cuda.synchronize()
        junk = torch.randn(1 << 20, device=ug.device).sum()  # eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        assert torch.equal(ug, ue), rep
    z.close()
