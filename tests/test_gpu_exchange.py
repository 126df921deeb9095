"""Same-level VC ghost exchange between blocks on the device (z4c_exchange)
against the oracle's brute-force exchange, and a bt_rk4 evolution of a
periodic domain split into blocks against the same domain as one block
(src/comm/comm_channel.cpp:1277-1330, src/comm/node_multiplicity.cpp)."""
import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from z4c_amd import Params, Z4cError  # noqa: E402
from z4c_amd import awa  # noqa: E402
from z4c_amd.initial_data import gauge_wave_adm, perturbed_flat  # noqa: E402
from z4c_amd.z4c import MeshBlockGeom, Z4c  # noqa: E402


def _grid(parts, n, h, lo):
    sub = [n // p for p in parts]
    return [MeshBlockGeom.cube(tuple(sub), h, tuple(lo + np.array([a, b, c]) * np.array(sub) * h))
            for c in range(parts[2]) for b in range(parts[1]) for a in range(parts[0])]


@pytest.mark.parametrize("parts,periodic", [((2, 2, 1), (1, 1, 1)), ((2, 1, 2), (1, 0, 1)),
                                            ((1, 1, 1), (1, 1, 1))])
def test_exchange_matches_oracle(parts, periodic):
    h, n, lo = 0.125, 8, -0.5
    gs = _grid(parts, n, h, lo)
    rng = np.random.default_rng(31)
    host = [rng.standard_normal((22,) + g.shape) for g in gs]
    z = Z4c(gs, Params())
    z.set_exchange([lo] * 3, [lo + n * h] * 3, periodic)
    u = z.upload(host)
    z.exchange(u)
    torch.cuda.synchronize()
    got = z.download(u, 22)
    og = [orc.Geom.block(g.nx, g.dx, g.xmin) for g in gs]
    exp = orc.exchange(og, [np.ascontiguousarray(x) for x in host], [lo] * 3,
                       [lo + n * h] * 3, periodic)
    for b in range(len(gs)):
        assert np.array_equal(got[b], exp[b]), b


def test_exchange_rejects_mixed_spacing_off_the_block_tree():
    """Blocks of several spacings go through the level-aware plan, which
    needs equal blocks of at least 8 cells on a 2:1-balanced tree."""
    gs = [MeshBlockGeom.cube(4, 0.25, 0.0), MeshBlockGeom.cube(4, 0.125, 1.0)]
    z = Z4c(gs, Params())
    with pytest.raises(Z4cError):
        z.set_exchange([0.0] * 3, [2.0] * 3)


def test_split_gauge_wave_evolution_equals_single_block():
    """The periodic gauge wave on one 16^3 block (periodic self-fill) and on
    2 x 1 x 2 blocks (exchange) evolve to the same state: two bt_rk4 steps
    with the ghost update after every stage."""
    n, nsteps = 16, 2
    geom, params, z1, u1 = awa.gauge_wave_setup(n=n, gauge="puncture")
    dt = z1.NewBlockTimeStep(0.25)
    un1, k1 = z1.make_rk_registers()
    for _ in range(nsteps):
        z1.rk4_step(dt, u1, un1, k1, periodic=True)
    whole = z1.download(u1, 22)[0]

    gs = _grid((2, 1, 2), n, 1.0 / n, -0.5)
    z = Z4c(gs, params)
    z.set_exchange([-0.5] * 3, [0.5] * 3, (1, 1, 1))
    adm = z.upload([gauge_wave_adm(g, A=0.01, d=1.0) for g in gs])
    u = z.zeros()
    z.ADMToZ4c(adm, u)
    z.exchange(u)
    nn = gs[0].nn
    z.AlgConstr(u, rng=(0, nn[0] - 1, 0, nn[1] - 1, 0, nn[2] - 1), skip_physical=True)
    z.PrepareZ4cDerivatives(u)
    un, k = z.make_rk_registers()
    for _ in range(nsteps):
        z.rk4_step(dt, u, un, k, exchange=True)
    torch.cuda.synchronize()
    parts = z.download(u, 22)
    # against the oracle's multi-block stage sequence with its exchange
    og = [orc.Geom.block(g.nx, g.dx, g.xmin) for g in gs]
    op = orc.Params(**params.as_dict())
    lo, hi, per = [-0.5] * 3, [0.5] * 3, [1, 1, 1]
    us = [orc.adm_to_z4c(op, o, np.ascontiguousarray(gauge_wave_adm(g, A=0.01, d=1.0)))
          for o, g in zip(og, gs)]
    orc.exchange(og, us, lo, hi, per)
    for o, x in zip(og, us):
        orc.algconstr_range(op, o, x, 0, o.nn1 - 1, 0, o.nn2 - 1, 0, o.nn3 - 1, 1)
    for _ in range(nsteps):
        un_h = [x.copy() for x in us]
        ks = [[] for _ in us]
        for st in range(4):
            for b, o in enumerate(og):
                ks[b].append(orc.rhs(op, o, us[b]))
                w = orc.BT_RK4_A[st + 1][:st + 1] if st < 3 else orc.BT_RK4_B
                us[b] = orc.bt_sum(o, un_h[b], ks[b], w, dt)
            orc.exchange(og, us, lo, hi, per)
            for o, x in zip(og, us):
                orc.algconstr(op, o, x)
    for b in range(len(gs)):
        num = np.abs(parts[b] - us[b]).reshape(22, -1).max(axis=1)
        den = np.maximum(np.abs(us[b]).reshape(22, -1).max(axis=1), 0.01)
        assert np.all(num / den <= 1e-11), (b, num / den)
    # and, physically, the same state as the single block: the two differ
    # only in the summation order of the 4- and 8-fold shared vertices,
    # which the second-derivative operators amplify in Theta and A~
    sub = n // 2
    for b, g in enumerate(gs):
        a, c = b % 2, b // 2
        i0, k0 = a * sub, c * sub
        ref = whole[:, k0:k0 + g.shape[0], :, i0:i0 + g.shape[2]]
        num = np.abs(parts[b] - ref).reshape(22, -1).max(axis=1)
        den = np.maximum(np.abs(ref).reshape(22, -1).max(axis=1), 0.01)
        assert np.all(num / den <= 1e-10), (b, num / den)


@pytest.mark.parametrize("faces,outflow", [(63, 0), (2 | 16, 0), (0, 1 | 4 | 8), (1, 32)])
def test_physical_fill_matches_oracle(faces, outflow):
    gs = [MeshBlockGeom.cube((6, 7, 5), 0.1, (0.2, -0.1, 0.3), faces=faces, outflow=outflow),
          MeshBlockGeom.cube(4, 0.2, (3.0, 3.0, 3.0)),
          MeshBlockGeom.cube(5, 0.1, (-2.0, 1.0, 0.5), faces=faces)]
    rng = np.random.default_rng(41)
    host = [rng.standard_normal((22,) + g.shape) for g in gs]
    z = Z4c(gs, Params())
    u = z.upload(host)
    z.physical_bc(u)
    torch.cuda.synchronize()
    got = z.download(u, 22)
    for b, g in enumerate(gs):
        bits = g.sommerfeld_faces | g.outflow_faces
        fl = [(bits >> f) & 1 for f in range(6)]
        exp = orc.physical_bc(orc.Geom.block(g.nx, g.dx, g.xmin), np.ascontiguousarray(host[b]), fl)
        assert np.array_equal(got[b], exp), b


def test_sommerfeld_block_evolution_with_physical_fill():
    """One block with gr_sommerfeld on every face: bt_rk4 steps with the
    physical ghost fill after every stage, against the oracle's steps."""
    g = MeshBlockGeom.cube(12, 1.0 / 12, (0.3, -0.45, 0.2), faces=63)
    p = Params(lapse_oplog=2.0, shift_Gamma=1.0, shift_eta=2.0, diss=0.02, damp_kappa1=0.02,
               chi_div_floor=1e-5)
    # smooth data (the gauge wave): the 5-point extrapolation into the ghosts
    # keeps it smooth, so two steps stay well conditioned
    op = orc.Params(**p.as_dict())
    og = orc.Geom.block(g.nx, g.dx, g.xmin)
    u0 = orc.adm_to_z4c(op, og, np.ascontiguousarray(gauge_wave_adm(g, A=0.01, d=1.0)))
    z = Z4c([g], p)
    u = z.upload([u0])
    un, k = z.make_rk_registers()
    dt = 0.25 / 12
    exp = u0.copy()
    for _ in range(2):
        z.rk4_step(dt, u, un, k, physical=True)
        # the oracle's stage sequence with the Sommerfeld face RHS
        # (Z4cBoundaryRHS after Z4cRHS) and the physical fill after SumBT
        n0, ks, cur = exp.copy(), [], exp
        for st in range(4):
            kk = orc.rhs(op, og, cur)
            orc.boundary_rhs(op, og, cur, [1] * 6, kk)
            ks.append(kk)
            w = orc.BT_RK4_A[st + 1][:st + 1] if st < 3 else orc.BT_RK4_B
            cur = orc.bt_sum(og, n0, ks, w, dt)
            orc.physical_bc(og, cur, [1] * 6)
            orc.algconstr(op, og, cur)
        exp = cur
    torch.cuda.synchronize()
    got = z.download(u, 22)[0]
    # Interior vertices: the RHS, the face RHS and the update.  1e-10: the
    # vertices next to the faces read ghosts that the fill extrapolated from
    # the previous stage, so rounding-level differences of the RHS (operation
    # order of the KO and advective sums) come back amplified by the
    # extrapolation weights (up to 10 per layer).
    sl = (slice(None),) + (slice(og.il, og.iu + 1),) * 3
    num = np.abs(got[sl] - exp[sl]).reshape(22, -1).max(axis=1)
    den = np.maximum(np.abs(exp[sl]).reshape(22, -1).max(axis=1), 0.01)
    assert np.all(num / den <= 1e-10), num / den
    # Ghosts: the fill reads the outermost ghost layer (iu+4), which it never
    # writes, and repeated 5-point extrapolation (coefficient sum 31 per layer)
    # grows the far ghost corners to O(1e2) within a step; rounding
    # differences grow with them, so the ghosts are compared at 1e-8 of the
    # field's largest value.
    num = np.abs(got - exp).reshape(22, -1).max(axis=1)
    den = np.maximum(np.abs(exp).reshape(22, -1).max(axis=1), 0.01)
    assert np.all(num / den <= 1e-8), num / den
    # and the interior stays at the gauge wave's amplitude
    assert np.abs(got[sl][8:14]).max() < 0.1


def _fused_vs_stages(gs, params, host, dt, nsteps, setup=None, **flags):
    """The fused step with a ghost update against rk4_step (four
    z4c_rk4_stage calls) with the same update."""
    z = Z4c(gs, params)
    if setup:
        setup(z)
    ua, ub = z.upload(host), z.upload(host)
    un, k = z.make_rk_registers()
    work = z.make_step_work()
    z.PrepareZ4cDerivatives(ua)
    for _ in range(nsteps):
        z.rk4_step(dt, ua, un, k, **flags)
        z.rk4_step_fused(dt, ub, work, **flags)
    torch.cuda.synchronize()
    A, Bf = z.download(ua, 22), z.download(ub, 22)
    for b in range(len(gs)):
        num = np.abs(Bf[b] - A[b]).reshape(22, -1).max(axis=1)
        den = np.maximum(np.abs(A[b]).reshape(22, -1).max(axis=1), 1e-300)
        assert np.all(num / den <= 1e-14), (b, flags, num / den)
    return A


def test_fused_step_with_exchange_equals_stage_sequence():
    """Split periodic gauge wave (2 x 1 x 2 blocks): the fused step with the
    exchange after every stage (stages in place) against the stage
    sequence, two steps."""
    n = 16
    gs = _grid((2, 1, 2), n, 1.0 / n, -0.5)
    params = Params(lapse_oplog=2.0, shift_Gamma=1.0, shift_eta=2.0, diss=0.02,
                    damp_kappa1=0.02)
    op = orc.Params(**params.as_dict())
    host = [orc.adm_to_z4c(op, orc.Geom.block(g.nx, g.dx, g.xmin),
                           np.ascontiguousarray(gauge_wave_adm(g, A=0.05, d=1.0))) for g in gs]
    _fused_vs_stages(gs, params, host, 0.25 / n, 2,
                     setup=lambda z: z.set_exchange([-0.5] * 3, [0.5] * 3, (1, 1, 1)),
                     exchange=True)


def test_fused_step_with_periodic_fill_equals_stage_sequence():
    n = 12
    g = MeshBlockGeom.cube(n, 1.0 / n, -0.5)
    params = Params(lapse_oplog=2.0, shift_Gamma=1.0, shift_eta=2.0, diss=0.02)
    op = orc.Params(**params.as_dict())
    u0 = orc.adm_to_z4c(op, orc.Geom.block(g.nx, g.dx, g.xmin),
                        np.ascontiguousarray(gauge_wave_adm(g, A=0.05, d=1.0)))
    orc.periodic_fill(orc.Geom.block(g.nx, g.dx, g.xmin), u0)
    _fused_vs_stages([g], params, [u0], 0.25 / n, 2, periodic=True)


@pytest.mark.parametrize("faces,outflow", [(63, 0), (0, 63)])
def test_fused_step_with_physical_fill_equals_stage_sequence(faces, outflow):
    """Physical ghost fill after every stage: with Sommerfeld faces (stage
    states in the work registers) and with outflow faces only (in place)."""
    g = MeshBlockGeom.cube(12, 1.0 / 12, (0.3, -0.45, 0.2), faces=faces, outflow=outflow)
    p = Params(lapse_oplog=2.0, shift_Gamma=1.0, shift_eta=2.0, diss=0.02, damp_kappa1=0.02,
               chi_div_floor=1e-5)
    op = orc.Params(**p.as_dict())
    u0 = orc.algconstr(op, orc.Geom.block(g.nx, g.dx, g.xmin),
                       np.ascontiguousarray(perturbed_flat(g, seed=12, amp=(1e-3, 1e-2))))
    _fused_vs_stages([g], p, [u0], 0.25 / 12, 2, physical=True)
