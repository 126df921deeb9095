"""Low-storage (2S) integrators on the GPU (z4c_ls_stage, z4c_ls_step)
against the oracle's low-storage step (oracle.ls_step: the GR_Z4c
IntegrateZ4c sequence, src/task_list/gr/task_list_gr_z4c.cpp:170-194, over
the tableaus of src/task_list/task_list_integrators_low_storage.cpp,
pinned by tests/test_oracle_lowstorage.py).

Bars: the stage sequence and the fused step equal the oracle to 1e-12
(per-variable norm-relative), and each other bit for bit (same sums in the
same order at every vertex)."""
import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from z4c_amd import Params  # noqa: E402
from z4c_amd.initial_data import gauge_wave_adm, perturbed_flat  # noqa: E402
from z4c_amd.z4c import MeshBlockGeom, Z4c  # noqa: E402

MP_GAUGE = dict(lapse_oplog=2.0, lapse_harmonicf=1.0, lapse_advect=1.0, shift_Gamma=1.0,
                shift_advect=1.0, shift_eta=2.0, diss=0.02, damp_kappa1=0.02,
                chi_div_floor=1e-5)
NAMES = ["rk4", "rk3", "rk2", "vl2", "rk1"]


def _rel(a, b):
    num = np.abs(a - b).reshape(a.shape[0], -1).max(axis=1)
    den = np.maximum(np.abs(b).reshape(b.shape[0], -1).max(axis=1), 1e-300)
    return num / den


def _states(gs, p, seed=0):
    op = orc.Params(**p.as_dict())
    return [orc.algconstr(op, orc.Geom.block(g.nx, g.dx, g.xmin),
                          np.ascontiguousarray(perturbed_flat(g, seed=seed + b)))
            for b, g in enumerate(gs)]


@pytest.mark.parametrize("name", NAMES)
def test_ls_step_against_oracle_two_blocks(name):
    gs = [MeshBlockGeom.cube(12, 1.0 / 12, (-0.5, -0.5, -0.5)),
          MeshBlockGeom.cube((10, 8, 14), (0.07, 0.09, 0.05), (1.0, 0.0, -0.3))]
    p = Params(**MP_GAUGE)
    us = _states(gs, p, seed=21)
    dt = 0.25 * 0.05
    op = orc.Params(**p.as_dict())
    exp = [orc.ls_step(op, orc.Geom.block(g.nx, g.dx, g.xmin), u.copy(), dt, name)
           for g, u in zip(gs, us)]
    z = Z4c(gs, p)
    # the task list's stage sequence
    ua = z.upload(us)
    u1, rhs = z.zeros(), z.zeros()
    z.ls_step_stages(name, dt, ua, u1, rhs)
    # the fused step, twice on the same context (work reused)
    ub = z.upload(us)
    work = z.make_ls_work()
    z.ls_step(name, dt, ub, work)
    uc = z.upload(us)
    z.ls_step(name, dt, uc, work)
    torch.cuda.synchronize()
    assert torch.equal(ua, ub), name
    assert torch.equal(ub, uc), name
    for b, got in enumerate(z.download(ua, 22)):
        r = _rel(got, exp[b])
        print(f"{name} block {b}: max rel diff vs oracle {r.max():.3e}")
        assert np.all(r <= 1e-12), r
    z.close()


def test_ls_step_sommerfeld_faces():
    """Sommerfeld faces: the stage reads neighbouring values of its input, so
    the fused step alternates two state buffers; equal to the stage
    sequence and to the oracle with Z4cBoundaryRHS on the flagged faces."""
    faces = 0b101101  # ix1, ix2, ox2, ox3
    flags = [bool(faces >> f & 1) for f in range(6)]
    g = MeshBlockGeom.cube(10, 0.1, (-0.5, -0.5, -0.5), faces=faces)
    p = Params(**MP_GAUGE)
    (u,) = _states([g], p, seed=5)
    dt = 0.025
    op = orc.Params(**p.as_dict())
    exp = orc.ls_step(op, orc.Geom.block(g.nx, g.dx, g.xmin), u.copy(), dt, "rk4", faces=flags)
    z = Z4c([g], p)
    ua = z.upload([u])
    u1, rhs = z.zeros(), z.zeros()
    z.ls_step_stages("rk4", dt, ua, u1, rhs)
    for name, nst in (("rk4", 4), ("rk3", 3)):
        ub = z.upload([u])
        z.ls_step(name, dt, ub, z.make_ls_work())
        uc = z.upload([u])
        z.ls_step_stages(name, dt, uc, z.zeros(), z.zeros())
        torch.cuda.synchronize()
        assert torch.equal(ub, uc), name  # odd stage count ends in the work buffer
    torch.cuda.synchronize()
    r = _rel(z.download(ua, 22)[0], exp)
    print(f"Sommerfeld: max rel diff vs oracle {r.max():.3e}")
    assert np.all(r <= 1e-12), r
    z.close()


def test_ls_step_periodic_gauge_wave():
    """Periodic single block with the ghost fill after every stage, 3 rk4
    steps: fused, stage sequence and oracle (floor 0.01 on the scale, as the
    bt_rk4 gauge-wave checks)."""
    n = 12
    g = MeshBlockGeom.cube(n, 1.0 / n, (-0.5, -0.5, -0.5))
    p = Params(lapse_oplog=0.0, lapse_harmonic=1.0, shift_Gamma=0.0, shift_eta=0.0,
               shift_advect=0.0, diss=0.02, chi_div_floor=1e-5, damp_kappa1=0.02)
    op = orc.Params(**p.as_dict())
    og = orc.Geom.block(g.nx, g.dx, g.xmin)
    u = orc.adm_to_z4c(op, og, np.ascontiguousarray(gauge_wave_adm(g, A=0.01, d=1.0)))
    orc.periodic_fill(og, u)
    orc.algconstr(op, og, u)
    dt = 0.25 / n
    exp = u.copy()
    for _ in range(3):
        orc.ls_step(op, og, exp, dt, "rk4", between_stages=lambda c: orc.periodic_fill(og, c))
    z = Z4c([g], p)
    ua, ub = z.upload([u]), z.upload([u])
    u1, rhs, work = z.zeros(), z.zeros(), z.make_ls_work()
    for _ in range(3):
        z.ls_step_stages("rk4", dt, ua, u1, rhs, periodic=True)
        z.ls_step("rk4", dt, ub, work, periodic=True)
    torch.cuda.synchronize()
    assert torch.equal(ua, ub)
    got = z.download(ua, 22)[0]
    num = np.abs(got - exp).reshape(22, -1).max(axis=1)
    den = np.maximum(np.abs(exp).reshape(22, -1).max(axis=1), 0.01)
    assert np.all(num / den <= 1e-11), num / den
    z.close()


def test_ls_rejects_bad_input():
    from z4c_amd._capi import LsTableau, Z4cError
    g = MeshBlockGeom.cube(8, 0.1, 0.0)
    z = Z4c([g], Params())
    u = z.zeros()
    with pytest.raises(Z4cError):
        z.ls_tableau("ssprk5_4")
    t = z.ls_tableau("rk4")
    t.gamma_3[1] = 0.5  # a 3S* stage: the fused step has no u2
    with pytest.raises(Z4cError):
        z.ls_step(t, 0.01, u, z.make_ls_work())
    with pytest.raises(Z4cError):
        z.ls_stage("rk4", 5, 0.01, u, z.zeros(), z.zeros())
    bad = LsTableau()
    bad.nstages = 0
    with pytest.raises(Z4cError):
        z.ls_step(bad, 0.01, u, z.make_ls_work())
    z.close()
