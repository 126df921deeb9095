"""The C++ host (gr-athena_amd/host/z4c_host.hpp, driven by the compiled
z4c_host_demo program) against the oracle: a C++ caller of
include/z4c_mi355x.h registers two MeshBlocks with different spacings per
axis, evaluates Z4cRHS of the initial state, and runs the GR_Z4c step loop
(per stage through z4c_rk4_stage / z4c_ls_stage, or fused) with the
periodic same-level exchange after every stage.  The program runs as a
child process; its results are compared with the oracle's multi-block
stage sequence (1e-12 norm-relative for the RHS; 1e-11 with a 0.01 floor
after the steps, as the other gauge-wave evolutions)."""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

from z4c_amd import Params  # noqa: E402
from z4c_amd.build import HOST_DEMO  # noqa: E402
from z4c_amd.initial_data import gauge_wave_adm  # noqa: E402
from z4c_amd.z4c import MeshBlockGeom  # noqa: E402

NX = (8, 6, 6)
LO, HI = (-0.5, -0.25, -0.25), (0.5, 0.25, 0.25)
PARAMS = dict(lapse_oplog=0.0, lapse_harmonic=1.0, shift_Gamma=0.0, shift_eta=0.0,
              shift_advect=0.0, diss=0.02, chi_div_floor=1e-5, damp_kappa1=0.02)


def _blocks():
    """Two blocks along x: [-0.5, 0] and [0, 0.5] in x, the whole domain in
    y and z; spacing (1/16, 1/12, 1/12)."""
    out = []
    for x0, x1 in ((-0.5, 0.0), (0.0, 0.5)):
        xmin, xmax = (x0, LO[1], LO[2]), (x1, HI[1], HI[2])
        dx = tuple((xmax[d] - xmin[d]) / NX[d] for d in range(3))
        out.append((xmin, xmax, MeshBlockGeom(NX, dx, xmin)))
    return out


def _initial(op, ogs, geoms):
    us = [orc.adm_to_z4c(op, og, np.ascontiguousarray(gauge_wave_adm(g, A=0.01, d=1.0)))
          for og, g in zip(ogs, geoms)]
    orc.exchange(ogs, us, LO, HI, (1, 1, 1))
    for og, u in zip(ogs, us):
        orc.algconstr(op, og, u)
    return us


def _oracle_steps(integ, op, ogs, us, dt, nsteps):
    for _ in range(nsteps):
        if integ == "bt_rk4":
            un = [u.copy() for u in us]
            ks = [[] for _ in us]
            for s in range(4):
                for b, og in enumerate(ogs):
                    ks[b].append(orc.rhs(op, og, us[b]))
                w = orc.BT_RK4_A[s + 1][:s + 1] if s < 3 else orc.BT_RK4_B
                us = [orc.bt_sum(og, un[b], ks[b], w, dt) for b, og in enumerate(ogs)]
                orc.exchange(ogs, us, LO, HI, (1, 1, 1))
                for og, u in zip(ogs, us):
                    orc.algconstr(op, og, u)
        else:
            u1 = [np.zeros_like(u) for u in us]
            for delta, g1, g2, g3, beta in orc.LS_TABLEAUS[integ]:
                rs = [orc.rhs(op, og, u) for og, u in zip(ogs, us)]
                for b, og in enumerate(ogs):
                    orc.weighted_ave(og, u1[b], us[b], None, [1.0, delta, 0.0])
                    orc.weighted_ave(og, us[b], u1[b], None, [g1, g2, g3])
                    orc.add_rhs(og, rs[b], beta * dt, us[b])
                orc.exchange(ogs, us, LO, HI, (1, 1, 1))
                for og, u in zip(ogs, us):
                    orc.algconstr(op, og, u)
    return us


def _write_input(path, blocks, us, p, integ, dt, nsteps, fused, ghosts=2):
    with open(path, "wb") as f:
        f.write(b"Z4CH" + struct.pack("<4i", len(blocks), nsteps, ghosts, int(fused)))
        f.write(integ.encode().ljust(16, b"\0"))
        f.write(struct.pack("<d", dt) + struct.pack("<6d", *LO, *HI) + struct.pack("<4i", 1, 1, 1, 0))
        f.write(bytes(p))
        for xmin, xmax, _ in blocks:
            f.write(struct.pack("<4i", *NX, 0) + struct.pack("<6d", *xmin, *xmax) +
                    struct.pack("<6i", *([0] * 6)))
        for u in us:
            f.write(np.ascontiguousarray(u, dtype="<f8").tobytes())


def _rel(a, b, floor):
    num = np.abs(a - b).reshape(22, -1).max(axis=1)
    den = np.maximum(np.abs(b).reshape(22, -1).max(axis=1), floor)
    return num / den


@pytest.mark.parametrize("integ,fused", [("bt_rk4", False), ("rk4", True), ("bt_rk4", True),
                                         ("rk3", False)])
def test_host_program_against_oracle(tmp_path, integ, fused):
    assert os.path.exists(HOST_DEMO), "build the host program first (__graft_entry__.build())"
    blocks = _blocks()
    geoms = [g for _, _, g in blocks]
    p = Params(**PARAMS)
    op = orc.Params(**p.as_dict())
    ogs = [orc.Geom.block(g.nx, g.dx, g.xmin) for g in geoms]
    us = _initial(op, ogs, geoms)
    dt, nsteps = 0.25 / 16, 2
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(inp, blocks, us, p, integ, dt, nsteps, fused)
    print(f"running {os.path.basename(HOST_DEMO)}: {integ}, fused={fused}", flush=True)
    r = subprocess.run([HOST_DEMO, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    data = np.fromfile(out, dtype="<f8")
    n = [22 * int(np.prod(g.shape)) for g in geoms]
    assert data.size == 2 * sum(n)
    rhs_got = [data[sum(n[:b]):sum(n[:b + 1])].reshape((22,) + g.shape) for b, g in enumerate(geoms)]
    off = sum(n)
    u_got = [data[off + sum(n[:b]):off + sum(n[:b + 1])].reshape((22,) + g.shape)
             for b, g in enumerate(geoms)]
    for b, (og, u) in enumerate(zip(ogs, us)):
        exp = orc.rhs(op, og, u)
        s = (slice(None),) + og.interior()
        r_ = _rel(rhs_got[b][s], exp[s], 1e-300)
        assert np.all(r_ <= 1e-12), (b, r_)
    exp_u = _oracle_steps(integ, op, ogs, [u.copy() for u in us], dt, nsteps)
    for b in range(len(geoms)):
        r_ = _rel(u_got[b], exp_u[b], 0.01)
        print(f"block {b}: max rel diff after {nsteps} steps {r_.max():.3e}")
        assert np.all(r_ <= 1e-11), (b, r_)


def test_host_program_adjoint_equals_python_binding(tmp_path):
    """The C++ host's adjoint calls (z4c_rhs_vjp, z4c_exchange_vjp,
    z4c_rk4_step_ghosts_vjp through Z4cBatch) equal the Python binding's on
    the same blocks, state and cotangents, bit for bit: the RHS VJP with the
    right-hand side as the cotangent and its parameter gradients, the
    exchange's transpose of it, and the VJP of one fused step with the
    exchange after every stage (the state as the cotangent)."""
    import torch
    from z4c_amd.z4c import Z4c
    assert os.path.exists(HOST_DEMO), "build the host program first (__graft_entry__.build())"
    blocks = _blocks()
    geoms = [g for _, _, g in blocks]
    p = Params(**dict(PARAMS, shift_eta=1.5, damp_kappa2=0.05))
    op = orc.Params(**p.as_dict())
    ogs = [orc.Geom.block(g.nx, g.dx, g.xmin) for g in geoms]
    us = _initial(op, ogs, geoms)
    dt = 0.25 / 16
    inp, out, adj = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "adj.bin"
    _write_input(inp, blocks, us, p, "bt_rk4", dt, 0, True)
    r = subprocess.run([HOST_DEMO, str(inp), str(out), str(adj)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(adj, dtype="<f8")
    n = sum(22 * int(np.prod(g.shape)) for g in geoms)
    assert got.size == 3 * n + 2 * 3
    z = Z4c(geoms, p)
    z.set_exchange(LO, HI, (True, True, True))
    u = z.upload(us)
    rhs = z.zeros()
    z.Z4cRHS(u, None, rhs)
    z.Z4cBoundaryRHS(u, None, rhs)
    ubar, pbar = z.rhs_vjp(u, rhs, 1)
    ex = ubar.clone()
    z.exchange_vjp(ex)
    sbar, spbar = z.rk4_step_ghosts_vjp(dt, u, u, 1)
    torch.cuda.synchronize()
    host = lambda t: np.concatenate([z.download(t, 22)[b].reshape(-1) for b in range(len(geoms))])
    assert np.array_equal(got[:n], host(ubar))
    assert np.array_equal(got[n:n + 3], pbar.cpu().numpy().reshape(-1))
    assert np.array_equal(got[n + 3:2 * n + 3], host(ex))
    assert np.array_equal(got[2 * n + 3:3 * n + 3], host(sbar))
    assert np.array_equal(got[3 * n + 3:], spbar.cpu().numpy().reshape(-1))
    assert np.abs(got[:n]).max() > 0 and np.abs(got[2 * n + 3:3 * n + 3]).max() > 0
    z.close()

