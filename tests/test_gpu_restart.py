"""Checkpoint and resume through a restart file (src/outputs/restart.cpp):
a run that writes the device state to an .rst file after one fused bt_rk4
step, reads it back into a fresh context and steps on equals the run that
never stopped, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from z4c_amd import Params  # noqa: E402
from z4c_amd import restart as rst  # noqa: E402
from z4c_amd.initial_data import perturbed_flat  # noqa: E402
from z4c_amd.z4c import MeshBlockGeom, Z4c  # noqa: E402

SIZE = dict(x1min=-1.0, x2min=-0.5, x3min=-0.5, x1max=1.0, x2max=0.5, x3max=0.5,
            x1rat=1.0, x2rat=1.0, x3rat=1.0, nx1=32, nx2=16, nx3=16)


def test_resume_from_restart_equals_uninterrupted_run(tmp_path):
    p = Params(lapse_oplog=2.0, shift_eta=2.0, damp_kappa1=0.02, diss=0.02,
               chi_div_floor=1e-5)
    h = 1.0 / 16
    gs = [MeshBlockGeom.cube(16, h, (x, -0.5, -0.5)) for x in (-1.0, 0.0)]
    dt = 0.25 * h
    z = Z4c(gs, p)
    u = z.upload([perturbed_flat(g, seed=b) for b, g in enumerate(gs)])
    work = z.make_step_work()
    z.rk4_step_fused(dt, u, work)
    torch.cuda.synchronize()
    params = rst.parameter_dump([("mesh", [("nx1", "32", ""), ("nx2", "16", ""), ("nx3", "16", "")]),
                                 ("meshblock", [("nx1", "16", ""), ("nx2", "16", ""),
                                                ("nx3", "16", "")])])
    r = rst.from_device(z, u, params, SIZE, dt, dt, 1, [(0, 0, 0, 0), (1, 0, 0, 0)])
    path = tmp_path / "z4c.00001.rst"
    rst.write_restart(path, r)
    z.rk4_step_fused(dt, u, work)
    torch.cuda.synchronize()
    a = z.download(u, 22)

    q = rst.read_restart(path)
    assert q.ncycle == 1 and q.time == dt and q.locs == r.locs
    z2 = Z4c(gs, p)
    u2 = rst.to_device(z2, q)
    z2.rk4_step_fused(q.dt, u2, z2.make_step_work())
    torch.cuda.synchronize()
    b = z2.download(u2, 22)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
