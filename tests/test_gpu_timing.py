"""The kernel timer of the C-ABI (z4c_set_timing / z4c_get_timing): every
pass launch of a fused RK4 step is recorded on its stream."""
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from z4c_amd import Params  # noqa: E402
from z4c_amd.initial_data import perturbed_flat  # noqa: E402
from z4c_amd.z4c import MeshBlockGeom, Z4c  # noqa: E402


def test_pass_timer_counts_fused_step():
    gs = [MeshBlockGeom.cube(8, 0.125, 0.0), MeshBlockGeom.cube(6, 0.25, 1.0)]
    z = Z4c(gs, Params(diss=0.02))
    u = z.upload([perturbed_flat(g, seed=b) for b, g in enumerate(gs)])
    work = z.make_step_work()
    z.set_timing(True)
    z.rk4_step_fused(0.01, u, work)
    z.set_timing(False)
    t = z.get_timing()
    counts = {k: v[1] for k, v in t.items()}
    # the fused step forms the curvature part inside the assembly kernel
    assert counts == {"stencil": 4, "curvature": 0, "assembly": 4, "ghosts": 1,
                      "prepare": 0, "bt_sum": 0, "algconstr": 0}
    assert all(v[0] > 0 for v in t.values() if v[1] > 0)
    # the record is cleared by reading it; nothing is recorded while off
    z.rk4_step_fused(0.01, u, work)
    assert all(v[1] == 0 for v in z.get_timing().values())


def test_stencil_configurations_agree_and_autotune():
    """Every stencil workgroup configuration gives bit-identical right-hand
    sides; the autotuner only picks among them."""
    gs = [MeshBlockGeom.cube(16, 0.0625, 0.0), MeshBlockGeom.cube(24, 0.05, 2.0),
          MeshBlockGeom.cube(12, 0.1, -2.0), MeshBlockGeom.cube(32, 0.03, 5.0)]
    p = Params(diss=0.02, damp_kappa1=0.02)
    z = Z4c(gs, p)
    u = z.upload([perturbed_flat(g, seed=3 + b) for b, g in enumerate(gs)])
    outs = []
    for th in (192, 256, 320, 384, 576, 640):
        z.set_stencil_threads(th)
        assert set(z.stencil_plan().values()) == {th}
        r = z.zeros()
        z.Z4cRHS(u, None, r)
        outs.append(r)
    plan = z.autotune(u, reps=2)
    assert set(plan.values()) <= {192, 256, 320, 384, 576, 640}
    r = z.zeros()
    z.Z4cRHS(u, None, r)
    outs.append(r)
    torch.cuda.synchronize()
    for o in outs[1:]:
        assert torch.equal(o, outs[0])


@pytest.mark.parametrize("faces", [63, 0])
def test_graph_step_replays_equal_eager_steps(faces):
    """The captured fused step, replayed several times with eager work and
    synchronisations in between, gives the eager step's state every time
    (with a Sommerfeld block: stage registers; without: in place)."""
    gs = [MeshBlockGeom.cube(8, 0.125, 0.3, faces=faces), MeshBlockGeom.cube(12, 0.1, 3.0)]
    p = Params(diss=0.02, damp_kappa1=0.02, shift_eta=2.0)
    z = Z4c(gs, p)
    host = [perturbed_flat(g, seed=40 + b) for b, g in enumerate(gs)]
    ua, ub = z.upload(host), z.upload(host)
    wa, wb = z.make_step_work(), z.make_step_work()
    z.AlgConstr(ua)
    z.AlgConstr(ub)
    for rep in range(4):
        z.rk4_step_fused(0.005, ua, wa)
        z.rk4_step_fused(0.005, ub, wb, graph=True)
        scratch = torch.ones(1000, dtype=torch.float64, device=ua.device) * rep  # eager work
        torch.cuda.synchronize()
        assert torch.equal(ua, ub), rep
        assert float(scratch.sum()) == 1000.0 * rep
    # a new dt is captured anew
    z.rk4_step_fused(0.0025, ua, wa)
    z.rk4_step_fused(0.0025, ub, wb, graph=True)
    torch.cuda.synchronize()
    assert torch.equal(ua, ub)
