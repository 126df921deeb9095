"""The batched uneven-block workload at its full size (2048 VC MeshBlocks:
70% 16^3, 20% 24^3, 10% 32^3 cells, levels 0-7, Brill-Lindquist binary),
through properties that do not depend on the size:

* the fused in-place bt_rk4 step equals the four-stage sequence of the
  GR_Z4c task list (z4c_rk4_stage) on every block to rounding;
* the right-hand side of the whole batch, at one block of each size class,
  equals the oracle's to the RHS tolerance (1e-12 of each variable's range);
* restriction followed by whole-interior prolongation (the AMR refinement
  form) reproduces every coincident vertex exactly.
"""
import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from z4c_amd import Params  # noqa: E402
from z4c_amd.initial_data import brill_lindquist_adm, octree_blocks  # noqa: E402
from z4c_amd.z4c import Z4c  # noqa: E402

MP_GAUGE = dict(lapse_oplog=2.0, lapse_harmonicf=1.0, lapse_advect=1.0, shift_Gamma=1.0,
                shift_advect=1.0, shift_eta=2.0, diss=0.02, damp_kappa1=0.02,
                chi_div_floor=1e-5)


@pytest.fixture(scope="module")
def workload():
    blocks = octree_blocks(nblocks=2048, seed=0)
    p = Params(**MP_GAUGE)
    z = Z4c(blocks, p)
    adm = z.zeros(17)
    for b, g in enumerate(blocks):
        z.block_view(adm, b, 17).copy_(torch.from_numpy(brill_lindquist_adm(g)))
    u = z.zeros()
    z.ADMToZ4c(adm, u)
    torch.cuda.synchronize()
    del adm
    yield blocks, p, z, u
    del z
    torch.cuda.empty_cache()


def test_fused_step_equals_stage_sequence_full_size(workload):
    blocks, p, z, u0 = workload
    dt = 0.25 * min(min(g.dx) for g in blocks)
    ua = u0.clone()
    ub = u0.clone()
    un, k = z.make_rk_registers()
    z.PrepareZ4cDerivatives(ua)
    z.rk4_step(dt, ua, un, k)
    del un, k
    work = z.make_step_work()
    z.rk4_step_fused(dt, ub, work)
    torch.cuda.synchronize()
    del work
    diff = (ua - ub).abs()
    worst = 0.0
    for b in range(0, len(blocks), 97):
        va, vb = z.block_view(ua, b, 22), z.block_view(diff, b, 22)
        den = va.abs().reshape(22, -1).amax(dim=1).clamp_min(1e-300)
        worst = max(worst, float((vb.reshape(22, -1).amax(dim=1) / den).max()))
    assert worst <= 1e-14, worst
    assert float(diff.max()) <= 1e-12 * float(ua.abs().max())


def test_rhs_against_oracle_one_block_per_size(workload):
    blocks, p, z, u = workload
    rhs = z.zeros()
    z.Z4cRHS(u, None, rhs)
    torch.cuda.synchronize()
    op = orc.Params(**p.as_dict())
    seen = set()
    for b, g in enumerate(blocks):
        if g.nx in seen:
            continue
        seen.add(g.nx)
        og = orc.Geom.block(g.nx, g.dx, g.xmin)
        ub = np.ascontiguousarray(z.block_view(u, b, 22).cpu().numpy())
        exp = orc.rhs(op, og, ub)
        got = z.block_view(rhs, b, 22).cpu().numpy()
        sl = (slice(None),) + tuple(slice(4, 5 + g.nx[d]) for d in (2, 1, 0))
        num = np.abs(got[sl] - exp[sl]).reshape(22, -1).max(axis=1)
        den = np.abs(exp[sl]).reshape(22, -1).max(axis=1)
        r = np.where(den > 0, num / np.where(den > 0, den, 1), num)
        assert np.all(r <= 1e-12), (g.nx, r)
    assert len(seen) == 3


def test_restrict_then_prolongate_interior_full_size(workload):
    blocks, p, z, u = workload
    c = z.coarse_full(22, 0.0)
    z.RestrictInterior(u, c, 22)
    boxes = []
    for b in range(len(blocks)):
        cs = z.coarse_shape(b)
        boxes.append((b, 4, cs[2] - 5, 4, cs[1] - 5, 4, cs[0] - 5))
    z.set_prolongation(boxes)
    v = u.clone()
    z.prolongate(c, v, 22)
    torch.cuda.synchronize()
    for b in range(0, len(blocks), 211):
        g = blocks[b]
        a = z.block_view(u, b, 22)[:, 4:5 + g.nx[2]:2, 4:5 + g.nx[1]:2, 4:5 + g.nx[0]:2]
        bb = z.block_view(v, b, 22)[:, 4:5 + g.nx[2]:2, 4:5 + g.nx[1]:2, 4:5 + g.nx[0]:2]
        assert torch.equal(a, bb), b
    z.set_prolongation([])
