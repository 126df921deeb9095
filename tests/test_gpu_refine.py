"""VC restriction and prolongation between levels on the device
(z4c_restrict / z4c_prolongate, src/mesh/mesh_refinement_vc.cpp) against the
oracle bit for bit, the neighbour ranges against ProlongationRange
(src/comm/refinement_ops.hpp:48-72), and at a 16^3 block the level transfer
of a quintic (exact for the 6-point Lagrange weights) on all 26 ghost boxes."""
import itertools

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from z4c_amd import Params, Z4cError  # noqa: E402
from z4c_amd.z4c import MeshBlockGeom, Z4c  # noqa: E402

OFFSETS = [o for o in itertools.product((-1, 0, 1), repeat=3) if o != (0, 0, 0)]


def _blocks():
    return [MeshBlockGeom((8, 8, 8), (0.1, 0.1, 0.1), (0.0, 0.0, 0.0)),
            MeshBlockGeom((12, 8, 10), (0.05, 0.07, 0.09), (0.3, -0.2, 0.1)),
            MeshBlockGeom((6, 10, 8), (0.2, 0.1, 0.15), (-1.0, 0.5, 0.2))]


def _og(g):
    return orc.Geom.block(g.nx, g.dx, g.xmin)


def _coarse_host(z, nvar, rng):
    return [rng.standard_normal((nvar,) + z.coarse_shape(b)) for b in range(len(z.blocks))]


def _upload_coarse(z, arrays):
    offs, total = z.coarse_layout()
    nvar = arrays[0].shape[0]
    host = np.concatenate([a.reshape(-1) for a in arrays])
    assert host.size == nvar * total
    assert offs[1] == int(np.prod(z.coarse_shape(0)))
    return torch.from_numpy(host).to(z.device)


@pytest.mark.parametrize("nvar", [22, 3])
def test_restriction_matches_oracle(nvar):
    gs = _blocks()
    z = Z4c(gs, Params())
    rng = np.random.default_rng(5)
    host = [rng.standard_normal((nvar,) + g.shape) for g in gs]
    u = z.upload(host)
    c = z.coarse_full(nvar, float("nan"))
    z.RestrictInterior(u, c, nvar)
    torch.cuda.synchronize()
    for b, g in enumerate(gs):
        got = z.coarse_view(c, b, nvar).cpu().numpy()
        exp = orc.restrict_vc(_og(g), np.ascontiguousarray(host[b]))
        assert np.array_equal(np.isnan(got), np.isnan(exp)), b
        w = ~np.isnan(exp)
        assert np.array_equal(got[w], exp[w]), b


def test_prolongation_range_matches_oracle():
    gs = _blocks()
    z = Z4c(gs, Params())
    for b, g in enumerate(gs):
        for ox in OFFSETS:
            for lx in ((0, 0, 0), (1, 1, 1), (2, 3, 5)):
                got = z.ProlongationRange(b, ox, lx)
                exp = orc.prolongation_range(_og(g), ox, [v & 1 for v in lx])
                assert got == (b,) + tuple(exp), (b, ox, lx)


@pytest.mark.parametrize("parity", [(0, 0, 0), (1, 0, 1)])
def test_neighbour_prolongation_matches_oracle(parity):
    """All 26 coarser-neighbour boxes of every block in one launch (boxes
    overlap on edges and corners) against the oracle applied box by box."""
    nvar = 22
    gs = _blocks()
    z = Z4c(gs, Params())
    rng = np.random.default_rng(11)
    ch = _coarse_host(z, nvar, rng)
    fine0 = [rng.standard_normal((nvar,) + g.shape) for g in gs]
    boxes = [z.ProlongationRange(b, ox, parity) for b in range(len(gs)) for ox in OFFSETS]
    z.set_prolongation(boxes)
    u = z.upload(fine0)
    z.prolongate(_upload_coarse(z, ch), u, nvar)
    torch.cuda.synchronize()
    got = z.download(u, nvar)
    for b, g in enumerate(gs):
        exp = np.ascontiguousarray(fine0[b]).copy()
        for box in boxes:
            if box[0] == b:
                orc.prolongate_vc(_og(g), np.ascontiguousarray(ch[b]), exp, box[1:])
        assert np.array_equal(got[b], exp), b
        assert not np.array_equal(got[b], fine0[b])


def test_whole_block_prolongation_matches_oracle():
    """The AMR refinement form: the coarse interior box of each block
    (src/comm/amr_registry.cpp:760-797), with 1 and 22 fields."""
    gs = _blocks()
    z = Z4c(gs, Params())
    rng = np.random.default_rng(2)
    for nvar in (1, 22):
        ch = _coarse_host(z, nvar, rng)
        boxes = []
        for b, g in enumerate(gs):
            cs = z.coarse_shape(b)
            boxes.append((b, 4, cs[2] - 5, 4, cs[1] - 5, 4, cs[0] - 5))
        z.set_prolongation(boxes)
        fine0 = [np.full((nvar,) + g.shape, -7.0) for g in gs]
        u = z.upload(fine0)
        z.prolongate(_upload_coarse(z, ch), u, nvar)
        torch.cuda.synchronize()
        got = z.download(u, nvar)
        for b, g in enumerate(gs):
            exp = fine0[b].copy()
            orc.prolongate_vc(_og(g), np.ascontiguousarray(ch[b]), exp, boxes[b][1:])
            assert np.array_equal(got[b], exp), (nvar, b)
            inner = got[b][:, 4:-4, 4:-4, 4:-4]
            assert np.all(inner != -7.0)


def test_large_block_prolongation_matches_oracle():
    """A block whose coarse buffer (33 x 29 x 31 vertices) exceeds what one
    workgroup stages: the whole-block box and the 26 neighbour boxes are
    split into pieces, which must write exactly the oracle's values."""
    nvar = 3
    gs = [MeshBlockGeom((48, 40, 44), (0.02, 0.025, 0.03), (0.1, -0.3, 0.2)),
          MeshBlockGeom((8, 8, 8), (0.1, 0.1, 0.1), (0.0, 0.0, 0.0))]
    z = Z4c(gs, Params())
    rng = np.random.default_rng(13)
    ch = _coarse_host(z, nvar, rng)
    cs = z.coarse_shape(0)
    for boxes in ([(0, 4, cs[2] - 5, 4, cs[1] - 5, 4, cs[0] - 5)],
                  [z.ProlongationRange(0, ox, (1, 0, 1)) for ox in OFFSETS] +
                  [z.ProlongationRange(1, ox, (0, 1, 1)) for ox in OFFSETS]):
        z.set_prolongation(boxes)
        fine0 = [rng.standard_normal((nvar,) + g.shape) for g in gs]
        u = z.upload(fine0)
        z.prolongate(_upload_coarse(z, ch), u, nvar)
        torch.cuda.synchronize()
        got = z.download(u, nvar)
        for b, g in enumerate(gs):
            exp = fine0[b].copy()
            for box in boxes:
                if box[0] == b:
                    orc.prolongate_vc(_og(g), np.ascontiguousarray(ch[b]), exp, box[1:])
            assert np.array_equal(got[b], exp), b


def test_out_of_range_boxes_are_rejected():
    gs = _blocks()
    z = Z4c(gs, Params())
    cs = z.coarse_shape(0)
    for bad in [(0, 1, 5, 4, 8, 4, 8),              # reads below the coarse buffer
                (0, 4, cs[2] - 2, 4, 8, 4, 8),      # writes past the fine block
                (3, 4, 8, 4, 8, 4, 8),              # no such block
                (0, 6, 5, 4, 8, 4, 8)]:             # empty range
        with pytest.raises(Z4cError):
            z.set_prolongation([bad])
    odd = Z4c([MeshBlockGeom((7, 8, 8), (0.1,) * 3, (0.0,) * 3)], Params())
    with pytest.raises(Z4cError):
        odd.coarse_layout()


def test_level_transfer_of_a_quintic_16cubed():
    """Restrict a quintic to the coarse interior, give the coarse ghosts the
    quintic (as a coarser neighbour's data would), prolongate all 26 ghost
    boxes: every fine ghost equals the quintic to rounding, and the interior
    is untouched."""
    nvar = 22
    g = MeshBlockGeom.cube(16, 1.0 / 16, (-0.5, -0.25, 0.0))
    z = Z4c([g], Params())
    rng = np.random.default_rng(9)
    cs = [rng.standard_normal((nvar, 6)) for _ in range(3)]
    x = g.coords()

    def P(xx, yy, zz):
        return (np.polynomial.polynomial.polyval(xx, cs[0].T)[:, None, None, :]
                * np.polynomial.polynomial.polyval(yy, cs[1].T)[:, None, :, None]
                * np.polynomial.polynomial.polyval(zz, cs[2].T)[:, :, None, None])

    fine = P(x[0], x[1], x[2])
    og = _og(g)
    cg = og.coarse()
    xc = [cg.x0[d] + np.arange((cg.nn1, cg.nn2, cg.nn3)[d]) * cg.dx[d] for d in range(3)]
    coarse_exact = P(xc[0], xc[1], xc[2])
    base = fine.copy()
    base[:, :4] = base[:, -4:] = 0.0
    base[:, :, :4] = base[:, :, -4:] = 0.0
    base[:, :, :, :4] = base[:, :, :, -4:] = 0.0
    u = z.upload([base])
    cint = z.coarse_full(nvar, float("nan"))
    z.RestrictInterior(u, cint, nvar)
    # coarse interior from the restriction, coarse ghosts from the "neighbour"
    shp = (nvar,) + z.coarse_shape(0)
    cc = torch.from_numpy(np.ascontiguousarray(coarse_exact)).to(z.device).view(shp).clone()
    cc[:, 4:-4, 4:-4, 4:-4] = cint.view(shp)[:, 4:-4, 4:-4, 4:-4]
    assert np.allclose(cc.cpu().numpy(), coarse_exact, rtol=1e-13, atol=1e-13 * np.abs(fine).max())
    z.set_prolongation([z.ProlongationRange(0, ox, (1, 0, 1)) for ox in OFFSETS])
    z.prolongate(cc.reshape(-1).contiguous(), u, nvar)
    torch.cuda.synchronize()
    got = z.download(u, nvar)[0]
    scale = np.abs(fine).reshape(nvar, -1).max(axis=1)
    err = np.abs(got - fine).reshape(nvar, -1).max(axis=1) / scale
    assert np.all(err < 1e-12), err
    assert np.array_equal(got[:, 4:-4, 4:-4, 4:-4], base[:, 4:-4, 4:-4, 4:-4])
