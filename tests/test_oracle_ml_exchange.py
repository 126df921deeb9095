"""Pins of the multi-level ghost-fill oracle (oracle/ml_exchange.py).

* One level: equals the same-level exchange oracle (itself pinned by brute
  force over vertex classes).
* A constant field stays exactly constant at every fine vertex inside the
  domain and at the coarse-buffer vertices the prolongation reads: a wrong
  multiplicity, a missing or doubled contribution or an unfilled ghost
  shows up as a different value.
* A tensor product of quintics (which the restriction copies and the VC
  prolongation reproduces) is reproduced at every fine vertex inside the
  domain and at the coarse-buffer vertices the prolongation reads, on two-
  and three-level meshes with open faces: a wrong index range, offset,
  parity, fi choice or neighbour gives a value of the polynomial at another
  point.

The coarse buffer is not checked elsewhere: its multiplicity is the fine
array's (NodeMultiplicity::ApplyDivisionCoarse), so next to a finer
neighbour, which sends nothing to the coarse buffer, coarse vertices that
no prolongation reads are divided by one contribution too many or stay
zero, as in the reference.
"""
import numpy as np
import pytest

import ml_exchange as mlx
import oracle as orc
from ml_meshes import open_faces, quartic, quintic, refined_blocks


def _mesh(spec, lo, hi, periodic):
    blocks = [mlx.Block(*b) for b in spec]
    return mlx.Mesh(blocks, lo, hi, periodic)


def _coords(b, coarse=False):
    n = (b.nx[0] // 2 + 2 * mlx.CNG + 1) if coarse else (b.nx[0] + 2 * mlx.NG + 1)
    h = [2 * d if coarse else d for d in b.dx]
    g = mlx.CNG if coarse else mlx.NG
    ax = [b.xmin[d] - g * h[d] + np.arange(n) * h[d] for d in range(3)]
    z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return x, y, z


def _fill_interiors(m, fn, nvar=2):
    us = []
    for b in m.blocks:
        x, y, z = _coords(b)
        u = np.stack([fn(x, y, z) * (1.0 + 0.5 * v) for v in range(nvar)])
        # ghosts hold garbage before the update
        g = np.ones(u.shape[1:], bool)
        g[mlx.NG:-mlx.NG, mlx.NG:-mlx.NG, mlx.NG:-mlx.NG] = False
        u[:, g] = 1e30
        us.append(np.ascontiguousarray(u))
    return us


def test_one_level_equals_same_level_exchange():
    spec = refined_blocks(2, 8, [0.0] * 3, [1.0] * 3, lambda c, L: False, 1)
    m = _mesh(spec, [0.0] * 3, [1.0] * 3, (True, True, True))
    assert all(len(n) == 26 for n in m.nbrs)
    rng = np.random.default_rng(3)
    us = [np.ascontiguousarray(rng.standard_normal((3, 17, 17, 17))) for _ in m.blocks]
    ref = [u.copy() for u in us]
    orc.exchange([b.geom for b in m.blocks], ref, [0.0] * 3, [1.0] * 3, [1, 1, 1])
    m.ghost_fill(us)
    for a, b in zip(us, ref):
        np.testing.assert_allclose(a, b, rtol=0, atol=4e-15 * np.abs(b).max())


PERIODIC_2L = dict(nrb=4, nx=8, lo=[-2.0] * 3, hi=[2.0] * 3,
                   refine=lambda c, L: bool(np.all(c < -1.0)), levels=2)


@pytest.mark.parametrize("case", ["periodic_corner", "open_centre", "open_3level"])
def test_constant_field_stays_constant(case):
    m = _case(case)
    us = _fill_interiors(m, lambda x, y, z: 0.0 * x + 1.75, nvar=1)
    cs = m.ghost_fill(us)
    for q, (u, c) in enumerate(zip(us, cs)):
        x, y, z = _coords(m.blocks[q])
        inside = _inside(m, x, y, z)
        assert np.all(u[0][inside] == 1.75), q
        if any(n["level"] < m.level[q] for n in m.nbrs[q]):
            assert np.all(c[0][_read(m, q)] == 1.75), q


@pytest.mark.parametrize("case", ["open_centre", "open_3level"])
def test_quintic_is_reproduced(case):
    m = _case(case)
    us = _fill_interiors(m, quintic)
    cs = m.ghost_fill(us)
    nprol = 0
    for q, (u, c) in enumerate(zip(us, cs)):
        x, y, z = _coords(m.blocks[q])
        inside = _inside(m, x, y, z)
        exp = quintic(x, y, z)
        for v in range(u.shape[0]):
            err = np.abs(u[v] - exp * (1.0 + 0.5 * v))[inside].max()
            assert err <= 1e-13, (q, v, err)
        if any(n["level"] < m.level[q] for n in m.nbrs[q]):
            nprol += 1
            cx, cy, cz = _coords(m.blocks[q], coarse=True)
            err = np.abs(c[0] - quintic(cx, cy, cz))[_read(m, q)].max()
            assert err <= 1e-13, (q, err)
    assert nprol > 0


def test_neighbour_lists_follow_the_tree():
    """Counts on the two-level open mesh: fine blocks next to the coarse
    region see coarser neighbours only where they sit in that corner or
    edge of their parent; coarse blocks see four finer blocks across a
    face, two across an edge and one across a corner."""
    m = _case("open_centre")
    for q in range(len(m.blocks)):
        kinds = {}
        for n in m.nbrs[q]:
            d = sum(1 for o in n["ox"] if o != 0)
            kinds.setdefault((d, n["level"] - m.level[q]), 0)
            kinds[(d, n["level"] - m.level[q])] += 1
        if m.level[q] == 0:
            for (d, rel), cnt in kinds.items():
                if rel == 1:
                    assert cnt % {1: 4, 2: 2, 3: 1}[d] == 0
        else:
            my = [2 * (m.loc[q][d] & 1) - 1 for d in range(3)]
            for n in m.nbrs[q]:
                if n["level"] < m.level[q]:
                    nz = [d for d in range(3) if n["ox"][d] != 0]
                    if len(nz) > 1:
                        assert all(n["ox"][d] == my[d] for d in nz)


def _case(name):
    if name == "periodic_corner":
        spec = refined_blocks(4, 8, [-2.0] * 3, [2.0] * 3, lambda c, L: bool(np.all(c < -1.0)), 2)
        return _mesh(spec, [-2.0] * 3, [2.0] * 3, (True, True, True))
    if name == "open_centre":
        spec = refined_blocks(4, 8, [-2.0] * 3, [2.0] * 3,
                              lambda c, L: bool(np.all(np.abs(c) < 1.0)), 2, (False,) * 3)
        return _mesh(spec, [-2.0] * 3, [2.0] * 3, (False,) * 3)
    if name == "open_3level":
        spec = refined_blocks(4, 8, [-2.0] * 3, [2.0] * 3,
                              lambda c, L: bool(np.all(np.abs(c) < (1.0 if L == 0 else 0.5))), 3,
                              (False,) * 3)
        return _mesh(spec, [-2.0] * 3, [2.0] * 3, (False,) * 3)
    if name == "open_corner":
        spec = refined_blocks(4, 8, [-2.0] * 3, [2.0] * 3,
                              lambda c, L: bool(np.all(c < -1.0)), 2, (False,) * 3)
        return _mesh(spec, [-2.0] * 3, [2.0] * 3, (False,) * 3)
    raise KeyError(name)


def test_coarse_physical_fill_at_open_faces():
    """Fine blocks in a refined corner of an open domain prolongate next to
    the domain boundary, reading coarse-buffer ghosts beyond it; the coarse
    level's physical fill (5-point extrapolation over the coarse VC ranges)
    makes a quartic exact at every fine vertex inside the domain, and
    without it those vertices are wrong."""
    m = _case("open_corner")
    faces = [open_faces(m, q) for q in range(len(m.blocks))]
    fine_at_face = [q for q in range(len(m.blocks))
                    if m.level[q] == 1 and any(faces[q]) and
                    any(n["level"] < 1 for n in m.nbrs[q])]
    assert fine_at_face
    us = _fill_interiors(m, quartic)
    m.ghost_fill(us, faces=faces)
    for q, u in enumerate(us):
        x, y, z = _coords(m.blocks[q])
        inside = _inside(m, x, y, z)
        err = np.abs(u[0] - quartic(x, y, z))[inside].max()
        assert err <= 1e-13, (q, err)
    bad = _fill_interiors(m, quartic)
    m.ghost_fill(bad)
    worst = max(np.abs(bad[q][0] - quartic(*_coords(m.blocks[q])))[
        _inside(m, *_coords(m.blocks[q]))].max() for q in fine_at_face)
    assert worst > 1e-3


def _read(m, q):
    """Coarse vertices the prolongation of block q certainly reads: every
    coarser neighbour's ProlongationRange box widened by the two coarse
    points on each side that the 6-point Lagrange stencils of its odd fine
    vertices reach."""
    shape = m.coarse_shape()
    mask = np.zeros(shape, bool)
    for n in m.nbrs[q]:
        if n["level"] >= m.level[q]:
            continue
        si, ei, sj, ej, sk, ek = orc.prolongation_range(m.blocks[q].geom, n["ox"], m.loc[q])
        mask[max(sk - 2, 0):ek + 3, max(sj - 2, 0):ej + 3, max(si - 2, 0):ei + 3] = True
    return mask


def _inside(m, x, y, z):
    eps = 1e-9
    ok = np.ones(x.shape, bool)
    for d, a in enumerate((x, y, z)):
        if not m.periodic[d]:
            ok &= (a >= m.lo[d] - eps) & (a <= m.hi[d] + eps)
    return ok
