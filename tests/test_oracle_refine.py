"""Pins the oracle's VC restriction / prolongation (src/mesh/mesh_refinement_vc.cpp,
comm/refinement_ops.*; CPU):

* restriction is injection: every coarse interior vertex equals the fine
  vertex at the same coordinate;
* for every one of the 26 neighbour offsets of a coarser neighbour and both
  logical-coordinate parities, prolongation over ProlongationRange's box
  writes exactly the fine ghost region that neighbour covers (4 ghost layers
  on each side with a nonzero offset; along a zero offset the interior plus
  the 4 ghosts on the side the parity selects) and nothing else;
* the values written reproduce any tensor product of quintics sampled on the
  coarse grid (the 6-point Lagrange weights are exact to degree 5), by the
  fine vertices' coordinates, and miss a degree-6 term.
"""
import itertools

import numpy as np
import pytest

import oracle as orc

NX = 8


def _geom():
    return orc.Geom.block(NX, (0.1, 0.12, 0.08), (0.3, -0.2, 0.5))


def _coords(g):
    return [g.x0[d] + np.arange((g.nn1, g.nn2, g.nn3)[d]) * g.dx[d] for d in range(3)]


def _poly(rng, deg):
    cs = [rng.standard_normal(deg + 1) for _ in range(3)]
    return lambda x, y, z: (np.polyval(cs[0], x) * np.polyval(cs[1], y) * np.polyval(cs[2], z))


def _sample(g, f):
    x = _coords(g)
    Zg, Yg, Xg = np.meshgrid(x[2], x[1], x[0], indexing="ij")
    return f(Xg, Yg, Zg)


def test_restriction_is_injection():
    g = _geom()
    cg = g.coarse()
    rng = np.random.default_rng(1)
    fine = rng.standard_normal((3,) + g.shape)
    c = orc.restrict_vc(g, fine)
    xf, xc = _coords(g), _coords(cg)
    for ck in range(cg.kl, cg.ku + 1):
        for cj in range(cg.jl, cg.ju + 1):
            for ci in range(cg.il, cg.iu + 1):
                idx = [int(round((xc[d][(ci, cj, ck)[d]] - xf[d][0]) / g.dx[d])) for d in range(3)]
                assert np.array_equal(c[:, ck, cj, ci], fine[:, idx[2], idx[1], idx[0]])
    assert np.isnan(c[0, 0, 0, 0])


def _expected_axis(g, a, ox, parity):
    lo, hi = (g.il, g.jl, g.kl)[a], (g.iu, g.ju, g.ku)[a]
    if ox > 0:
        return set(range(hi + 1, hi + 5))
    if ox < 0:
        return set(range(lo - 4, lo))
    return set(range(lo, hi + 5)) if parity == 0 else set(range(lo - 4, hi + 1))


@pytest.mark.parametrize("parity", [(0, 0, 0), (1, 1, 1), (0, 1, 0)])
def test_prolongation_region_and_quintic_exactness(parity):
    g = _geom()
    cg = g.coarse()
    rng = np.random.default_rng(7)
    P = _poly(rng, 5)
    coarse = _sample(cg, P)[None]
    exact = _sample(g, P)
    for ox in itertools.product((-1, 0, 1), repeat=3):
        if ox == (0, 0, 0):
            continue
        fine = np.full((1,) + g.shape, np.nan)
        rng_box = orc.prolongation_range(g, ox, parity)
        orc.prolongate_vc(g, coarse, fine, rng_box)
        written = ~np.isnan(fine[0])
        sets = [_expected_axis(g, a, ox[a], parity[a]) for a in range(3)]
        want = np.zeros(g.shape, bool)
        for k in sets[2]:
            for j in sets[1]:
                for i in sets[0]:
                    want[k, j, i] = True
        assert np.array_equal(written, want), ox
        err = np.abs(fine[0][written] - exact[written]).max()
        assert err <= 1e-12 * np.abs(exact[written]).max(), (ox, err)


def test_prolongation_misses_degree_six():
    g = _geom()
    cg = g.coarse()
    P = lambda x, y, z: (x - 0.31) ** 6 + 0.0 * y * z  # noqa: E731
    coarse = _sample(cg, P)[None]
    fine = np.full((1,) + g.shape, np.nan)
    orc.prolongate_vc(g, coarse, fine, orc.prolongation_range(g, (1, 0, 0), (0, 0, 0)))
    exact = _sample(g, P)
    w = ~np.isnan(fine[0])
    err = np.abs(fine[0][w] - exact[w]).max()
    assert err > 1e-9, err
