"""GPU z4c_weyl (psi4, src/z4c/calculate_weyl_scalars.cpp:44-628) against the
oracle (tests/test_oracle_weyl.py pins the oracle on exact spacetimes).

psi4 is a cancellation-dominated sum of O(1) curvature terms (second
derivatives ~1/h^2), so the bound is max|gpu - oracle| <= 1e-10 of the
largest term scale (max |R_ab| + |K|^2 ~ 1 here) over the interior."""
import numpy as np
import pytest
import torch

import oracle as orc
from z4c_amd import Params
from z4c_amd.initial_data import brill_lindquist_adm, perturbed_flat
from z4c_amd.z4c import MeshBlockGeom, Z4c

pytestmark = pytest.mark.gpu


def _interior(og, a):
    return a[:, og.kl:og.ku + 1, og.jl:og.ju + 1, og.il:og.iu + 1]


def _check(gs, p, states):
    op = orc.Params(**p.as_dict())
    z = Z4c(gs, p)
    u = z.upload(states)
    adm = z.zeros(17)
    z.Z4cToADM(u, adm)
    w = z.weyl(u, adm)
    torch.cuda.synchronize()
    got = z.download(w, 2)
    for b, g in enumerate(gs):
        og = orc.Geom.block(g.nx, g.dx, g.xmin)
        ref = orc.weyl(op, og, states[b])
        gi, ri = _interior(og, got[b]), _interior(og, ref)
        assert np.all(np.isfinite(gi))
        err = np.abs(gi - ri).max()
        assert err <= 1e-10 * max(1.0, np.abs(ri).max()), (b, err, np.abs(ri).max())
        # outside the interior the buffer holds NaN, as the reference's u_weyl
        assert np.isnan(got[b][0, 0, 0, 0]) and np.isnan(got[b][1, -1, -1, -1])
    return got


def test_weyl_two_punctures_and_uneven_batch():
    p = Params(chi_div_floor=1e-5)
    op = orc.Params(**p.as_dict())
    gs = [MeshBlockGeom.cube(12, 0.25, (0.4, -1.3, 0.7)), MeshBlockGeom.cube(8, 0.125, (3.0, 0.2, 0.3))]
    states = []
    for g in gs:
        og = orc.Geom.block(g.nx, g.dx, g.xmin)
        adm = brill_lindquist_adm(g, m=(0.5, 0.5), sep=3.0)
        states.append(orc.adm_to_z4c(op, og, np.ascontiguousarray(adm)))
    got = _check(gs, p, states)
    assert np.nanmax(np.abs(got[0])) > 1e-4  # a nonzero psi4


@pytest.mark.parametrize("pw", [-4.0, -2.0])
def test_weyl_perturbed_state(pw):
    """Generic perturbed data (nonzero K, A~, both chi_psi_power branches)."""
    p = Params(chi_div_floor=1e-5, chi_psi_power=pw)
    op = orc.Params(**p.as_dict())
    g = MeshBlockGeom.cube(10, 0.1, (0.15, -0.45, 0.35))
    og = orc.Geom.block(g.nx, g.dx, g.xmin)
    u = orc.algconstr(op, og, np.ascontiguousarray(perturbed_flat(g, seed=4, amp=(1e-3, 3e-2))))
    _check([g], p, [u])
