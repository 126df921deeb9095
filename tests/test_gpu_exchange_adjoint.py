"""Adjoint of the same-level ghost exchange (z4c_exchange_vjp) and of the
fused RK4 step with the exchange after every stage (z4c_rk4_step_ghosts_vjp).

The exchange E is linear, so its transpose is pinned by the dot-product
identity <E x, y> = <x, E^T y> on random x, y (a wrong weight, a dropped or
extra term or a wrong vertex fails it), on periodic and open block sets.
The step's VJP is pinned by central differences of <w, step(u)> in the
state and in the gauge parameters, as the frozen-ghost step VJP is."""
import numpy as np
import pytest

import oracle as orc
from ml_meshes import refined_blocks

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from z4c_amd import Params, Z4cError  # noqa: E402
from z4c_amd._capi import PBAR  # noqa: E402
from z4c_amd.initial_data import perturbed_flat  # noqa: E402
from z4c_amd.z4c import MeshBlockGeom, Z4c  # noqa: E402

PUNCTURE = dict(lapse_oplog=2.0, shift_Gamma=1.0, shift_eta=2.0, damp_kappa1=0.02,
                diss=0.02, chi_div_floor=1e-5)


def _grid(nb, n, h, periodic):
    """nb[0] x nb[1] x nb[2] blocks of n^3 cells, spacing h, from the origin."""
    geoms = [MeshBlockGeom.cube(n, h, (i * n * h, j * n * h, k * n * h))
             for k in range(nb[2]) for j in range(nb[1]) for i in range(nb[0])]
    hi = [nb[d] * n * h for d in range(3)]
    return geoms, [0.0] * 3, hi, periodic


MESHES = {
    "periodic_2x2x1": _grid((2, 2, 1), 8, 1.0 / 8, (True, True, True)),
    "open_x_2x1x1": _grid((2, 1, 1), 8, 1.0 / 8, (False, True, True)),
    "one_block_periodic": _grid((1, 1, 1), 8, 1.0 / 8, (True, True, True)),
}


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("nvar", [22, 3])
def test_exchange_vjp_dot_product(name, nvar):
    geoms, lo, hi, periodic = MESHES[name]
    z = Z4c(geoms, Params())
    z.set_exchange(lo, hi, periodic)
    rng = np.random.default_rng(17)
    x = z.upload([rng.standard_normal((nvar,) + g.shape) for g in geoms], nvar)
    y = z.upload([rng.standard_normal((nvar,) + g.shape) for g in geoms], nvar)
    ex = x.clone()
    z.exchange(ex, nvar)
    ety = y.clone()
    z.exchange_vjp(ety, nvar)
    ety2 = y.clone()
    z.exchange_vjp(ety2, nvar)
    torch.cuda.synchronize()
    lhs, rhs = float((ex * y).sum()), float((x * ety).sum())
    scale = float((ex.abs() * y.abs()).sum())
    print(f"{name}, {nvar} fields: <Ex, y> = {lhs:.15e}, <x, E^T y> = {rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-13 * scale, (lhs, rhs)
    assert torch.equal(ety, ety2)  # a fixed-order gather: bit-identical repeats
    # the transpose moves something (ghosts receive zero, sources more)
    assert not torch.equal(ety, y)
    z.close()


def test_exchange_vjp_rejects_levels():
    """The update across levels (prolongation included) has no transpose."""
    spec = refined_blocks(2, 8, [-0.5] * 3, [0.5] * 3, lambda c, L: bool(np.all(c < 0.0)), 2,
                          (True,) * 3)
    geoms = [MeshBlockGeom(tuple(n), tuple(dx), tuple(x0)) for n, dx, x0 in spec]
    z = Z4c(geoms, Params())
    z.set_exchange([-0.5] * 3, [0.5] * 3, (True,) * 3)
    ub = z.zeros()
    with pytest.raises(Z4cError):
        z.exchange_vjp(ub)
    with pytest.raises(Z4cError):
        z.rk4_step_ghosts_vjp(0.01, z.zeros(), ub, 1)
    z.close()


def _state(geoms, p, seed):
    """Perturbed flat state per block with a constant shift added (fixed
    upwind selection under the finite differences), after AlgConstr."""
    op = orc.Params(**p.as_dict())
    out = []
    for b, g in enumerate(geoms):
        og = orc.Geom.block(g.nx, g.dx, g.xmin)
        u = np.ascontiguousarray(perturbed_flat(g, seed=seed + b, amp=(1e-3, 3e-2)))
        for a, s in enumerate((0.3, -0.25, 0.2)):
            u[19 + a] += s
        out.append(orc.algconstr(op, og, u))
    return out


def test_step_ghosts_vjp_finite_differences():
    """Central differences (e = 1e-6) of <w, step(u)> for the step with the
    exchange after every stage on four periodic 8^3 blocks; eps_floor far
    below zero for the reason test_gpu_adjoint's step test gives."""
    geoms, lo, hi, periodic = _grid((2, 1, 2), 8, 1.0 / 16, (True, True, True))
    p = Params(**dict(PUNCTURE, eps_floor=-1e30, damp_kappa2=0.05))
    z = Z4c(geoms, p)
    z.set_exchange(lo, hi, periodic)
    u_d = z.upload(_state(geoms, p, seed=11))
    dt = 0.25 / 16
    rng = np.random.default_rng(4)
    w_d = z.upload([rng.standard_normal((22,) + g.shape) for g in geoms])
    v_d = z.upload([rng.standard_normal((22,) + g.shape) for g in geoms])
    ubar, pbar = z.rk4_step_ghosts_vjp(dt, u_d, w_d, 1)
    ubar2, _ = z.rk4_step_ghosts_vjp(dt, u_d, w_d, 1)
    assert torch.equal(ubar, ubar2)
    work = z.make_step_work()
    eps = 1e-6

    def phi(uu, **kw):
        q = Params(**dict(p.as_dict(), **kw))
        out = uu.clone()
        saved = z.params
        z.params = q
        z.rk4_step_fused(dt, out, work, exchange=True)
        z.params = saved
        return float((out * w_d).sum())

    fd = (phi(u_d + eps * v_d) - phi(u_d - eps * v_d)) / (2 * eps)
    an = float((ubar * v_d).sum())
    print(f"state: fd {fd:.12e}, adjoint {an:.12e}", flush=True)
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
    for t, name in enumerate(PBAR):
        x = getattr(p, name)
        h = 1e-3 * max(1.0, abs(x))
        fd = (-phi(u_d, **{name: x + 2 * h}) + 8 * phi(u_d, **{name: x + h})
              - 8 * phi(u_d, **{name: x - h}) + phi(u_d, **{name: x - 2 * h})) / (12 * h)
        an = float(pbar[0, t])
        print(f"{name}: fd {fd:.12e}, adjoint {an:.12e}", flush=True)
        assert abs(fd - an) <= 1e-6 * abs(an), (name, fd, an)
    z.close()


def test_step_ghosts_vjp_rejects_physical_faces():
    geoms, lo, hi, _ = _grid((2, 1, 1), 8, 1.0 / 8, (False, True, True))
    geoms[0].outflow_faces = 1
    p = Params(**PUNCTURE)
    z = Z4c(geoms, p)
    z.set_exchange(lo, hi, (False, True, True))
    u = z.zeros()
    with pytest.raises(Z4cError):
        z.rk4_step_ghosts_vjp(0.01, u, z.zeros(), 1)
    z.close()
