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


ML_CASES = {
    "periodic_corner": dict(nrb=4, lo=-2.0, hi=2.0, periodic=(True,) * 3, levels=2,
                            refine=lambda c, L: bool(np.all(c < -1.0))),
    "open_centre": dict(nrb=4, lo=-2.0, hi=2.0, periodic=(False,) * 3, levels=2,
                        refine=lambda c, L: bool(np.all(np.abs(c) < 1.0))),
    "open_3level": dict(nrb=4, lo=-2.0, hi=2.0, periodic=(False,) * 3, levels=3,
                        refine=lambda c, L: bool(np.all(np.abs(c) < (1.0 if L == 0 else 0.5)))),
    "open_corner": dict(nrb=4, lo=-2.0, hi=2.0, periodic=(False,) * 3, levels=2,
                        refine=lambda c, L: bool(np.all(c < -1.0))),
    "periodic_2level": dict(nrb=2, lo=-0.5, hi=0.5, periodic=(True,) * 3, levels=2,
                            refine=lambda c, L: bool(np.all(c < 0.0))),
}


def _ml_geoms(case, nx=8):
    """The block tree of an ML case; every face on an open domain boundary
    carries the outflow fill (the update across levels needs it there)."""
    c = ML_CASES[case]
    spec = refined_blocks(c["nrb"], nx, [c["lo"]] * 3, [c["hi"]] * 3, c["refine"], c["levels"],
                          c["periodic"])
    geoms = [MeshBlockGeom(tuple(n), tuple(dx), tuple(x0)) for n, dx, x0 in spec]
    h0 = max(g.dx[0] for g in geoms)
    nrb = c["nrb"]
    for g in geoms:
        lev = int(round(np.log2(h0 / g.dx[0])))
        m = nrb << lev
        f = 0
        for d in range(3):
            lx = int(round((g.xmin[d] - c["lo"]) / (g.nx[d] * g.dx[d])))
            if not c["periodic"][d]:
                f |= (1 << (2 * d)) if lx == 0 else 0
                f |= (2 << (2 * d)) if lx == m - 1 else 0
        g.outflow_faces = f
    return c, geoms


@pytest.mark.parametrize("case", ["periodic_corner", "open_centre", "open_3level", "open_corner"])
@pytest.mark.parametrize("nvar", [22, 3])
def test_ml_exchange_vjp_dot_product(case, nvar):
    """<G x, y> = <x, G^T y> for the update across levels (restriction,
    gather, coarse physical fill, prolongation)."""
    c, geoms = _ml_geoms(case)
    z = Z4c(geoms, Params())
    z.set_exchange([c["lo"]] * 3, [c["hi"]] * 3, c["periodic"])
    rng = np.random.default_rng(23)
    x = z.upload([rng.standard_normal((nvar,) + g.shape) for g in geoms], nvar)
    y = z.upload([rng.standard_normal((nvar,) + g.shape) for g in geoms], nvar)
    gx = x.clone()
    z.exchange(gx, nvar)
    gty = y.clone()
    z.exchange_vjp(gty, nvar)
    gty2 = y.clone()
    z.exchange_vjp(gty2, nvar)
    torch.cuda.synchronize()
    lhs, rhs = float((gx * y).sum()), float((x * gty).sum())
    scale = float((gx.abs() * y.abs()).sum())
    print(f"{case}, {nvar} fields, {len(geoms)} blocks: <Gx, y> = {lhs:.15e}, "
          f"<x, G^T y> = {rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs)
    assert torch.equal(gty, gty2)
    z.close()


def test_ml_exchange_vjp_dot_product_large_blocks():
    """<G x, y> = <x, G^T y> across levels on 72^3-cell blocks, whose
    prolongation transpose tiles need more than 76 KiB of LDS (the chooser's
    fallback to the device's LDS)."""
    c, geoms = _ml_geoms("periodic_2level", nx=72)
    nvar = 3
    z = Z4c(geoms, Params())
    z.set_exchange([c["lo"]] * 3, [c["hi"]] * 3, c["periodic"])
    rng = np.random.default_rng(31)
    x = z.upload([rng.standard_normal((nvar,) + g.shape) for g in geoms], nvar)
    y = z.upload([rng.standard_normal((nvar,) + g.shape) for g in geoms], nvar)
    gx = x.clone()
    z.exchange(gx, nvar)
    gty = y.clone()
    z.exchange_vjp(gty, nvar)
    torch.cuda.synchronize()
    lhs, rhs = float((gx * y).sum()), float((x * gty).sum())
    scale = float((gx.abs() * y.abs()).sum())
    print(f"{len(geoms)} blocks of 72^3: <Gx, y> = {lhs:.15e}, <x, G^T y> = {rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs)
    z.close()


def test_ml_exchange_vjp_context_reuse():
    """One context: the transpose data are built on the first call, serve
    any field count, and are rebuilt by a second z4c_set_exchange; every
    result equals a fresh context's bit for bit."""
    c, geoms = _ml_geoms("open_3level")
    rng = np.random.default_rng(29)
    ys = {nv: [rng.standard_normal((nv,) + g.shape) for g in geoms] for nv in (22, 3)}

    def fresh(nv):
        z = Z4c(geoms, Params())
        z.set_exchange([c["lo"]] * 3, [c["hi"]] * 3, c["periodic"])
        y = z.upload(ys[nv], nv)
        z.exchange_vjp(y, nv)
        torch.cuda.synchronize()
        out = y.clone()
        z.close()
        return out

    ref = {nv: fresh(nv) for nv in (22, 3)}
    z = Z4c(geoms, Params())
    z.set_exchange([c["lo"]] * 3, [c["hi"]] * 3, c["periodic"])
    for nv in (22, 3, 22):
        y = z.upload(ys[nv], nv)
        z.exchange_vjp(y, nv)
        torch.cuda.synchronize()
        assert torch.equal(y, ref[nv]), nv
    z.set_exchange([c["lo"]] * 3, [c["hi"]] * 3, c["periodic"])
    y = z.upload(ys[3], 3)
    z.exchange_vjp(y, 3)
    torch.cuda.synchronize()
    assert torch.equal(y, ref[3])
    z.close()


def test_physical_bc_vjp_dot_product():
    """<F x, y> = <x, F^T y> for the physical fill: every face of one block
    outflow, two faces of a second, and a third block with Sommerfeld faces
    (the fill extrapolates their ghosts the same way) and one outflow face."""
    g0 = MeshBlockGeom.cube(8, 1.0 / 8, (0.0, 0.0, 0.0))
    g0.outflow_faces = 63
    g1 = MeshBlockGeom((10, 8, 12), (0.1, 0.12, 0.08), (2.0, 0.0, 0.0))
    g1.outflow_faces = 2 | 16
    g2 = MeshBlockGeom((8, 10, 8), (0.1, 0.1, 0.1), (-3.0, 0.0, 0.0))
    g2.sommerfeld_faces = 1 | 8 | 32
    g2.outflow_faces = 4
    geoms = [g0, g1, g2]
    z = Z4c(geoms, Params())
    rng = np.random.default_rng(29)
    x = z.upload([rng.standard_normal((22,) + g.shape) for g in geoms])
    y = z.upload([rng.standard_normal((22,) + g.shape) for g in geoms])
    fx = x.clone()
    z.physical_bc(fx)
    fty = y.clone()
    z.physical_bc_vjp(fty)
    torch.cuda.synchronize()
    lhs, rhs = float((fx * y).sum()), float((x * fty).sum())
    scale = float((fx.abs() * y.abs()).sum())
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs)
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


def test_step_ghosts_vjp_rejects_unfilled_and_sommerfeld_faces():
    """Outflow faces need the fill in the step (ghosts = 6); Sommerfeld
    faces have no adjoint here."""
    geoms, lo, hi, _ = _grid((2, 1, 1), 8, 1.0 / 8, (False, True, True))
    geoms[0].outflow_faces = 1
    p = Params(**PUNCTURE)
    z = Z4c(geoms, p)
    z.set_exchange(lo, hi, (False, True, True))
    u = z.zeros()
    with pytest.raises(Z4cError):
        z.rk4_step_ghosts_vjp(0.01, u, z.zeros(), 1)
    z.close()
    geoms, lo, hi, _ = _grid((2, 1, 1), 8, 1.0 / 8, (False, True, True))
    geoms[0].sommerfeld_faces = 1
    z = Z4c(geoms, p)
    z.set_exchange(lo, hi, (False, True, True))
    with pytest.raises(Z4cError):
        z.rk4_step_ghosts_vjp(0.01, z.zeros(), z.zeros(), 1, physical=True)
    z.close()


def _smooth_state(z, geoms, p):
    """The gauge wave of period 4 along x (smooth across all blocks) through
    ADMToZ4c, a constant shift added (fixed upwind selection), AlgConstr."""
    from z4c_amd.initial_data import gauge_wave_adm
    adm = z.upload([gauge_wave_adm(g, A=0.05, d=4.0) for g in geoms], 17)
    u = z.zeros()
    z.ADMToZ4c(adm, u)
    for b in range(len(geoms)):
        v = z.block_view(u, b, 22)
        for a, sh in enumerate((0.3, -0.25, 0.2)):
            v[19 + a] += sh
    z.AlgConstr(u)
    torch.cuda.synchronize()
    return u


@pytest.mark.parametrize("case,physical", [("periodic_2level", False), ("open_centre", True)])
def test_ml_step_ghosts_vjp_finite_differences(case, physical):
    """Central differences of <w, step(u)> for the step with the update
    across levels after every stage.  AlgConstr's conformal factor takes a
    constant branch where det(gt) - 1 < eps_floor
    (src/utils/floating_point.hpp:120-131); a constrained state has
    det(gt) = 1 to rounding, so at the default eps_floor (1e-12) the stage
    states sit on that kink and differences straddle it.  Periodic mesh:
    per-block bumps and eps_floor far below zero (the pow branch
    everywhere).  Open mesh with the outflow fill: a state smooth across all
    blocks (the fill extrapolates it outward without amplifying jumps),
    eps_floor = -0.5 (the kink at det(gt) = 1/2, away from the state, while
    the ghost layer the fill leaves at zero, the upper faces' outermost as
    in the reference, takes the finite constant branch) and w, v at interior
    vertices (ghost corners extrapolate that zero layer)."""
    c, geoms = _ml_geoms(case)
    if physical:
        p = Params(**dict(PUNCTURE, eps_floor=-0.5, damp_kappa2=0.05))
    else:
        p = Params(**dict(PUNCTURE, eps_floor=-1e30, damp_kappa2=0.05))
    z = Z4c(geoms, p)
    z.set_exchange([c["lo"]] * 3, [c["hi"]] * 3, c["periodic"])
    u_d = _smooth_state(z, geoms, p) if physical else z.upload(_state(geoms, p, seed=31))
    dt = 0.25 * min(g.dx[0] for g in geoms)
    rng = np.random.default_rng(6)

    def field(g):
        """Random values; on the open mesh at interior vertices only."""
        a = rng.standard_normal((22,) + g.shape)
        if physical:
            m = np.zeros(g.shape, bool)
            m[4:-4, 4:-4, 4:-4] = True
            a *= m
        return a

    w_d = z.upload([field(g) for g in geoms])
    v_d = z.upload([field(g) for g in geoms])
    ubar, pbar = z.rk4_step_ghosts_vjp(dt, u_d, w_d, 1, physical=physical)
    work = z.make_step_work()
    eps = 1e-6

    def phi(uu, **kw):
        q = Params(**dict(p.as_dict(), **kw))
        out = uu.clone()
        saved = z.params
        z.params = q
        z.rk4_step_fused(dt, out, work, exchange=True, physical=physical)
        z.params = saved
        return float((out * w_d).sum())

    # fourth order: on the open mesh the step's curvature leaves the
    # two-point difference ~1e-5 off at eps = 1e-6
    fd = (-phi(u_d + 2 * eps * v_d) + 8 * phi(u_d + eps * v_d) - 8 * phi(u_d - eps * v_d)
          + phi(u_d - 2 * eps * v_d)) / (12 * eps)
    an = float((ubar * v_d).sum())
    print(f"{case}: {len(geoms)} blocks; state: fd {fd:.12e}, adjoint {an:.12e}", flush=True)
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
    for t, name in enumerate(PBAR[:2]):
        x = getattr(p, name)
        # Theta ~ 0 in the constraint-satisfying open-mesh state leaves the
        # damping derivative ~1e-3 of the sum's magnitude: a step of 1e-2
        # keeps rounding in the sum below the tolerance
        h = 1e-2 * max(1.0, abs(x))
        fd = (-phi(u_d, **{name: x + 2 * h}) + 8 * phi(u_d, **{name: x + h})
              - 8 * phi(u_d, **{name: x - h}) + phi(u_d, **{name: x - 2 * h})) / (12 * h)
        an = float(pbar[0, t])
        print(f"{name}: fd {fd:.12e}, adjoint {an:.12e}", flush=True)
        assert abs(fd - an) <= 1e-6 * abs(an), (name, fd, an)
    z.close()


def test_step_ghosts_vjp_batched_cotangents():
    """Two cotangents in one z4c_rk4_step_ghosts_vjp call (the update across
    levels and the outflow fill after every stage) equal two single calls
    bit for bit, state and parameter cotangents."""
    c, geoms = _ml_geoms("open_centre")
    p = Params(**dict(PUNCTURE, eps_floor=-0.5, damp_kappa2=0.05))
    z = Z4c(geoms, p)
    z.set_exchange([c["lo"]] * 3, [c["hi"]] * 3, c["periodic"])
    u_d = _smooth_state(z, geoms, p)
    dt = 0.25 * min(g.dx[0] for g in geoms)
    rng = np.random.default_rng(41)
    ws = [z.upload([rng.standard_normal((22,) + g.shape) for g in geoms]) for _ in range(2)]
    both, pb = z.rk4_step_ghosts_vjp(dt, u_d, torch.cat(ws), 2, physical=True)
    n = ws[0].numel()
    for q in range(2):
        one, pb1 = z.rk4_step_ghosts_vjp(dt, u_d, ws[q], 1, physical=True)
        torch.cuda.synchronize()
        assert torch.equal(both[q * n:(q + 1) * n], one), q
        assert torch.equal(pb[q], pb1[0]), q
    assert not torch.equal(both[:n], both[n:])
    z.close()


def test_step_ghosts_vjp_single_level_outflow_finite_differences():
    """The step VJP with the same-level exchange and the outflow fill after
    every stage (ghosts = 6 on one spacing): 2 x 2 x 2 blocks of 8^3 cells on
    an open box, the gauge-wave state with a shift, eps_floor = -0.5 and w, v
    at interior vertices as in the open tree's check; fourth-order
    differences of <w, step(u)> in the state and the shift damping."""
    nb, n, h = 2, 8, 1.0 / 8
    geoms = []
    for k in range(nb):
        for j in range(nb):
            for i in range(nb):
                g = MeshBlockGeom.cube(n, h, (i * n * h, j * n * h, k * n * h))
                f = 0
                for d, idx in enumerate((i, j, k)):
                    f |= (1 << (2 * d)) if idx == 0 else (2 << (2 * d))
                g.outflow_faces = f
                geoms.append(g)
    p = Params(**dict(PUNCTURE, eps_floor=-0.5, damp_kappa2=0.05))
    z = Z4c(geoms, p)
    z.set_exchange([0.0] * 3, [nb * n * h] * 3, (False, False, False))
    u_d = _smooth_state(z, geoms, p)
    dt = 0.25 * h
    rng = np.random.default_rng(43)

    def field(g):
        a = rng.standard_normal((22,) + g.shape)
        m = np.zeros(g.shape, bool)
        m[4:-4, 4:-4, 4:-4] = True
        return a * m

    w_d = z.upload([field(g) for g in geoms])
    v_d = z.upload([field(g) for g in geoms])
    ubar, pbar = z.rk4_step_ghosts_vjp(dt, u_d, w_d, 1, physical=True)
    work = z.make_step_work()

    def phi(uu, **kw):
        saved = z.params
        z.params = Params(**dict(p.as_dict(), **kw))
        out = uu.clone()
        z.rk4_step_fused(dt, out, work, exchange=True, physical=True)
        z.params = saved
        return float((out * w_d).sum())

    eps = 1e-6
    fd = (-phi(u_d + 2 * eps * v_d) + 8 * phi(u_d + eps * v_d) - 8 * phi(u_d - eps * v_d)
          + phi(u_d - 2 * eps * v_d)) / (12 * eps)
    an = float((ubar * v_d).sum())
    print(f"state: fd {fd:.12e}, adjoint {an:.12e}", flush=True)
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
    x = p.shift_eta
    hh = 1e-2 * max(1.0, abs(x))
    fd = (-phi(u_d, shift_eta=x + 2 * hh) + 8 * phi(u_d, shift_eta=x + hh)
          - 8 * phi(u_d, shift_eta=x - hh) + phi(u_d, shift_eta=x - 2 * hh)) / (12 * hh)
    an = float(pbar[0, 0])
    print(f"shift_eta: fd {fd:.12e}, adjoint {an:.12e}", flush=True)
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
    z.close()

