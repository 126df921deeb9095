"""GPU adjoints (z4c_rhs_vjp, z4c_rk4_step_vjp, the autograd bindings) against
torch.autograd of the float64 restatement (oracle/z4c_torch.py, itself pinned
to the C oracle by tests/test_oracle_torch.py) and against central finite
differences of the GPU forward path.

Bounds: per-field max|gpu - ref| / max|ref| <= 1e-10 against autograd (the
two evaluate the same derivative in different operation orders; second-
derivative weights ~1/dx^2 amplify rounding); finite differences with
eps = 1e-6 agree to 1e-6 relative.
"""
import numpy as np
import pytest
import torch

import oracle as orc
import z4c_torch as zt
from z4c_amd import Params, adjoint
from z4c_amd._capi import PBAR, Z4cError
from z4c_amd.initial_data import perturbed_flat
from z4c_amd.z4c import MeshBlockGeom, Z4c

pytestmark = pytest.mark.gpu

FULL = dict(lapse_harmonic=0.3, lapse_K=0.5, shift_alpha2Gamma=0.4, shift_H=0.3,
            damp_kappa1=0.02, damp_kappa2=0.1, chi_div_floor=1e-5, diss=0.02)
PUNCTURE = dict(lapse_oplog=2.0, shift_Gamma=1.0, shift_eta=2.0, damp_kappa1=0.02,
                diss=0.02, chi_div_floor=1e-5)


def _state(g, p, seed, amp=(1e-2, 1e-1), shift=None):
    """Perturbed flat state after AlgConstr; ``shift`` adds a constant to
    beta^i, which keeps the upwind selection of the advective stencils
    (sign of beta^i) fixed under finite-difference perturbations: the
    advective term beta^a Lsel_a(u) is continuous but not differentiable
    where beta^a changes sign."""
    op = orc.Params(**p.as_dict())
    og = orc.Geom.block(g.nx, g.dx, g.xmin)
    u = np.ascontiguousarray(perturbed_flat(g, seed=seed, amp=amp))
    if shift is not None:
        for a in range(3):
            u[19 + a] += shift[a]
    return og, orc.algconstr(op, og, u)


def _regs(z, arrays):
    """nw single-block arrays -> nw consecutive registers."""
    return torch.cat([z.upload([a]) for a in arrays])


def _pdict(p):
    d = p.as_dict()
    pt = {k: torch.tensor(d[k], dtype=torch.float64, requires_grad=True) for k in PBAR}
    d.update(pt)
    return d, [pt[k] for k in PBAR]


def _ref_rhs_vjp(og, p, u, w):
    ut = torch.from_numpy(u).requires_grad_(True)
    d, pts = _pdict(p)
    s = (zt.rhs(ut, og, d) * torch.from_numpy(w)).sum()
    gr = torch.autograd.grad(s, [ut] + pts)
    return gr[0].numpy(), np.array([x.item() for x in gr[1:]])


def _ref_step_vjp(og, p, u, w, dt):
    ut = torch.from_numpy(u).requires_grad_(True)
    d, pts = _pdict(p)
    s = (zt.rk4_step(ut, og, d, dt) * torch.from_numpy(w)).sum()
    gr = torch.autograd.grad(s, [ut] + pts)
    return gr[0].numpy(), np.array([x.item() for x in gr[1:]])


def _rel(got, ref):
    num = np.abs(got - ref).reshape(22, -1).max(axis=1)
    den = np.maximum(np.abs(ref).reshape(22, -1).max(axis=1), 1e-300)
    return num / den


@pytest.mark.parametrize("kw", [FULL, dict(FULL, chi_psi_power=-2.0), PUNCTURE])
def test_rhs_vjp_matches_autograd(kw):
    g = MeshBlockGeom.cube(10, 1.0 / 10, (0.1, -0.2, 0.05))
    p = Params(**kw)
    og, u = _state(g, p, seed=3)
    rng = np.random.default_rng(5)
    nw = 3
    w = rng.standard_normal((nw,) + u.shape)
    z = Z4c([g], p)
    ubar, pbar = z.rhs_vjp(z.upload([u]), _regs(z, w), nw)
    torch.cuda.synchronize()
    got = ubar.cpu().numpy().reshape((nw,) + u.shape)
    pb = pbar.cpu().numpy()
    for c in range(nw):
        ref, pref = _ref_rhs_vjp(og, p, u, w[c])
        assert np.all(_rel(got[c], ref) <= 1e-10), (c, _rel(got[c], ref))
        assert np.allclose(pb[c], pref, rtol=1e-10, atol=1e-12 * np.abs(pref).max()), (pb[c], pref)


def test_rhs_vjp_uneven_batch_and_batching():
    """Two blocks of different size and spacing in one context: each block's
    VJP is its own (blocks are independent); nw cotangents in one call equal
    nw single calls bit for bit."""
    gs = [MeshBlockGeom.cube(8, 1.0 / 8, (0.1, -0.2, 0.05)),
          MeshBlockGeom.cube(12, 1.0 / 16, (-0.4, 0.3, 0.2))]
    p = Params(**FULL)
    states, ogs = [], []
    for b, g in enumerate(gs):
        og, u = _state(g, p, seed=11 + b)
        states.append(u)
        ogs.append(og)
    rng = np.random.default_rng(7)
    nw = 2
    ws = [[rng.standard_normal(u.shape) for u in states] for _ in range(nw)]
    z = Z4c(gs, p)
    u_d = z.upload(states)
    w_d = torch.cat([z.upload(ws[c]) for c in range(nw)])
    ubar, pbar = z.rhs_vjp(u_d, w_d, nw)
    torch.cuda.synchronize()
    reg = 22 * z.total_points
    for c in range(nw):
        got = z.download(ubar[c * reg:(c + 1) * reg], 22)
        ptot = np.zeros(3)
        for b in range(2):
            ref, pref = _ref_rhs_vjp(ogs[b], p, states[b], ws[c][b])
            assert np.all(_rel(got[b], ref) <= 1e-10), (c, b, _rel(got[b], ref))
            ptot += pref
        assert np.allclose(pbar[c].cpu().numpy(), ptot, rtol=1e-10)
        one, p1 = z.rhs_vjp(u_d, w_d[c * reg:(c + 1) * reg], 1)
        assert torch.equal(one, ubar[c * reg:(c + 1) * reg])
        assert torch.equal(p1[0], pbar[c])


def test_rhs_vjp_finite_differences():
    """<ubar, v> = d/de <w, rhs(u + e v)> by central differences (e = 1e-6) of
    the GPU forward right-hand side.  The right-hand side is affine in each
    of shift_eta, damp_kappa1, damp_kappa2, so their differences use a wide
    step (exact up to rounding)."""
    g = MeshBlockGeom.cube(12, 1.0 / 12, (0.1, -0.2, 0.05))
    p = Params(**PUNCTURE)
    og, u = _state(g, p, seed=21, shift=(0.3, -0.25, 0.2))
    rng = np.random.default_rng(2)
    w = rng.standard_normal(u.shape)
    v = rng.standard_normal(u.shape)
    z = Z4c([g], p)
    u_d, w_d, v_d = z.upload([u]), z.upload([w]), z.upload([v])
    ubar, pbar = z.rhs_vjp(u_d, w_d, 1)
    eps = 1e-6

    def phi(uu, **kw):
        q = Params(**dict(p.as_dict(), **kw))
        r = z.zeros()
        saved = z.params
        z.params = q
        z.Z4cRHS(uu, None, r)
        z.params = saved
        return float((r * w_d).sum())

    fd = (phi(u_d + eps * v_d) - phi(u_d - eps * v_d)) / (2 * eps)
    an = float((ubar * v_d).sum())
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
    for t, name in enumerate(PBAR):
        x = getattr(p, name)
        h = 0.01
        fd = (phi(u_d, **{name: x + h}) - phi(u_d, **{name: x - h})) / (2 * h)
        an = float(pbar[0, t])
        assert abs(fd - an) <= 1e-6 * abs(an), (name, fd, an)


def test_step_vjp_matches_autograd():
    g = MeshBlockGeom.cube(10, 1.0 / 10, (0.1, -0.2, 0.05))
    p = Params(**PUNCTURE)
    og, u = _state(g, p, seed=4, amp=(1e-3, 3e-2))
    dt = 0.25 / 10
    rng = np.random.default_rng(9)
    nw = 2
    w = rng.standard_normal((nw,) + u.shape)
    z = Z4c([g], p)
    ubar, pbar = z.rk4_step_vjp(dt, z.upload([u]), _regs(z, w), nw)
    torch.cuda.synchronize()
    got = ubar.cpu().numpy().reshape((nw,) + u.shape)
    for c in range(nw):
        ref, pref = _ref_step_vjp(og, p, u, w[c], dt)
        assert np.all(_rel(got[c], ref) <= 1e-9), (c, _rel(got[c], ref))
        assert np.allclose(pbar[c].cpu().numpy(), pref, rtol=1e-9, atol=1e-12 * np.abs(pref).max())


def test_step_vjp_several_blocks():
    """Two blocks of different sizes and two cotangents: every block's
    adjoint reads the stage right-hand sides at its own offset (the second
    block's AlgConstr linearisation point is u_n + a dt k of that block)."""
    g0 = MeshBlockGeom.cube(8, 1.0 / 10, (-0.6, 0.1, 0.2))
    g1 = MeshBlockGeom.cube(10, 1.0 / 10, (0.1, -0.2, 0.05))
    p = Params(**PUNCTURE)
    og0, u0 = _state(g0, p, seed=11, amp=(1e-3, 3e-2))
    og1, u1 = _state(g1, p, seed=4, amp=(1e-3, 3e-2))
    dt = 0.25 / 10
    rng = np.random.default_rng(5)
    nw = 2
    w0 = rng.standard_normal((nw,) + u0.shape)
    w1 = rng.standard_normal((nw,) + u1.shape)
    z = Z4c([g0, g1], p)
    w_d = torch.cat([z.upload([w0[c], w1[c]]) for c in range(nw)])
    ubar, pbar = z.rk4_step_vjp(dt, z.upload([u0, u1]), w_d, nw)
    torch.cuda.synchronize()
    reg = ubar.numel() // nw
    for c in range(nw):
        got = ubar[c * reg:(c + 1) * reg]
        p_ref = 0.0
        for b, (og, u, w) in enumerate(((og0, u0, w0[c]), (og1, u1, w1[c]))):
            ref, pref = _ref_step_vjp(og, p, u, w, dt)
            blk = z.block_view(got, b, 22).cpu().numpy()
            assert np.all(_rel(blk, ref) <= 1e-9), (c, b, _rel(blk, ref))
            p_ref = p_ref + pref
        assert np.allclose(pbar[c].cpu().numpy(), p_ref, rtol=1e-9,
                           atol=1e-12 * np.abs(p_ref).max())


def test_step_vjp_repeated_calls_on_one_context():
    """A second step VJP on the same context (its scratch reused) equals the
    same call on a fresh context, bit for bit."""
    g = MeshBlockGeom.cube(8, 1.0 / 8, (0.1, -0.2, 0.05))
    p = Params(**PUNCTURE)
    _, u1 = _state(g, p, seed=12, amp=(1e-3, 3e-2))
    _, u2 = _state(g, p, seed=13, amp=(1e-3, 3e-2))
    rng = np.random.default_rng(6)
    w1, w2 = rng.standard_normal(u1.shape), rng.standard_normal(u2.shape)
    dt = 0.25 / 8
    z = Z4c([g], p)
    z.rk4_step_vjp(dt, z.upload([u1]), z.upload([w1]), 1)
    got, pgot = z.rk4_step_vjp(dt, z.upload([u2]), z.upload([w2]), 1)
    z2 = Z4c([g], p)
    exp, pexp = z2.rk4_step_vjp(dt, z2.upload([u2]), z2.upload([w2]), 1)
    torch.cuda.synchronize()
    assert torch.equal(got, exp)
    assert torch.equal(pgot, pexp)


def test_step_vjp_finite_differences():
    """Central differences of the GPU fused step (e = 1e-6).  eps_floor is set
    far below zero so that AlgConstr's regularize_nth_root takes its pow
    branch at every vertex: the reference's constant branch for
    det g~ < 1 + eps_floor has a derivative jump at the switch, and every
    vertex that AlgConstr has already projected (det g~ = 1 to rounding, all
    ghost vertices of a step) sits on it, so a finite difference straddles
    it there.  The adjoint differentiates the branch taken."""
    g = MeshBlockGeom.cube(10, 1.0 / 10, (0.1, -0.2, 0.05))
    p = Params(**dict(PUNCTURE, eps_floor=-1e30, damp_kappa2=0.05))
    og, u = _state(g, p, seed=6, amp=(1e-3, 3e-2), shift=(0.3, -0.25, 0.2))
    dt = 0.25 / 10
    rng = np.random.default_rng(3)
    w = rng.standard_normal(u.shape)
    v = rng.standard_normal(u.shape)
    z = Z4c([g], p)
    u_d, w_d, v_d = z.upload([u]), z.upload([w]), z.upload([v])
    ubar, pbar = z.rk4_step_vjp(dt, u_d, w_d, 1)
    work = z.make_step_work()
    eps = 1e-6

    def phi(uu, **kw):
        q = Params(**dict(p.as_dict(), **kw))
        out = uu.clone()
        saved = z.params
        z.params = q
        z.rk4_step_fused(dt, out, work)
        z.params = saved
        return float((out * w_d).sum())

    fd = (phi(u_d + eps * v_d) - phi(u_d - eps * v_d)) / (2 * eps)
    an = float((ubar * v_d).sum())
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
    # the step is a polynomial in the parameters: 4th-order differences
    for t, name in enumerate(PBAR):
        x = getattr(p, name)
        h = 1e-3 * max(1.0, abs(x))
        fd = (-phi(u_d, **{name: x + 2 * h}) + 8 * phi(u_d, **{name: x + h})
              - 8 * phi(u_d, **{name: x - h}) + phi(u_d, **{name: x - 2 * h})) / (12 * h)
        an = float(pbar[0, t])
        assert abs(fd - an) <= 1e-6 * abs(an), (name, fd, an)


def test_autograd_functions():
    """Z4cRHSFunction / Z4cStepFunction backward passes are the C-ABI VJPs and
    carry the gauge-parameter gradients."""
    g = MeshBlockGeom.cube(8, 1.0 / 8, (0.1, -0.2, 0.05))
    p = Params(**PUNCTURE)
    og, u = _state(g, p, seed=8, amp=(1e-3, 3e-2))
    rng = np.random.default_rng(4)
    w = rng.standard_normal(u.shape)
    z = Z4c([g], p)
    u_d, w_d = z.upload([u]), z.upload([w])
    ut = u_d.clone().requires_grad_(True)
    eta = torch.tensor(p.shift_eta, dtype=torch.float64, device=u_d.device, requires_grad=True)
    k1 = torch.tensor(p.damp_kappa1, dtype=torch.float64, device=u_d.device, requires_grad=True)
    (adjoint.rhs(z, ut, eta=eta, kappa1=k1) * w_d).sum().backward()
    ref, pref = _ref_rhs_vjp(og, p, u, w)
    assert np.all(_rel(z.download(ut.grad, 22)[0], ref) <= 1e-10)
    assert np.isclose(eta.grad.item(), pref[0], rtol=1e-10)
    assert np.isclose(k1.grad.item(), pref[1], rtol=1e-10)
    dt = 0.25 / 8
    ut2 = u_d.clone().requires_grad_(True)
    eta2 = torch.tensor(p.shift_eta, dtype=torch.float64, device=u_d.device, requires_grad=True)
    (adjoint.rk4_step(z, ut2, dt, eta=eta2) * w_d).sum().backward()
    ref, pref = _ref_step_vjp(og, p, u, w, dt)
    assert np.all(_rel(z.download(ut2.grad, 22)[0], ref) <= 1e-9)
    assert np.isclose(eta2.grad.item(), pref[0], rtol=1e-9)


def test_vjp_rejects_unsupported():
    g = MeshBlockGeom.cube(8, 1.0 / 8, (0.1, -0.2, 0.05), faces=63)
    z = Z4c([g], Params(**PUNCTURE))
    u = z.zeros()
    with pytest.raises(Z4cError):
        z.rk4_step_vjp(0.01, u, z.zeros(), 1)
    with pytest.raises(Z4cError):
        z.rhs_vjp(u, z.zeros(), 0)
    with pytest.raises(Z4cError):
        z.rhs_vjp(u, z.zeros(), 1, params=Params(cowling=1))


def _ref_hnorm(og, p, u):
    ut = torch.from_numpy(u).requires_grad_(True)
    J = zt.hamiltonian_norm(ut, og, p.as_dict())
    (g,) = torch.autograd.grad(J, [ut])
    return J.item(), g.numpy()


def test_hamiltonian_norm_and_vjp():
    g = MeshBlockGeom.cube(10, 1.0 / 10, (0.1, -0.2, 0.05))
    p = Params(**PUNCTURE)
    og, u = _state(g, p, seed=12)
    z = Z4c([g], p)
    u_d = z.upload([u])
    J = z.hamiltonian_norm(u_d)
    ubar = z.hamiltonian_norm_vjp(u_d)
    Jr, gr = _ref_hnorm(og, p, u)
    assert abs(J - Jr) <= 1e-11 * abs(Jr), (J, Jr)
    assert np.all(_rel(z.download(ubar, 22)[0], gr) <= 1e-9), _rel(z.download(ubar, 22)[0], gr)


def test_hamiltonian_sensitivity_through_steps():
    """J after 3 steps, adjoint through the last 2, against torch.autograd of
    the restatement over the same 2 steps from the same state."""
    from z4c_amd.sensitivity import hamiltonian_sensitivity
    g = MeshBlockGeom.cube(8, 1.0 / 8, (0.1, -0.2, 0.05))
    p = Params(**dict(PUNCTURE, damp_kappa2=0.05))
    og, u = _state(g, p, seed=13, amp=(1e-3, 3e-2))
    dt = 0.25 / 8
    z = Z4c([g], p)
    u_d = z.upload([u])
    out = hamiltonian_sensitivity(z, u_d, dt, nsteps=3, window=2, graph=False)
    w = z.upload([u])
    z.rk4_step_fused(dt, w, z.make_step_work())
    start = z.download(w, 22)[0]
    ut = torch.from_numpy(start).requires_grad_(True)
    d, pts = _pdict(p)
    x = ut
    for _ in range(2):
        x = zt.rk4_step(x, og, d, dt)
    Jr = zt.hamiltonian_norm(x, og, d)
    gr = torch.autograd.grad(Jr, [ut] + pts)
    assert abs(out["J"] - Jr.item()) <= 1e-9 * abs(Jr.item())
    assert np.all(_rel(z.download(out["ubar"], 22)[0], gr[0].numpy()) <= 1e-8)
    for t, name in enumerate(PBAR):
        assert np.isclose(out["pbar"][name], gr[1 + t].item(), rtol=1e-8, atol=1e-14), name
