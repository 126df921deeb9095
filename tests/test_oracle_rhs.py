"""Pins the oracle's Z4cRHS, ADMConstraints, ADMToZ4c/Z4cToADM and AlgConstr
(CPU).  Reference: src/z4c/calculate_z4c_rhs.cpp:29-916,
src/z4c/adm_z4c.cpp:46-1120, src/z4c/z4c.cpp:677-991.

* 6th-order convergence of the oracle RHS and constraints to an independent
  continuum evaluation (tests/continuum.py) on closed-form fields that switch
  on every term (matter, all gauge options, both chi_psi_power branches);
* Minkowski gives an identically zero RHS;
* the 1-D gauge wave (an exact solution, src/z4c/awa_z4c.cpp:298-351) gives
  RHS = analytic time derivative to O(h^6);
* KO dissipation adds exactly sum_a Diss_a(u);
* ADM <-> Z4c round trip; AlgConstr invariants and the regularize_nth_root
  quirk (src/utils/floating_point.hpp:118-131).
"""
import numpy as np
import pytest
import sympy as sp

import continuum as ct
import oracle as orc

X0 = np.array([0.13, -0.21, 0.37])
FULL = dict(lapse_harmonic=0.3, lapse_K=0.5, shift_alpha2Gamma=0.4, shift_H=0.3,
            damp_kappa1=0.02, damp_kappa2=0.1, chi_div_floor=1e-5, diss=0.0)


@pytest.fixture(scope="module")
def state7():
    ex, mat = ct.random_state(7)
    return ex, mat, ct.ContinuumZ4c(ex, mat)


def _block_at(x0, h, nx=8):
    return orc.Geom.block(nx, h, x0 - (nx // 2) * h)


def _conv(errs):
    e0, e1 = errs
    ok = (e1 < 1e-8) | ((e0 / np.maximum(e1, 1e-300) > 40) & (e0 / np.maximum(e1, 1e-300) < 100))
    return ok


def test_rhs_converges_to_continuum(state7):
    ex, mat, cz = state7
    o = orc.Params(**FULL)
    exact = cz.rhs(X0, o)
    errs = []
    for h in (1 / 20, 1 / 40):
        g = _block_at(X0, h)
        u = ct.sample([ex[n] for n in ct.ORDER], g)
        m = ct.sample(mat, g)
        got = orc.rhs(o, g, u, m)[:, 8, 8, 8]
        errs.append(np.abs(got - exact))
    errs = np.array(errs)
    assert np.all(_conv(errs)), (errs[0] / errs[1])
    assert np.max(errs[1] / (np.abs(exact) + 1.0)) < 2e-6


def test_rhs_general_chi_power():
    """chi = psi^p with p = -2: the non-default oopsi4 / DA / Lchi branches."""
    ex, mat = ct.random_state(11, with_matter=False)
    cz = ct.ContinuumZ4c(ex, None, p=-2.0)
    o = orc.Params(**FULL)
    o.chi_psi_power = -2.0
    exact = cz.rhs(X0, o)
    errs = []
    for h in (1 / 20, 1 / 40):
        g = _block_at(X0, h)
        u = ct.sample([ex[n] for n in ct.ORDER], g)
        errs.append(np.abs(orc.rhs(o, g, u)[:, 8, 8, 8] - exact))
    errs = np.array(errs)
    assert np.all(_conv(errs)), (errs[0] / errs[1])


def test_constraints_converge_to_continuum(state7):
    ex, mat, cz = state7
    o = orc.Params(**FULL)
    H, Md, M = cz.constraints(X0)
    errs = []
    for h in (1 / 20, 1 / 40):
        g = _block_at(X0, h)
        u = ct.sample([ex[n] for n in ct.ORDER], g)
        m = ct.sample(mat, g)
        adm = orc.z4c_to_adm(o, g, u)
        con = orc.adm_constraints(o, g, u, adm, m)[:, 8, 8, 8]
        errs.append(np.abs(np.array([con[1], *con[4:7], con[2]]) - np.array([H, *Md, M])))
    errs = np.array(errs)
    assert np.all(_conv(errs)), errs
    # C = H^2 + M + Theta^2 + 4 Z, Z from Gamma~ - Gamma~(g~): check the Z part
    g = _block_at(X0, 1 / 40)
    u = ct.sample([ex[n] for n in ct.ORDER], g)
    adm = orc.z4c_to_adm(o, g, u)
    con = orc.adm_constraints(o, g, u, adm, ct.sample(mat, g))[:, 8, 8, 8]
    th = u[17, 8, 8, 8]
    Zc = (con[0] - con[1] ** 2 - con[2] - th * th) / 4.0
    np.testing.assert_allclose(Zc, con[3], rtol=1e-10)


def test_minkowski_rhs_is_zero():
    g = orc.Geom.block(8, 0.1, -0.4)
    u = orc.zeros(22, g)
    u[0] = 1.0
    for n in (1, 4, 6):
        u[n] = 1.0
    u[18] = 1.0
    for kw in (dict(), FULL, dict(lapse_harmonic=1.0, lapse_oplog=0.0, diss=0.02)):
        r = orc.rhs(orc.Params(**kw), g, u)
        assert np.all(r == 0.0)
    adm = orc.z4c_to_adm(orc.Params(), g, u)
    con = orc.adm_constraints(orc.Params(), g, u, adm)
    assert np.all(con == 0.0)


def _gauge_wave(A=0.01, d=1.0):
    x, t = sp.symbols("x t", real=True)
    H = 1 - A * sp.sin(2 * sp.pi * (x - t) / d)
    Kxx = -sp.diff(H, t) / (2 * sp.sqrt(H))
    chi = H ** sp.Rational(-1, 3)
    K = Kxx / H
    z = {"chi": chi, "gxx": H ** sp.Rational(2, 3), "gyy": H ** sp.Rational(-1, 3),
         "gzz": H ** sp.Rational(-1, 3), "Khat": K,
         "Axx": sp.Rational(2, 3) * chi * Kxx, "Ayy": -chi * Kxx / (3 * H),
         "Azz": -chi * Kxx / (3 * H), "Gamx": -sp.diff(H ** sp.Rational(-2, 3), x),
         "alpha": sp.sqrt(H)}
    return x, t, z


def test_gauge_wave_rhs_is_time_derivative():
    """Config-1 gauge (harmonic slicing f = alpha, zero shift): RHS = d_t."""
    x, t, z = _gauge_wave()
    o = orc.Params(lapse_oplog=0.0, lapse_harmonic=1.0, shift_Gamma=0.0, shift_eta=0.0,
                   shift_advect=0.0, damp_kappa1=0.02, chi_div_floor=1e-5, diss=0.0)
    x0 = 0.173
    errs = []
    for h in (1 / 32, 1 / 64):
        g = orc.Geom.block(8, h, np.array([x0, 0.0, 0.0]) - 4 * h)
        Xg, _, _ = g.mesh()
        u = orc.zeros(22, g)
        for n, name in enumerate(ct.ORDER):
            if name in z:
                u[n] = sp.lambdify(x, z[name].subs(t, 0), "numpy")(Xg)
        r = orc.rhs(o, g, u)[:, 8, 8, 8]
        exact = np.zeros(22)
        for n, name in enumerate(ct.ORDER):
            if name in z:
                exact[n] = float(sp.diff(z[name], t).subs({t: 0, x: x0}))
        errs.append(np.abs(r - exact))
    errs = np.array(errs)
    assert np.max(errs[1]) < 1e-8
    big = errs[0] > 1e-12
    assert np.all((errs[0][big] / errs[1][big]) > 40)


def test_dissipation_is_additive():
    rng = np.random.default_rng(5)
    g = orc.Geom.block(6, 0.05, 0.0)
    u = orc.zeros(22, g)
    u[0] = 1.0; u[1] = u[4] = u[6] = 1.0; u[18] = 1.0
    u += 1e-3 * rng.standard_normal(u.shape)
    r0 = orc.rhs(orc.Params(diss=0.0), g, u)
    r1 = orc.rhs(orc.Params(diss=0.02), g, u)
    k, j, i = 5, 6, 7
    p = i + g.nn1 * (j + g.nn2 * k)
    for n in (0, 9, 18, 21):
        kd = sum(0.02 * orc.fd_point(5, g, np.ascontiguousarray(u[n]), p, a) for a in range(3))
        np.testing.assert_allclose(r1[n, k, j, i] - r0[n, k, j, i], kd, rtol=1e-9, atol=1e-15)


def test_adm_z4c_round_trip():
    rng = np.random.default_rng(1)
    g = orc.Geom.block(6, 0.1, 0.0)
    adm = orc.zeros(17, g)
    adm[0] = 1.0 + 0.1 * rng.random(g.shape)
    adm[1:4] = 0.1 * rng.standard_normal((3,) + g.shape)
    for n, (a, b) in enumerate(ct.SYM):
        adm[4 + n] = (2.0 if a == b else 0.0) + 0.1 * rng.standard_normal(g.shape)
        adm[10 + n] = 0.1 * rng.standard_normal(g.shape)
    o = orc.Params(eps_floor=1e-12)
    u = orc.adm_to_z4c(o, g, adm)
    back = orc.z4c_to_adm(o, g, u)
    # det(gamma) = chi^-3 * det(g~) and det(g~)=1 after AlgConstr (detg>1 here)
    np.testing.assert_allclose(back[4:16], adm[4:16], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(back[0:4], adm[0:4], rtol=0, atol=0)
    np.testing.assert_allclose(u[17], 0.0)


def test_algconstr_invariants_and_quirk():
    rng = np.random.default_rng(2)
    g = orc.Geom.block(4, 0.1, 0.0)
    u = orc.zeros(22, g)
    for n, (a, b) in enumerate(ct.SYM):
        u[1 + n] = (1.3 if a == b else 0.0) + 0.05 * rng.standard_normal(g.shape)
        u[8 + n] = 0.1 * rng.standard_normal(g.shape)
    o = orc.Params()
    v = orc.algconstr(o, g, u.copy())
    gm = ct._sym(list(v[1:7]))
    det = np.einsum("ij...->...ij", gm)
    d = np.linalg.det(det)
    np.testing.assert_allclose(d, 1.0, rtol=1e-14)
    A = ct._sym(list(v[8:14]))
    gi = np.linalg.inv(det)
    tr = np.einsum("...ij,ij...->...", gi, A)
    assert np.max(np.abs(tr)) < 1e-15
    # idempotent up to the regularisation of the second pass: det is now
    # 1 (+/- roundoff), so eps < eps_floor and g~ *= (1 - eps_floor/3)
    w = orc.algconstr(o, g, v.copy())
    np.testing.assert_allclose(w[1:7], v[1:7] * (1.0 + 1e-12 / -3.0), rtol=1e-15)
    np.testing.assert_allclose(w[8:14], v[8:14], atol=1e-15)
    # det(g~) < 1 is not projected: g~ only scaled by (1 - eps_floor/3)
    u2 = orc.zeros(22, g)
    u2[1] = u2[4] = u2[6] = 0.9
    w2 = orc.algconstr(o, g, u2.copy())
    np.testing.assert_array_equal(w2[1], 0.9 * (1.0 + 1e-12 / -3.0))
