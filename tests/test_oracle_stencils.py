"""Pins the oracle's finite-difference operators (CPU).

Reference: src/utils/finite_differencing.tpp (DBG_SYMMETRIZE_FD branches,
enabled by src/defs.hpp.in:18) with the NGHOST=4 stencils of
src/utils/finite_differencing_stencils.tpp.

The weights the oracle applies are extracted by probing it with unit impulses
and compared with weights generated independently by Fornberg's algorithm in
exact rational arithmetic; KO dissipation is compared with the binomial 8th
difference.  Polynomial exactness and convergence order follow.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as orc


def fornberg(x0, xs, m):
    """Exact weights for the m-th derivative at x0 on nodes xs (Fornberg 1988)."""
    n = len(xs)
    c = [[[Fraction(0)] * n for _ in range(n)] for _ in range(m + 1)]
    c[0][0][0] = Fraction(1)
    c1 = Fraction(1)
    c4 = xs[0] - x0
    for i in range(1, n):
        mn = min(i, m)
        c2 = Fraction(1)
        c5 = c4
        c4 = xs[i] - x0
        for j in range(i):
            c3 = xs[i] - xs[j]
            c2 *= c3
            if j == i - 1:
                for k in range(mn, 0, -1):
                    c[k][i][i] = c1 * (k * c[k - 1][i - 1][i - 1] - c5 * c[k][i - 1][i - 1]) / c2
                c[0][i][i] = -c1 * c5 * c[0][i - 1][i - 1] / c2
            for k in range(mn, 0, -1):
                c[k][i][j] = (c4 * c[k][i - 1][j] - k * c[k - 1][i - 1][j]) / c3
            c[0][i][j] = c4 * c[0][i - 1][j] / c3
        c1 = c2
    return [c[m][n - 1][j] for j in range(n)]


def probe_weights(kind, offsets, d=0, h=1.0):
    """Apply oracle stencil `kind` at the block centre to unit impulses."""
    g = orc.Geom.block(12, h, 0.0)
    u = np.zeros(g.shape)
    c = [6 + 4, 6 + 4, 6 + 4]  # centre vertex (k, j, i)
    stride = [1, g.nn1, g.nn1 * g.nn2][d]
    p = c[2] + g.nn1 * (c[1] + g.nn2 * c[0])
    w = []
    flat = u.reshape(-1)
    for off in offsets:
        flat[:] = 0.0
        flat[p + off * stride] = 1.0
        w.append(orc.fd_point(kind, g, u, p, d))
    return np.array(w)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_centered_first_and_second(d):
    offs = list(range(-3, 4))
    xs = [Fraction(o) for o in offs]
    w1 = fornberg(Fraction(0), xs, 1)
    w2 = fornberg(Fraction(0), xs, 2)
    np.testing.assert_allclose(probe_weights(0, offs, d), [float(x) for x in w1], rtol=1e-15, atol=1e-16)
    np.testing.assert_allclose(probe_weights(1, offs, d), [float(x) for x in w2], rtol=1e-15, atol=1e-15)
    # the 1st-derivative stencil is antisymmetric with zero centre weight
    assert w1[3] == 0 and all(w1[i] == -w1[6 - i] for i in range(7))
    # exact rationals as printed in the reference's tables (LCM-factored)
    assert [x * 60 for x in w1[:3]] == [-1, 9, -45]
    assert [x * 180 for x in w2[:4]] == [2, -27, 270, -490]


@pytest.mark.parametrize("d", [0, 1, 2])
def test_lopsided(d):
    offs_l = list(range(-4, 3))
    offs_r = list(range(-2, 5))
    wl = fornberg(Fraction(0), [Fraction(o) for o in offs_l], 1)
    wr = fornberg(Fraction(0), [Fraction(o) for o in offs_r], 1)
    np.testing.assert_allclose(probe_weights(3, offs_l, d), [float(x) for x in wl], rtol=1e-15, atol=1e-16)
    np.testing.assert_allclose(probe_weights(4, offs_r, d), [float(x) for x in wr], rtol=1e-15, atol=1e-16)
    assert [x * 60 for x in wl] == [1, -8, 30, -80, 35, 24, -2]
    assert [x * 60 for x in wr] == [2, -24, -35, 80, -30, 8, -1]


@pytest.mark.parametrize("d", [0, 1, 2])
def test_ds_and_dissipation(d):
    wd = probe_weights(2, [-1, 0, 1], d)
    np.testing.assert_array_equal(wd, [-0.5, 0.0, 0.5])
    offs = list(range(-4, 5))
    wk = probe_weights(5, offs, d)
    # Diss = -2^-8 * (8th binomial difference) / h, with diss = 1
    binom = np.array([1, -8, 28, -56, 70, -56, 28, -8, 1], float)
    np.testing.assert_array_equal(wk, -binom / 256.0)


def test_spacing_scaling():
    h = 0.37
    offs = list(range(-3, 4))
    np.testing.assert_allclose(probe_weights(0, offs, 1, h) * h, probe_weights(0, offs, 1, 1.0), rtol=2e-15, atol=1e-16)
    np.testing.assert_allclose(probe_weights(1, offs, 2, h) * h * h, probe_weights(1, offs, 2, 1.0), rtol=2e-15, atol=1e-14)
    # Diss scales as 1/h (idx, not idx^8: finite_differencing.tpp:48)
    np.testing.assert_allclose(probe_weights(5, list(range(-4, 5)), 0, h) * h,
                               probe_weights(5, list(range(-4, 5)), 0, 1.0), rtol=2e-15)


def test_lx_upwind_select():
    """Lx uses the left-biased stencil for v<0, right-biased otherwise."""
    g = orc.Geom.block(12, 0.1, 0.0)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(g.shape)
    p = 10 + g.nn1 * (10 + g.nn2 * 10)
    for d in range(3):
        dl = orc.fd_point(3, g, u, p, d)
        dr = orc.fd_point(4, g, u, p, d)
        lib = orc.lib()
        for v in (-0.7, 0.0, 0.3):
            got = lib.oracle_lx_point(orc.C.byref(g), orc._p(u), p, d, v)
            exp = v * (dl if v < 0 else dr)
            np.testing.assert_allclose(got, exp, rtol=1e-14, atol=1e-300)


def test_polynomial_exactness_and_order():
    g = orc.Geom.block(16, 1.0, 0.0)
    X, Y, Z = g.mesh()
    p = 12 + g.nn1 * (12 + g.nn2 * 12)
    x0 = X.reshape(-1)[p]
    for deg in range(0, 8):
        u = np.ascontiguousarray((X - x0) ** deg)
        d1 = orc.fd_point(0, g, u, p, 0)
        d2 = orc.fd_point(1, g, u, p, 0)
        exp1 = 1.0 if deg == 1 else 0.0
        exp2 = 2.0 if deg == 2 else 0.0
        if deg <= 6:
            assert abs(d1 - exp1) < 1e-9
        if deg <= 7:
            assert abs(d2 - exp2) < 1e-8
        dk = orc.fd_point(5, g, u, p, 0)
        assert abs(dk) < 1e-7  # KO annihilates degree <= 7
    # 6th-order convergence on a sine
    errs = []
    for n in (16, 32):
        h = 1.0 / n
        gg = orc.Geom.block(n, h, 0.0)
        X, _, _ = gg.mesh()
        u = np.ascontiguousarray(np.sin(2 * np.pi * X))
        q = (4 + n // 2) + gg.nn1 * (8 + gg.nn2 * 8)
        x = X.reshape(-1)[q]
        errs.append(abs(orc.fd_point(0, gg, u, q, 0) - 2 * np.pi * np.cos(2 * np.pi * x)))
    assert 50 < errs[0] / errs[1] < 80


def test_lx_selection_is_upwind_by_stability():
    """Pins Lx's selection rule by a property, not by restating it: for the
    advection the right-hand side carries, du/dt = v Lx(v, u), the
    semi-discrete operator on a periodic line must not amplify any Fourier
    mode (real part of its symbol <= 0 for every wavenumber, for either
    sign of v), which holds only if the stencil's extra points lie upwind.
    The opposite choice is checked to amplify, so a swapped selection fails."""
    g = orc.Geom.block(12, 1.0, 0.0)
    p = 10 + g.nn1 * (10 + g.nn2 * 10)
    lib = orc.lib()
    offs = np.arange(-4, 5)
    theta = np.linspace(0.0, np.pi, 721)
    for d in range(3):
        stride = [1, g.nn1, g.nn1 * g.nn2][d]

        def symbol_re(v):
            w = []
            u = np.zeros(g.shape)
            flat = u.reshape(-1)
            for m in offs:
                flat[:] = 0.0
                flat[p + m * stride] = 1.0
                w.append(lib.oracle_lx_point(orc.C.byref(g), orc._p(u), p, d, v))
            return (np.array(w)[None, :] * np.cos(np.outer(theta, offs))).sum(1)

        for v in (-0.7, 0.3):
            re = symbol_re(v)
            assert re.max() <= 1e-13, (d, v, re.max())
            assert re.min() < -1e-3  # dissipative, not neutral
            # the other one-sided stencil at the same v (the one the oracle
            # selects for -v): v D_other = -Lx(-v, u); it amplifies
            re_dn = -symbol_re(-v)
            assert re_dn.max() > 1e-3
