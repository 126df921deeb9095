"""Pins the oracle's Z4cWeyl (psi4, src/z4c/calculate_weyl_scalars.cpp:44-628)
on exact vacuum spacetimes (CPU):

* Minkowski: psi4 = 0 exactly;
* flat spacetime on the curved slice t = f(x, y, z): gamma_ij = delta_ij -
  f_i f_j and K_ij = -N f_ij (N = (1 - |grad f|^2)^-1/2) are O(1), every
  term of the Gauss / Codazzi projection is O(1), and psi4 must cancel to
  zero: it converges to 0 at 6th order;
* static Schwarzschild in isotropic coordinates, centred away from the
  origin of the coordinate tetrad (so psi4 != 0): psi4 equals
  -E_ab mbar^a mbar^b with mbar = (v - i w)/sqrt(2) of the same tetrad and
  the electric Weyl tensor computed a different way, from the static lapse,
  E_ab = D_a D_b N / N (a static vacuum identity), to 1e-8.  This fixes
  the sign convention and every relative weight of the Ricci and Gauss terms.
"""
import numpy as np
import sympy as sp

import oracle as orc
from z4c_amd import Params
from z4c_amd.z4c import MeshBlockGeom

X, Y, Z = sp.symbols("x y z", real=True)
XYZ = (X, Y, Z)
SYM = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
P = Params(chi_div_floor=1e-5)
OP = orc.Params(**P.as_dict())


def _grid(n, h, xmin):
    g = MeshBlockGeom.cube(n, h, xmin)
    og = orc.Geom.block(g.nx, g.dx, g.xmin)
    x = [og.x0[d] + np.arange(og.nn1) * og.dx[d] for d in range(3)]
    Zg, Yg, Xg = np.meshgrid(x[2], x[1], x[0], indexing="ij")
    return og, x, (Xg, Yg, Zg)


def _psi4(og, adm):
    u = orc.adm_to_z4c(OP, og, np.ascontiguousarray(adm))
    return orc.weyl(OP, og, u, adm)


def test_minkowski_is_zero():
    og, _, _ = _grid(8, 0.125, (0.1, -0.2, 0.05))
    u = np.zeros((22,) + og.shape)
    u[0] = u[1] = u[4] = u[6] = u[18] = 1.0
    w = orc.weyl(OP, og, u)
    inner = w[:, og.kl:og.ku + 1, og.jl:og.ju + 1, og.il:og.iu + 1]
    assert np.all(inner == 0.0)
    assert np.isnan(w[0, 0, 0, 0])


def test_flat_spacetime_on_curved_slice_converges_to_zero():
    A, sig, c = 0.3, 0.7, (0.2, -0.1, 0.15)
    f = A * sp.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / sig ** 2)
    fi = [sp.diff(f, v) for v in XYZ]
    Nn = 1 / sp.sqrt(1 - sum(q * q for q in fi))
    gam = [sp.lambdify(XYZ, (1 if a == b else 0) - fi[a] * fi[b], "numpy") for a, b in SYM]
    K = [sp.lambdify(XYZ, -Nn * sp.diff(f, XYZ[a], XYZ[b]), "numpy") for a, b in SYM]
    errs = []
    for n in (16, 32):
        og, _, (Xg, Yg, Zg) = _grid(n, 1.0 / n, (-0.3, -0.4, -0.35))
        adm = np.zeros((17,) + og.shape)
        adm[0] = 1.0
        for q in range(6):
            adm[4 + q] = gam[q](Xg, Yg, Zg)
            adm[10 + q] = K[q](Xg, Yg, Zg)
        assert np.abs(adm[10:16]).max() > 1.0  # the K terms are O(1)
        errs.append(np.nanmax(np.abs(_psi4(og, adm))))
    assert errs[0] < 1e-4 and 40 < errs[0] / errs[1] < 100, errs


def test_schwarzschild_off_centre_matches_electric_weyl():
    M, c = 1.0, (0.9, -0.4, 0.3)
    r = sp.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
    psi = 1 + M / (2 * r)
    Nl = (1 - M / (2 * r)) / (1 + M / (2 * r))
    g = sp.eye(3) * psi ** 4
    gi = g.inv()
    Gam = [[[sum(gi[k, l] * (sp.diff(g[l, i], XYZ[j]) + sp.diff(g[l, j], XYZ[i])
                             - sp.diff(g[i, j], XYZ[l])) / 2 for l in range(3))
             for j in range(3)] for i in range(3)] for k in range(3)]
    E = sp.Matrix(3, 3, lambda i, j: (sp.diff(Nl, XYZ[i], XYZ[j])
                                      - sum(Gam[k][i][j] * sp.diff(Nl, XYZ[k]) for k in range(3))) / Nl)
    Ef = sp.lambdify(XYZ, E, "numpy")
    psi4f = sp.lambdify(XYZ, psi ** 4, "numpy")

    def expected(xx, yy, zz):
        G = np.eye(3) * psi4f(xx, yy, zz)
        x1 = xx + (1e-8 if xx * xx + yy * yy < 1e-10 else 0.0)
        u = np.array([x1, yy, zz])
        v = np.array([x1 * zz, yy * zz, -x1 * x1 - yy * yy])
        w = np.array([-yy, x1, 0.0])
        w = w / np.sqrt(w @ G @ w)
        u = u - (w @ G @ u) * w
        u = u / np.sqrt(u @ G @ u)
        v = v - (w @ G @ v) * w - (u @ G @ v) * u
        v = v / np.sqrt(v @ G @ v)
        mb = (v - 1j * w) / np.sqrt(2)
        return -(mb @ np.array(Ef(xx, yy, zz), dtype=float) @ mb)

    errs = []
    for n in (16, 32):
        og, x, (Xg, Yg, Zg) = _grid(n, 0.6 / n, (2.0, 1.5, 1.0))
        rr = np.sqrt((Xg - c[0]) ** 2 + (Yg - c[1]) ** 2 + (Zg - c[2]) ** 2)
        adm = np.zeros((17,) + og.shape)
        adm[4] = adm[7] = adm[9] = adm[16] = (1 + M / (2 * rr)) ** 4
        adm[0] = 1.0
        w = _psi4(og, adm)
        e = 0.0
        for (i, j, k) in [(og.il + 2, og.jl + 3, og.kl + 1), (og.iu - 1, og.ju - 2, og.ku - 3),
                          ((og.il + og.iu) // 2,) * 3]:
            ex = expected(x[0][i], x[1][j], x[2][k])
            assert abs(ex) > 1e-3
            e = max(e, abs(w[0, k, j, i] + 1j * w[1, k, j, i] - ex) / abs(ex))
        errs.append(e)
    # truncation (h^6) is below the rounding of the second derivatives
    # (~1e-16/h^2 relative to psi4 ~ 1e-3) already at n = 16: an O(1)
    # mistake in any term would show as an O(1) relative error
    assert max(errs) < 1e-8, errs
