"""Differentiable CPU restatement of the Z4c right-hand side (test
infrastructure: the checker of the adjoint kernels, never the product).

A float64 ``torch`` (CPU) restatement of the vacuum ``Z4c::Z4cRHS``
(src/z4c/calculate_z4c_rhs.cpp:29-916), of ``Z4c::AlgConstr``
(src/z4c/z4c.cpp:677-799) and of one ``bt_rk4`` step with ghosts held at
u_n (src/task_list/task_list_integrators_butcher.cpp:37-49,542-680), written
over whole tensors so that ``torch.autograd`` differentiates it.  Its
autograd vector-Jacobian products are the reference gradients of the GPU
adjoint; its forward values are pinned against the C oracle (which follows
the reference operation by operation) by the CPU tests.

Differences from the C oracle are in operation order only: the finite
differences are tensor sums over shifted views, and the mixed second
derivatives are the same nested first derivatives Dx_a(Dx_b u), a < b
(calculate_z4c_rhs.cpp:132-175), evaluated over a box widened along a.
The finite-difference weights are computed here from Fornberg's recursion
in exact rational arithmetic (the reference tabulates the same weights in
src/utils/finite_differencing_stencils.tpp:24-26,55-57,109-111,155-157,211-213).
"""
from __future__ import annotations

import math
from fractions import Fraction

import torch

NGHOST = 4
NVAR = 22
V_CHI, V_GXX, V_KHAT, V_AXX, V_GAMX, V_THETA, V_ALPHA, V_BETAX = 0, 1, 7, 8, 14, 17, 18, 19
SYM = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
PAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def fornberg(m, offsets):
    """Weights of the m-th derivative at 0 on integer ``offsets`` (exact)."""
    xs = [Fraction(o) for o in offsets]
    n = len(xs)
    c = [[Fraction(0)] * (m + 1) for _ in range(n)]
    c[0][0] = Fraction(1)
    c1 = Fraction(1)
    c4 = xs[0]
    for i in range(1, n):
        mn = min(i, m)
        c2 = Fraction(1)
        c5 = c4
        c4 = xs[i]
        for j in range(i):
            c3 = xs[i] - xs[j]
            c2 *= c3
            if j == i - 1:
                for k in range(mn, 0, -1):
                    c[i][k] = c1 * (k * c[i - 1][k - 1] - c5 * c[i - 1][k]) / c2
                c[i][0] = -c1 * c5 * c[i - 1][0] / c2
            for k in range(mn, 0, -1):
                c[j][k] = (c4 * c[j][k] - k * c[j][k - 1]) / c3
            c[j][0] = c4 * c[j][0] / c3
        c1 = c2
    return [float(c[i][m]) for i in range(n)]


# centred 7-point first and second derivatives, lop-sided advective
# derivatives on -4..2 and -2..4, and the 9-point KO difference scaled by
# 2^-8 * (-1) (finite_differencing.tpp:24-32)
OFF7 = list(range(-3, 4))
W_D1 = fornberg(1, OFF7)
W_D2 = fornberg(2, OFF7)
OFF_L = list(range(-4, 3))
OFF_R = list(range(-2, 5))
W_L = fornberg(1, OFF_L)
W_R = fornberg(1, OFF_R)
OFF9 = list(range(-4, 5))
W_KO = [(-1.0) ** o * math.comb(8, o + 4) * -(2.0 ** -8) for o in OFF9]


class Box:
    """Index box [lo, hi] (inclusive, per axis x, y, z) of a padded block."""

    def __init__(self, lo, hi):
        self.lo = list(lo)
        self.hi = list(hi)

    def grow(self, a, e):
        lo, hi = list(self.lo), list(self.hi)
        lo[a] -= e
        hi[a] += e
        return Box(lo, hi)

    def take(self, f, a=0, o=0):
        """View of array f (..., nn3, nn2, nn1) on the box shifted by o along a."""
        lo, hi = list(self.lo), list(self.hi)
        lo[a] += o
        hi[a] += o
        return f[..., lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]


def interior(geom):
    return Box((geom.il, geom.jl, geom.kl), (geom.iu, geom.ju, geom.ku))


def d1(f, box, a, dx):
    """Dx along a of f on box."""
    return sum(w * box.take(f, a, o) for w, o in zip(W_D1, OFF7) if w != 0.0) / dx[a]


def d2(f, box, a, dx):
    return sum(w * box.take(f, a, o) for w, o in zip(W_D2, OFF7)) / (dx[a] * dx[a])


def dmix(f, box, a, b, dx):
    """Dx_a(Dx_b f), a < b, on box: the inner derivative over box widened by
    3 along a, then the outer one."""
    wide = box.grow(a, 3)
    inner = d1(f, wide, b, dx)
    n = [box.hi[d] - box.lo[d] + 1 for d in range(3)]
    dims = {0: -1, 1: -2, 2: -3}
    out = 0.0
    for w, o in zip(W_D1, OFF7):
        if w == 0.0:
            continue
        out = out + w * inner.narrow(dims[a], 3 + o, n[a])
    return out / dx[a]


def lx(f, box, a, dx, vx):
    left = sum(w * box.take(f, a, o) for w, o in zip(W_L, OFF_L)) / dx[a]
    right = sum(w * box.take(f, a, o) for w, o in zip(W_R, OFF_R)) / dx[a]
    return torch.where(vx < 0, vx * left, vx * right)


def ko(f, box, a, dx, diss):
    return sum(w * box.take(f, a, o) for w, o in zip(W_KO, OFF9)) * (diss / dx[a])


def _p(p, name):
    """Parameter ``name`` of p (dict or z4c_params-like); tensors are kept so
    that autograd can differentiate with respect to them."""
    v = p[name] if isinstance(p, dict) else getattr(p, name)
    return v if torch.is_tensor(v) else float(v)


def rhs(u, geom, p):
    """Vacuum Z4cRHS of the padded state ``u`` (22, nn3, nn2, nn1): the
    right-hand side on the interior vertices, zero elsewhere."""
    return _rhs(u, geom, p)[0]


def hamiltonian_norm(u, geom, p):
    """sum over the cell loop (vertices il..iu-1 per axis, as the history
    output integrates, src/outputs/history.cpp:178-205) of dx dy dz Ht^2,
    with Ht = R + 2/3 K^2 - A~_ab A~^ab the Hamiltonian constraint formed
    inside the right-hand side (calculate_z4c_rhs.cpp:606-624)."""
    _, Ht = _rhs(u, geom, p)
    vol = float(geom.dx[0]) * float(geom.dx[1]) * float(geom.dx[2])
    return vol * (Ht[:-1, :-1, :-1] ** 2).sum()


def _rhs(u, geom, p):
    pw = _p(p, "chi_psi_power")
    diss = _p(p, "diss")
    k1, k2 = _p(p, "damp_kappa1"), _p(p, "damp_kappa2")
    dx = [float(geom.dx[a]) for a in range(3)]
    B = interior(geom)
    at = lambda n: B.take(u[n])  # noqa: E731

    chi = at(V_CHI)
    g = [at(V_GXX + i) for i in range(6)]
    A = [at(V_AXX + i) for i in range(6)]
    Khat, Theta, alpha = at(V_KHAT), at(V_THETA), at(V_ALPHA)
    Gam = [at(V_GAMX + i) for i in range(3)]
    beta = [at(V_BETAX + i) for i in range(3)]

    dv = lambda n, a: d1(u[n], B, a, dx)  # noqa: E731

    def ddv(n, a, b):
        if a == b:
            return d2(u[n], B, a, dx)
        a, b = min(a, b), max(a, b)
        return dmix(u[n], B, a, b, dx)

    # metric inverse (linear_algebra.hpp: Det3Metric / Inv3Metric)
    gxx, gxy, gxz, gyy, gyz, gzz = g
    detg = (-(gxz * gxz) * gyy + 2 * gxy * gxz * gyz - gxx * (gyz * gyz)
            - (gxy * gxy) * gzz + gxx * gyy * gzz)
    oodet = 1.0 / detg
    gu = [(-(gyz * gyz) + gyy * gzz) * oodet, (gxz * gyz - gxy * gzz) * oodet,
          (-gxz * gyy + gxy * gyz) * oodet, (-(gxz * gxz) + gxx * gzz) * oodet,
          (gxy * gxz - gxx * gyz) * oodet, (-(gxy * gxy) + gxx * gyy) * oodet]
    G = lambda a, b: g[SYM[a][b]]  # noqa: E731
    GU = lambda a, b: gu[SYM[a][b]]  # noqa: E731

    dg = [[dv(V_GXX + SYM[a][b], c) for (a, b) in PAIRS] for c in range(3)]  # dg[c][ab]
    dG = lambda c, a, b: dg[c][SYM[a][b]]  # noqa: E731
    # Christoffels (calculate_z4c_rhs.cpp:300-340)
    Gddd = [[0.5 * (dG(a, b, c) + dG(b, a, c) - dG(c, a, b)) for (a, b) in PAIRS]
            for c in range(3)]
    Gudd = [[sum(GU(c, d) * Gddd[d][ab] for d in range(3)) for ab in range(6)]
            for c in range(3)]
    Gu = [sum(GU(b, c) * Gudd[a][SYM[b][c]] for b in range(3) for c in range(3))
          for a in range(3)]

    # conformal Ricci (:357-396)
    dGam = [[dv(V_GAMX + c, b) for b in range(3)] for c in range(3)]  # dGam[c][b]
    RR = []
    for (a, b) in PAIRS:
        r = 0.0
        for c in range(3):
            r = r + 0.5 * (G(c, a) * dGam[c][b] + G(c, b) * dGam[c][a]
                           + Gu[c] * (Gddd[a][SYM[b][c]] + Gddd[b][SYM[a][c]]))
        for c in range(3):
            for d in range(3):
                r = r - 0.5 * GU(c, d) * ddv(V_GXX + SYM[a][b], c, d)
        for c in range(3):
            for d in range(3):
                for e in range(3):
                    r = r + GU(c, d) * (Gudd[e][SYM[c][a]] * Gddd[b][SYM[e][d]]
                                        + Gudd[e][SYM[c][b]] * Gddd[a][SYM[e][d]]
                                        + Gudd[e][SYM[a][d]] * Gddd[e][SYM[c][b]])
        RR.append(r)

    # conformal-factor terms (:402-458)
    chi_g = torch.where(chi < _p(p, "chi_div_floor"), torch.full_like(chi, _p(p, "chi_div_floor")),
                        chi)
    oochi = 1.0 / chi_g
    oopsi4 = chi_g if pw == -4.0 else chi_g ** (-4.0 / pw)
    dchi = [dv(V_CHI, a) for a in range(3)]
    dphi = [dchi[a] * oochi * (1.0 / pw) for a in range(3)]
    Ddphi = []
    for (a, b) in PAIRS:
        t = ddv(V_CHI, a, b) * oochi * (1.0 / pw) - pw * dphi[a] * dphi[b]
        for c in range(3):
            t = t - Gudd[c][SYM[a][b]] * dphi[c]
        Ddphi.append(t)
    trphi = sum(GU(c, d) * (Ddphi[SYM[c][d]] + 2.0 * dphi[c] * dphi[d])
                for c in range(3) for d in range(3))
    RR = [RR[SYM[a][b]] + (4.0 * dphi[a] * dphi[b] - 2.0 * Ddphi[SYM[a][b]]
                           - 2.0 * G(a, b) * trphi) for (a, b) in PAIRS]

    # D_a D_b alpha (:484-532)
    dalpha = [dv(V_ALPHA, a) for a in range(3)]
    Dda = sum(GU(c, d) * dphi[c] * dalpha[d] for c in range(3) for d in range(3))
    DDa = []
    for (a, b) in PAIRS:
        t = (ddv(V_ALPHA, a, b) - 2.0 * (dphi[a] * dalpha[b] + dphi[b] * dalpha[a])
             + 2.0 * G(a, b) * Dda)
        for c in range(3):
            t = t - Gudd[c][SYM[a][b]] * dalpha[c]
        DDa.append(t)
    Ddalpha = sum(oopsi4 * GU(a, b) * DDa[SYM[a][b]] for a in range(3) for b in range(3))

    # A~ contractions (:536-601)
    AD = lambda a, b: A[SYM[a][b]]  # noqa: E731
    Aud = [[sum(GU(c, d) * AD(d, b) for d in range(3)) for b in range(3)] for c in range(3)]
    AA = [sum(AD(a, c) * Aud[c][b] for c in range(3)) for (a, b) in PAIRS]
    trAA = sum(GU(a, b) * AA[SYM[a][b]] for a in range(3) for b in range(3))
    Auu = [sum(GU(b, c) * Aud[a][c] for c in range(3)) for (a, b) in PAIRS]
    AU = lambda a, b: Auu[SYM[a][b]]  # noqa: E731
    dK = [dv(V_KHAT, a) for a in range(3)]
    dTh = [dv(V_THETA, a) for a in range(3)]
    cpow = -6.0 / pw
    DA = []
    for a in range(3):
        t = 0.0
        for b in range(3):
            t = t - cpow * AU(a, b) * dchi[b] * oochi
            t = t - (1.0 / 3.0) * GU(a, b) * (2.0 * dK[b] + dTh[b])
        for b in range(3):
            for c in range(3):
                t = t + Gudd[a][SYM[b][c]] * AU(b, c)
        DA.append(t)

    R = sum(oopsi4 * GU(a, b) * RR[SYM[a][b]] for a in range(3) for b in range(3))
    K = Khat + 2.0 * Theta
    Ht = R + (2.0 / 3.0) * K * K - trAA

    dbeta = [[dv(V_BETAX + a, b) for a in range(3)] for b in range(3)]  # dbeta[b][a] = d_b beta^a
    div = dbeta[0][0] + dbeta[1][1] + dbeta[2][2]

    L = [sum(lx(u[n], B, a, dx, beta[a]) for a in range(3)) for n in range(NVAR)]
    D = [sum(ko(u[n], B, a, dx, diss) for a in range(3)) for n in range(NVAR)]

    out = [None] * NVAR
    out[V_CHI] = (L[V_CHI] + (1.0 / 6.0) * pw * chi_g * div
                  - (1.0 / 6.0) * pw * chi_g * alpha * K + D[V_CHI])
    out[V_KHAT] = (-Ddalpha + alpha * (trAA + (1.0 / 3.0) * K * K) + L[V_KHAT]
                   + k1 * (1 - k2) * alpha * Theta + D[V_KHAT])
    out[V_THETA] = L[V_THETA] + alpha * (0.5 * Ht - (2.0 + k2) * k1 * Theta) + D[V_THETA]

    ddbeta_d = [sum((1.0 / 3.0) * ddv(V_BETAX + b, a, b) for b in range(3)) for a in range(3)]
    for a in range(3):
        LG = L[V_GAMX + a] + (2.0 / 3.0) * Gu[a] * div
        for b in range(3):
            LG = LG + GU(a, b) * ddbeta_d[b] - Gu[b] * dbeta[b][a]
            for c in range(3):
                LG = LG + GU(b, c) * ddv(V_BETAX + a, b, c)
        t = 2.0 * alpha * DA[a] + LG - 2.0 * alpha * k1 * (Gam[a] - Gu[a])
        for b in range(3):
            t = t - 2.0 * AU(a, b) * dalpha[b]
        out[V_GAMX + a] = t + D[V_GAMX + a]

    for (a, b) in PAIRS:
        ab = SYM[a][b]
        Lg = L[V_GXX + ab] - (2.0 / 3.0) * g[ab] * div
        LA = L[V_AXX + ab] - (2.0 / 3.0) * A[ab] * div
        for c in range(3):
            Lg = Lg + dbeta[a][c] * G(b, c) + dbeta[b][c] * G(a, c)
            LA = LA + dbeta[b][c] * AD(a, c) + dbeta[a][c] * AD(b, c)
        out[V_GXX + ab] = -2.0 * alpha * A[ab] + Lg + D[V_GXX + ab]
        t = oopsi4 * (-DDa[ab] + alpha * RR[ab])
        t = t - (1.0 / 3.0) * g[ab] * (-Ddalpha + alpha * R)
        t = t + alpha * (K * A[ab] - 2.0 * AA[ab]) + LA
        out[V_AXX + ab] = t + D[V_AXX + ab]

    f = _p(p, "lapse_oplog") * _p(p, "lapse_harmonicf") + _p(p, "lapse_harmonic") * alpha
    out[V_ALPHA] = (_p(p, "lapse_advect") * L[V_ALPHA]
                    - f * alpha * (Khat + 2 * _p(p, "lapse_K") * Theta) + D[V_ALPHA])
    for a in range(3):
        r = _p(p, "shift_Gamma") * Gam[a] + _p(p, "shift_advect") * L[V_BETAX + a]
        if abs(_p(p, "shift_alpha2Gamma")) > 0.0:
            r = r + _p(p, "shift_alpha2Gamma") * alpha * alpha * Gam[a]
        if abs(_p(p, "shift_H")) > 0.0:
            for b in range(3):
                r = r + (_p(p, "shift_H") * alpha * chi_g * (0.5 * alpha * dchi[b] - dalpha[b])
                         * GU(a, b))
        r = r - _p(p, "shift_eta") * beta[a]
        out[V_BETAX + a] = r + D[V_BETAX + a]

    full = torch.zeros_like(u)
    full[:, B.lo[2]:B.hi[2] + 1, B.lo[1]:B.hi[1] + 1, B.lo[0]:B.hi[0] + 1] = torch.stack(out)
    return full, Ht


def algconstr(u, p):
    """AlgConstr (z4c.cpp:677-799) over the whole padded array, returning a
    new tensor."""
    eps = _p(p, "eps_floor")
    g = [u[V_GXX + i] for i in range(6)]
    A = [u[V_AXX + i] for i in range(6)]
    gxx, gxy, gxz, gyy, gyz, gzz = g
    detg = (-(gxz * gxz) * gyy + 2 * gxy * gxz * gyz - gxx * (gyz * gyz)
            - (gxy * gxy) * gzz + gxx * gyy * gzz)
    tf = int(_p(p, "tf_force_regularization"))
    if tf:
        detg = torch.where(detg > 0, detg, torch.ones_like(detg))
        project = torch.ones_like(detg, dtype=torch.bool)
    else:
        project = detg > eps
    s = torch.where(detg - 1.0 < eps, torch.full_like(detg, 1.0 + eps / -3.0),
                    detg.abs() ** (-1.0 / 3.0))
    s = torch.where(project, s, torch.ones_like(s))
    g2 = [s * gi for gi in g]
    gxx, gxy, gxz, gyy, gyz, gzz = g2
    Axx, Axy, Axz, Ayy, Ayz, Azz = A
    tr = (-2. * Ayz * gxx * gyz + Axx * gyy * gzz + gxx * (Azz * gyy + Ayy * gzz)
          + 2. * (gxz * (Ayz * gxy - Axz * gyy + Axy * gyz) + gxy * (Axz * gyz - Axy * gzz))
          - Azz * (gxy * gxy) - Ayy * (gxz * gxz) - Axx * (gyz * gyz))
    A2 = [torch.where(project, A[i] - (1.0 / 3.0) * tr * g2[i], A[i]) for i in range(6)]
    out = list(u.unbind(0))
    for i in range(6):
        out[V_GXX + i] = g2[i]
        out[V_AXX + i] = A2[i]
    if int(_p(p, "force_regularization")):
        chi = u[V_CHI]
        out[V_CHI] = torch.where(chi < eps, torch.full_like(chi, eps), chi)
        al = u[V_ALPHA]
        out[V_ALPHA] = torch.where(al >= 0, torch.where(al < eps, torch.full_like(al, eps), al),
                                   torch.where(al > -eps, torch.full_like(al, -eps), al))
    return torch.stack(out)


BT_A = [0.5, 0.5, 1.0]
BT_B = [1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0]


def rk4_step(u, geom, p, dt):
    """One bt_rk4 step with ghosts held at u_n (zero RHS there) and AlgConstr
    after every stage update, as the GR_Z4c task list orders them
    (task_list_gr_z4c.cpp:45-89)."""
    ks = []
    cur = u
    for s in range(4):
        k = rhs(cur, geom, p)
        ks.append(k)
        if s < 3:
            cur = algconstr(u + (BT_A[s] * dt) * k, p)
    acc = u
    for s in range(4):
        acc = acc + (BT_B[s] * dt) * ks[s]
    return algconstr(acc, p)
