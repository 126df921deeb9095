"""Independent continuum evaluation of the Z4c right-hand side and the ADM
constraints from closed-form fields (test helper, CPU only).

This is written from the Z4c equations of Bernuzzi & Hilditch (2010) and
Hilditch et al. (2013) in the form GR-Athena++ evolves them
(src/z4c/calculate_z4c_rhs.cpp), but computed a different way from the
oracle so that a dropped term, a sign or an index slip in the oracle shows
up as an O(1) mismatch instead of an O(h^6) one:

* spatial derivatives are exact (sympy), not finite differences;
* the Ricci tensor is the Ricci tensor of the *physical* metric
  gamma_ij = chi^(4/p) g~_ij computed from its own Christoffel symbols, plus
  the Z4c term g~_k(i d_j)(Gam~^k - Gam~_d^k) that replaces the metric
  derived Gamma~ by the evolved one -- not the conformal split
  R~_ij + R^chi_ij the oracle uses;
* D_i D_j alpha uses the physical Christoffels, not the conformal ones plus
  d phi terms;
* every contraction is an explicit numpy einsum.
"""
from __future__ import annotations

import numpy as np
import sympy as sp

X, Y, Z = sp.symbols("x y z", real=True)
XYZ = (X, Y, Z)
SYM = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


class Field:
    """A closed-form scalar field with exact 1st and 2nd derivatives."""

    def __init__(self, expr):
        self.expr = expr
        self.f = sp.lambdify(XYZ, expr, "numpy")
        d1 = [sp.diff(expr, v) for v in XYZ]
        self.d1 = [sp.lambdify(XYZ, e, "numpy") for e in d1]
        self.d2 = [[sp.lambdify(XYZ, sp.diff(d1[a], XYZ[b]), "numpy") for b in range(3)]
                   for a in range(3)]

    def val(self, x):
        return np.broadcast_to(np.asarray(self.f(*x), float), np.shape(x[0])).copy()

    def grad(self, x):
        return np.array([float(self.d1[a](*x)) for a in range(3)])

    def hess(self, x):
        return np.array([[float(self.d2[a][b](*x)) for b in range(3)] for a in range(3)])


def wave_expr(rng, amp, base=0.0, nwaves=2):
    e = sp.Float(base)
    for _ in range(nwaves):
        k = rng.uniform(0.4, 1.3, 3) * rng.choice([-1, 1], 3)
        ph = rng.uniform(0, 2 * np.pi)
        a = amp * rng.uniform(0.5, 1.0)
        e = e + sp.Float(a) * sp.sin(2 * sp.pi * (sp.Float(k[0]) * X + sp.Float(k[1]) * Y
                                                  + sp.Float(k[2]) * Z) + sp.Float(ph))
    return e


def random_state(seed, amp=0.08, with_matter=True):
    """Closed-form Z4c state: 22 field expressions (+10 matter) ordered as
    the reference's variable list (src/z4c/z4c.hpp:54-88)."""
    rng = np.random.default_rng(seed)
    ex = {}
    ex["chi"] = wave_expr(rng, amp, 1.0)
    for n, (a, b) in zip(["gxx", "gxy", "gxz", "gyy", "gyz", "gzz"], SYM):
        ex[n] = wave_expr(rng, amp, 1.0 if a == b else 0.0)
    ex["Khat"] = wave_expr(rng, amp)
    for n in ["Axx", "Axy", "Axz", "Ayy", "Ayz", "Azz"]:
        ex[n] = wave_expr(rng, amp)
    for n in ["Gamx", "Gamy", "Gamz", "Theta"]:
        ex[n] = wave_expr(rng, amp)
    ex["alpha"] = wave_expr(rng, amp, 1.0)
    for n in ["betax", "betay", "betaz"]:
        ex[n] = wave_expr(rng, amp)
    mat = None
    if with_matter:
        mat = [wave_expr(rng, 0.1 * amp, 0.02 if i == 0 else 0.0) for i in range(10)]
    return ex, mat


ORDER = ["chi", "gxx", "gxy", "gxz", "gyy", "gyz", "gzz", "Khat", "Axx", "Axy",
         "Axz", "Ayy", "Ayz", "Azz", "Gamx", "Gamy", "Gamz", "Theta", "alpha",
         "betax", "betay", "betaz"]


def sample(exprs, geom):
    """Evaluate expressions on the padded vertices of an oracle Geom."""
    Xg, Yg, Zg = geom.mesh()
    out = np.empty((len(exprs),) + geom.shape)
    for n, e in enumerate(exprs):
        f = sp.lambdify(XYZ, e, "numpy")
        out[n] = np.broadcast_to(f(Xg, Yg, Zg), geom.shape)
    return out


def _sym(vals6):
    m = np.empty((3, 3) + np.shape(vals6[0]))
    for n, (a, b) in enumerate(SYM):
        m[a, b] = vals6[n]
        m[b, a] = vals6[n]
    return m


def _christoffel(gm, dgm, ddgm=None):
    """Christoffels of the 2nd kind of metric gm (3x3), dgm[m,a,b] = d_m g_ab;
    with ddgm[m,n,a,b] also returns d_m Gamma^k_ij."""
    gi = np.linalg.inv(gm)
    low = 0.5 * (np.einsum("ijl->lij", dgm) + np.einsum("jil->lij", dgm) - dgm)
    # low[l,i,j] = 1/2 (d_i g_jl + d_j g_il - d_l g_ij)
    G = np.einsum("kl,lij->kij", gi, low)
    if ddgm is None:
        return gi, G
    dgi = -np.einsum("ka,mab,bl->mkl", gi, dgm, gi)
    dlow = 0.5 * (np.einsum("mijl->mlij", ddgm) + np.einsum("mjil->mlij", ddgm)
                  - ddgm)
    # dlow[m,l,i,j] = 1/2 (d_m d_i g_jl + d_m d_j g_il - d_m d_l g_ij)
    dG = np.einsum("mkl,lij->mkij", dgi, low) + np.einsum("kl,mlij->mkij", gi, dlow)
    return gi, G, dG


def _ricci(G, dG):
    # R_ij = d_k G^k_ij - d_j G^k_ik + G^k_kl G^l_ij - G^k_jl G^l_ik
    return (np.einsum("kkij->ij", dG) - np.einsum("jkik->ij", dG)
            + np.einsum("kkl,lij->ij", G, G) - np.einsum("kjl,lik->ij", G, G))


class ContinuumZ4c:
    def __init__(self, exprs, mat_exprs=None, p=-4.0):
        self.p = p
        self.F = {n: Field(exprs[n]) for n in ORDER}
        psi4 = exprs["chi"] ** sp.Float(4.0 / p)
        self.gam = [Field(psi4 * exprs[n]) for n in ["gxx", "gxy", "gxz", "gyy", "gyz", "gzz"]]
        Kfac = sp.Rational(1, 3) * (exprs["Khat"] + 2 * exprs["Theta"])
        self.Kdd = [Field(psi4 * exprs["A" + s] + Kfac * psi4 * exprs["g" + s])
                    for s in ["xx", "xy", "xz", "yy", "yz", "zz"]]
        self.M = [Field(e) for e in mat_exprs] if mat_exprs is not None else None

    def rhs(self, x, o):
        """The 22 continuum right-hand sides at point x for oracle params o."""
        F = self.F
        p = self.p
        gs = ["g" + s for s in ["xx", "xy", "xz", "yy", "yz", "zz"]]
        As = ["A" + s for s in ["xx", "xy", "xz", "yy", "yz", "zz"]]
        g = _sym([float(F[n].val(x)) for n in gs])
        dg = np.zeros((3, 3, 3))
        ddg = np.zeros((3, 3, 3, 3))
        for n, (a, b) in enumerate(SYM):
            gr = F[gs[n]].grad(x)
            hs = F[gs[n]].hess(x)
            dg[:, a, b] = dg[:, b, a] = gr
            ddg[:, :, a, b] = ddg[:, :, b, a] = hs
        A = _sym([float(F[n].val(x)) for n in As])
        dA = np.zeros((3, 3, 3))
        for n, (a, b) in enumerate(SYM):
            dA[:, a, b] = dA[:, b, a] = F[As[n]].grad(x)
        chi = float(F["chi"].val(x)); dchi = F["chi"].grad(x)
        alpha = float(F["alpha"].val(x)); dalpha = F["alpha"].grad(x); ddalpha = F["alpha"].hess(x)
        Khat = float(F["Khat"].val(x)); dKhat = F["Khat"].grad(x)
        Theta = float(F["Theta"].val(x)); dTheta = F["Theta"].grad(x)
        beta = np.array([float(F[n].val(x)) for n in ["betax", "betay", "betaz"]])
        dbeta = np.array([F[n].grad(x) for n in ["betax", "betay", "betaz"]]).T  # [k, i] = d_k beta^i
        ddbeta = np.array([F[n].hess(x) for n in ["betax", "betay", "betaz"]])  # [i, k, l]
        Gam = np.array([float(F[n].val(x)) for n in ["Gamx", "Gamy", "Gamz"]])
        dGam = np.array([F[n].grad(x) for n in ["Gamx", "Gamy", "Gamz"]]).T  # [k, i]

        # physical metric and its derivatives (exact)
        gam = _sym([float(f.val(x)) for f in self.gam])
        dgam = np.zeros((3, 3, 3)); ddgam = np.zeros((3, 3, 3, 3))
        for n, (a, b) in enumerate(SYM):
            dgam[:, a, b] = dgam[:, b, a] = self.gam[n].grad(x)
            ddgam[:, :, a, b] = ddgam[:, :, b, a] = self.gam[n].hess(x)
        gami, PG, dPG = _christoffel(gam, dgam, ddgam)
        Ric = _ricci(PG, dPG)

        # conformal metric-derived contracted Christoffel and its gradient
        gi, CG, dCG = _christoffel(g, dg, ddg)
        dgi = -np.einsum("ka,mab,bl->mkl", gi, dg, gi)
        Gd = np.einsum("ij,kij->k", gi, CG)
        dGd = np.einsum("mij,kij->mk", dgi, CG) + np.einsum("ij,mkij->mk", gi, dCG)
        # Z4c Ricci: evolved Gamma~ replaces the metric one in g~_k(i d_j) Gamma~^k
        Zt = dGam - dGd  # [m, k]
        Rz = 0.5 * (np.einsum("ki,jk->ij", g, Zt) + np.einsum("kj,ik->ij", g, Zt))
        Rij = Ric + Rz
        DDa = ddalpha - np.einsum("kij,k->ij", PG, dalpha)

        K = Khat + 2 * Theta
        psi4 = chi ** (4.0 / p)
        oopsi4 = 1.0 / psi4
        Auu = gi @ A @ gi
        AA = A @ gi @ A
        trAA = np.einsum("ij,ij->", A, Auu)
        trDDa = np.einsum("ij,ij->", gami, DDa)
        trR = np.einsum("ij,ij->", gami, Rij)
        divb = np.trace(dbeta)
        adv = lambda grad: float(beta @ grad)

        if self.M is not None:
            mv = [float(m.val(x)) for m in self.M]
            rho = mv[0]; Sd = np.array(mv[1:4]); Sdd = _sym(mv[4:10])
        else:
            rho = 0.0; Sd = np.zeros(3); Sdd = np.zeros((3, 3))
        Str = np.einsum("ij,ij->", gami, Sdd)
        k1, k2 = o.damp_kappa1, o.damp_kappa2

        out = {}
        out["chi"] = adv(dchi) + (p / 6.0) * chi * (divb - alpha * K)
        liedg = (np.einsum("k,kij->ij", beta, dg) + np.einsum("ik,jk->ij", g, dbeta)
                 + np.einsum("jk,ik->ij", g, dbeta) - (2.0 / 3.0) * g * divb)
        rg = -2 * alpha * A + liedg
        out["Khat"] = (-trDDa + alpha * (trAA + K * K / 3.0) + adv(dKhat)
                       + k1 * (1 - k2) * alpha * Theta + 4 * np.pi * alpha * (Str + rho))
        liedA = (np.einsum("k,kij->ij", beta, dA) + np.einsum("ik,jk->ij", A, dbeta)
                 + np.einsum("jk,ik->ij", A, dbeta) - (2.0 / 3.0) * A * divb)
        rA = (oopsi4 * (-DDa + alpha * Rij) - g / 3.0 * (-trDDa + alpha * trR)
              + alpha * (K * A - 2 * AA) + liedA
              - 8 * np.pi * alpha * (oopsi4 * Sdd - Str * g / 3.0))
        out["Theta"] = (adv(dTheta) + alpha * (0.5 * (trR + (2.0 / 3.0) * K * K - trAA)
                                                - (2 + k2) * k1 * Theta)
                        - 8 * np.pi * alpha * rho)
        rGam = (-2 * Auu @ dalpha
                + 2 * alpha * (np.einsum("ijk,jk->i", CG, Auu) + (6.0 / p) * Auu @ dchi / chi
                               - (gi @ (2 * dKhat + dTheta)) / 3.0 - 8 * np.pi * gi @ Sd)
                + np.einsum("j,ji->i", beta, dGam)
                - np.einsum("j,ji->i", Gd, dbeta) + (2.0 / 3.0) * Gd * divb
                + np.einsum("jk,ijk->i", gi, ddbeta)
                + gi @ np.einsum("kjk->j", ddbeta) / 3.0
                - 2 * alpha * k1 * (Gam - Gd))
        f = o.lapse_oplog * o.lapse_harmonicf + o.lapse_harmonic * alpha
        out["alpha"] = o.lapse_advect * adv(dalpha) - f * alpha * (Khat + 2 * o.lapse_K * Theta)
        rbeta = np.array([o.shift_Gamma * Gam[i] + o.shift_advect * adv(dbeta[:, i])
                          for i in range(3)])
        rbeta = rbeta + o.shift_alpha2Gamma * alpha ** 2 * Gam
        rbeta = rbeta + o.shift_H * alpha * chi * (gi @ (0.5 * alpha * dchi - dalpha))
        rbeta = rbeta - o.shift_eta * beta

        res = np.empty(22)
        res[0] = out["chi"]
        for n, (a, b) in enumerate(SYM):
            res[1 + n] = rg[a, b]
            res[8 + n] = rA[a, b]
        res[7] = out["Khat"]
        res[14:17] = rGam
        res[17] = out["Theta"]
        res[18] = out["alpha"]
        res[19:22] = rbeta
        return res

    def constraints(self, x):
        """H, M_i (covariant), |M|^2 from the physical metric and K_ij."""
        gam = _sym([float(f.val(x)) for f in self.gam])
        dgam = np.zeros((3, 3, 3)); ddgam = np.zeros((3, 3, 3, 3))
        K = _sym([float(f.val(x)) for f in self.Kdd]); dK = np.zeros((3, 3, 3))
        for n, (a, b) in enumerate(SYM):
            dgam[:, a, b] = dgam[:, b, a] = self.gam[n].grad(x)
            ddgam[:, :, a, b] = ddgam[:, :, b, a] = self.gam[n].hess(x)
            dK[:, a, b] = dK[:, b, a] = self.Kdd[n].grad(x)
        gi, G, dG = _christoffel(gam, dgam, ddgam)
        R = np.einsum("ij,ij->", gi, _ricci(G, dG))
        trK = np.einsum("ij,ij->", gi, K)
        KK = np.einsum("ij,jk,kl,li->", gi, K, gi, K)
        if self.M is not None:
            rho = float(self.M[0].val(x)); Sd = np.array([float(m.val(x)) for m in self.M[1:4]])
        else:
            rho = 0.0; Sd = np.zeros(3)
        H = R + trK ** 2 - KK - 16 * np.pi * rho
        # D_k K_ij = d_k K_ij - G^l_ki K_lj - G^l_kj K_il
        DK = dK - np.einsum("lki,lj->kij", G, K) - np.einsum("lkj,il->kij", G, K)
        # M^i = D_j K^ij - D^i K - 8 pi S^i
        Mu = (np.einsum("ia,jb,jab->i", gi, gi, DK) - np.einsum("ij,ab,jab->i", gi, gi, DK)
              - 8 * np.pi * gi @ Sd)
        Md = gam @ Mu
        return H, Md, float(Mu @ gam @ Mu)
