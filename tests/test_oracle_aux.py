"""Pins the oracle's remaining functions against closed forms, brute force
and invariants (CPU):

* SumBT with the bt_rk4 tableau reproduces the RK4 stability polynomial on
  y' = lambda y (src/task_list/task_list_integrators_butcher.cpp:37-49,
  573-680);
* WeightedAve and AddZ4cRHS are the linear combinations they name, over the
  owned vertices (src/mesh/weighted_ave.cpp:21-167,
  src/z4c/add_z4c_rhs.cpp:24-34); NewBlockTimeStep is cfl*min(dx)
  (src/z4c/new_blockdt_z4c.cpp:24-70);
* the periodic VC fill equals periodic identification of vertices with the
  duplicated boundary vertices averaged, by brute force over vertex classes
  (src/comm/index_utilities.hpp:116-139, src/comm/comm_channel.cpp:1277-1330);
* ranged AlgConstr acts as the whole-block one inside its range and nowhere
  else, and skips the interior when asked (src/z4c/z4c.cpp:801-991);
* the Sommerfeld face RHS converges at second order to the radiative
  condition d_t u = -c (u + x^a d_a u)/r with c = sqrt(2) for Khat and 1
  otherwise (src/z4c/boundary_conditions.cpp:66-174);
* the physical derivatives of PrepareZ4cDerivatives' aux output converge at
  sixth order to exact derivatives of alpha, beta and gamma = chi^(4/p) g~
  (src/z4c/adm_z4c.cpp:1044-1119), and PrepareAuxExtended's sqrt(det gamma)
  and gamma^ij equal numpy's det/inv of the ADM metric
  (src/z4c/adm_z4c.cpp:1139-1231).
"""
import itertools

import numpy as np
import pytest
import sympy as sp

import continuum as ct
import oracle as orc


def test_bt_rk4_tableau_gives_rk4_polynomial():
    g = orc.Geom.block(3, 0.1, 0.0)
    rng = np.random.default_rng(3)
    u0 = rng.standard_normal((2,) + g.shape)
    lam, dt = -0.7 + 0.0, 0.3
    ks, cur = [], u0
    for s in range(4):
        ks.append(lam * cur)
        w = orc.BT_RK4_A[s + 1][:len(ks)] if s < 3 else orc.BT_RK4_B
        cur = orc.bt_sum(g, u0, ks, w, dt)
    z = lam * dt
    poly = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    np.testing.assert_allclose(cur, u0 * poly, rtol=1e-14, atol=1e-15)
    # fourth order: the one-step error against exp(z) scales as dt^5
    e = []
    for h in (0.2, 0.1):
        z = lam * h
        e.append(abs(1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24 - np.exp(z)))
    assert 28 < e[0] / e[1] < 36


def test_linear_combinations_and_dt():
    g = orc.Geom.block(4, (0.1, 0.05, 0.2), 0.0)
    rng = np.random.default_rng(4)
    a, b, c = (rng.standard_normal((3,) + g.shape) for _ in range(3))
    for w in ([0.3, 0.7, 0.0], [1.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.2, 0.3, 0.5],
              [0.0, 1.0, 0.0]):
        got = orc.weighted_ave(g, a.copy(), b, c, w)
        exp = a.copy()  # WeightedAveVC spans the owned vertices is..ive only
        I = (slice(None),) + g.interior()
        exp[I] = w[0] * a[I] + w[1] * b[I] + w[2] * c[I]
        np.testing.assert_allclose(got, exp, rtol=1e-14, atol=1e-15)
    got = orc.add_rhs(g, b, 0.25, a.copy())
    exp = a.copy()
    exp[(slice(None),) + g.interior()] += 0.25 * b[(slice(None),) + g.interior()]
    np.testing.assert_array_equal(got, exp)
    assert orc.new_dt(g, 0.25) == 0.25 * 0.05


@pytest.mark.parametrize("nx", [5, 6])
def test_periodic_fill_is_vertex_identification(nx):
    g = orc.Geom.block(nx, 0.1, 0.0)
    rng = np.random.default_rng(nx)
    u = rng.standard_normal((2,) + g.shape)
    got = orc.periodic_fill(g, u.copy())
    # vertex class along an axis: (i - il) mod nx (il and iu are one vertex)
    cls = (np.arange(g.nn1) - g.il) % nx
    lo, hi = g.il, g.iu
    exp = np.empty_like(u)
    for n in range(u.shape[0]):
        sums = np.zeros((nx, nx, nx))
        cnt = np.zeros((nx, nx, nx))
        for k, j, i in itertools.product(range(lo, hi + 1), repeat=3):
            sums[cls[k], cls[j], cls[i]] += u[n, k, j, i]
            cnt[cls[k], cls[j], cls[i]] += 1
        mean = sums / cnt
        exp[n] = mean[np.ix_(cls, cls, cls)]
    np.testing.assert_allclose(got, exp, rtol=1e-14, atol=1e-15)


def _random_state(g, seed):
    rng = np.random.default_rng(seed)
    u = orc.zeros(22, g)
    for n, (a, b) in enumerate(ct.SYM):
        u[1 + n] = (1.2 if a == b else 0.0) + 0.05 * rng.standard_normal(g.shape)
        u[8 + n] = 0.1 * rng.standard_normal(g.shape)
    u[0] = 1.0 + 0.1 * rng.random(g.shape)
    u[18] = 1.0 + 0.1 * rng.random(g.shape)
    return u


def test_algconstr_range_acts_only_inside():
    g = orc.Geom.block(4, 0.1, 0.0)
    u = _random_state(g, 9)
    o = orc.Params()
    full = orc.algconstr(o, g, u.copy())
    rng_ = (2, 9, 0, 6, 5, 12)  # il,iu,jl,ju,kl,ku (padded indices)
    got = orc.algconstr_range(o, g, u.copy(), *rng_, 0)
    m = np.zeros(g.shape, bool)
    m[rng_[4]:rng_[5] + 1, rng_[2]:rng_[3] + 1, rng_[0]:rng_[1] + 1] = True
    np.testing.assert_array_equal(got[:, m], full[:, m])
    np.testing.assert_array_equal(got[:, ~m], u[:, ~m])
    got = orc.algconstr_range(o, g, u.copy(), *rng_, 1)
    phys = np.zeros(g.shape, bool)
    phys[g.interior()] = True
    np.testing.assert_array_equal(got[:, m & ~phys], full[:, m & ~phys])
    np.testing.assert_array_equal(got[:, ~(m & ~phys)], u[:, ~(m & ~phys)])


def _sommerfeld_fields():
    exprs = {}
    rng = np.random.default_rng(11)
    for n in [7, 17] + list(range(8, 14)) + [14, 15, 16]:
        exprs[n] = ct.wave_expr(rng, 0.1, base=0.05)
    return exprs


def test_sommerfeld_converges_to_radiative_condition():
    exprs = _sommerfeld_fields()
    fields = {n: ct.Field(e) for n, e in exprs.items()}
    errs = []
    for h in (0.05, 0.025):
        nx = 8
        # the checked face vertex sits at the same point for both spacings
        P = np.array([1.0, 0.1, 0.45])
        g = orc.Geom.block(nx, h, P - h * np.array([nx, nx // 2, nx // 2]))
        u = orc.zeros(22, g)
        X, Y, Zc = g.mesh()
        for n, f in fields.items():
            u[n] = f.f(X, Y, Zc) * np.ones(g.shape)
        u[0] = 1.0; u[18] = 1.0
        u[1] = u[4] = u[6] = 1.0
        out = orc.zeros(22, g)
        orc.boundary_rhs(orc.Params(), g, u, [0, 1, 0, 0, 0, 0], out)
        i, j, k = g.iu, g.jl + nx // 2, g.kl + nx // 2
        x = np.array([X[k, j, i], Y[k, j, i], Zc[k, j, i]])
        r = np.linalg.norm(x)
        e = []
        for n, f in fields.items():
            c = np.sqrt(2.0) if n == 7 else 1.0
            exact = -c * (float(f.f(*x)) + x @ f.grad(x)) / r
            e.append(abs(out[n, k, j, i] - exact))
        # variables outside the radiative set and other faces stay untouched
        assert not np.any(out[[0, 1, 18, 19], :, :, :])
        assert not np.any(out[:, :, :, g.il])
        errs.append(np.array(e))
    ratio = errs[0] / errs[1]
    assert np.all((ratio > 3.5) & (ratio < 4.6)), ratio


def test_prepared_aux_and_aux_extended():
    ex, mat = ct.random_state(7, with_matter=False)
    fields = [ct.Field(ex[n]) for n in ct.ORDER]
    x0 = np.array([0.13, -0.21, 0.37])
    o = orc.Params()
    chi_e = ex[ct.ORDER[0]]
    gam = [ex[ct.ORDER[1 + ab]] / chi_e for ab in range(6)]
    dgam = [[sp.lambdify(ct.XYZ, sp.diff(gam[ab], v), "numpy") for ab in range(6)]
            for v in ct.XYZ]
    errs = []
    for h in (1 / 20, 1 / 40):
        g = orc.Geom.block(8, h, x0 - 4 * h)
        u = ct.sample([ex[n] for n in ct.ORDER], g)
        _, _, aux = orc.prep_derivs(o, g, u, want_aux=True)
        c = (g.kl + 4, g.jl + 4, g.il + 4)
        got = aux[(slice(None),) + c]
        exact = np.zeros(30)
        exact[0:3] = fields[18].grad(x0)
        for cc in range(3):
            for m in range(3):
                exact[3 + cc * 3 + m] = fields[19 + m].grad(x0)[cc]
            for ab in range(6):
                exact[12 + cc * 6 + ab] = float(dgam[cc][ab](*x0))
        errs.append(np.abs(got - exact))
    e0, e1 = errs
    big = e0 > 1e-11
    assert np.all(e1[~big] < 1e-10)
    assert np.all((e0[big] / e1[big] > 40) & (e0[big] / e1[big] < 100)), e0[big] / e1[big]
    # aux_extended: sqrt(det gamma) and gamma^ij of the ADM metric, which
    # the reference forms assuming det g~ = 1, i.e. after AlgConstr has
    # projected g~ (det g~ > 1 here, so the projection applies)
    g = orc.Geom.block(4, 0.1, 0.0)
    u = orc.algconstr(o, g, _random_state(g, 12))
    adm, ax = orc.z4c_to_adm(o, g, u, want_aux_ext=True)
    gm = np.moveaxis(ct._sym(list(adm[4:10])), (0, 1), (-2, -1))
    np.testing.assert_allclose(ax[0], np.sqrt(np.linalg.det(gm)), rtol=1e-13)
    inv = np.linalg.inv(gm)
    for n, (a, b) in enumerate(ct.SYM):
        np.testing.assert_allclose(ax[1 + n], inv[..., a, b], rtol=1e-12, atol=1e-14)


def test_adm_to_z4c_against_closed_form():
    """chi = det(gamma)^(-1/3), g~ = chi gamma, Khat = gamma^ij K_ij,
    A~ = chi (K_ij - gamma_ij K/3) pointwise, and Gam~^a = -d_b g~^(ab)
    converging at sixth order (src/z4c/adm_z4c.cpp:46-192)."""
    ex, _ = ct.random_state(3, with_matter=False)
    x0 = np.array([0.13, -0.21, 0.37])
    # exact d_b g~^(ab) at x0 from exact first derivatives of gamma_ij:
    # g~^ab = d^(1/3) gamma^ab, d gamma^ab = -gamma^ac (d gamma_cd) gamma^db,
    # d det = det gamma^cd d gamma_cd
    fg = [ct.Field(ex[ct.ORDER[1 + n]]) for n in range(6)]
    G0 = ct._sym([f.val(tuple(x0)) for f in fg]).reshape(3, 3)
    dG = np.zeros((3, 3, 3))  # dG[c, a, b] = d_c gamma_ab
    for n, (a, b) in enumerate(ct.SYM):
        dG[:, a, b] = dG[:, b, a] = fg[n].grad(x0)
    Gi0 = np.linalg.inv(G0)
    det0 = np.linalg.det(G0)
    dGi = -np.einsum("ac,kcd,db->kab", Gi0, dG, Gi0)
    ddet = det0 * np.einsum("cd,kcd->k", Gi0, dG)
    d13 = det0 ** (1.0 / 3.0)
    dgt_inv = d13 * dGi + np.einsum("ab,k->kab", Gi0, ddet * d13 / (3.0 * det0))
    Gam_exact = -np.einsum("bab->a", dgt_inv)
    Kex = [ex[ct.ORDER[8 + n]] for n in range(6)]  # any smooth symmetric K_ij
    o = orc.Params()
    errs = []
    for h in (1 / 20, 1 / 40):
        g = orc.Geom.block(8, h, x0 - 4 * h)
        adm = orc.zeros(17, g)
        adm[0] = 1.0
        adm[4:10] = ct.sample([ex[ct.ORDER[1 + n]] for n in range(6)], g)
        adm[10:16] = ct.sample(Kex, g)
        u = orc.adm_to_z4c(o, g, adm)
        c = (g.kl + 4, g.jl + 4, g.il + 4)
        errs.append(np.abs(u[(slice(14, 17),) + c] - Gam_exact))
        # pointwise parts
        G = np.moveaxis(ct._sym(list(adm[4:10])), (0, 1), (-2, -1))
        K = np.moveaxis(ct._sym(list(adm[10:16])), (0, 1), (-2, -1))
        chi_n = np.linalg.det(G) ** (-1.0 / 3.0)
        Gi = np.linalg.inv(G)
        trK = np.einsum("...ij,...ij->...", Gi, K)
        np.testing.assert_allclose(u[0], chi_n, rtol=1e-14)
        np.testing.assert_allclose(u[7], trK, rtol=1e-11, atol=1e-13)
        for n, (a, b) in enumerate(ct.SYM):
            np.testing.assert_allclose(u[1 + n], chi_n * G[..., a, b], rtol=1e-12, atol=1e-14)
            np.testing.assert_allclose(u[8 + n], chi_n * (K[..., a, b] - G[..., a, b] * trK / 3),
                                       rtol=1e-10, atol=1e-13)
    ratio = errs[0] / errs[1]
    assert np.all((ratio > 40) & (ratio < 100)), ratio


def test_history_sums_known_fields():
    """The oracle's .hst sums: constant fields give value * cells * volume
    over the reference's cell loop (nx^3 cells, not (nx+1)^3 vertices)."""
    g = orc.Geom.block(6, 0.5, -1.5)
    con = orc.zeros(7, g)
    con[0], con[1], con[2], con[3], con[4:7] = 7.0, 2.0, 3.0, 5.0, 1.0
    u = orc.zeros(22, g)
    u[17] = 0.5
    out = orc.history(orc.Params(), g, u, con)
    V = 6 ** 3 * 0.5 ** 3
    np.testing.assert_allclose(out, np.array([4, 3, 1, 1, 1, 5, 0.25, 7]) * V, rtol=1e-14)
    out = orc.history(orc.Params(r_max_con=1.0), g, u, con)
    c = -1.5 + 0.25 + 0.5 * np.arange(6)
    inside = sum(1 for a in c for b in c for e in c if a * a + b * b + e * e <= 1.0)
    np.testing.assert_allclose(out[0], 4 * inside * 0.125, rtol=1e-14)


def _split(g_big, nx_big, parts, h, lo):
    """Geometries of the parts of a cube split per axis (VC: neighbours share
    their boundary vertices)."""
    sub = [nx_big // p for p in parts]
    geoms, starts = [], []
    for c in range(parts[2]):
        for b in range(parts[1]):
            for a in range(parts[0]):
                s = (a * sub[0], b * sub[1], c * sub[2])
                geoms.append(orc.Geom.block(sub, h, [lo + s[d] * h for d in range(3)]))
                starts.append(s)
    return geoms, starts


def test_exchange_single_periodic_block_equals_periodic_fill():
    g = orc.Geom.block(6, 0.1, 0.0)
    rng = np.random.default_rng(21)
    u = rng.standard_normal((2,) + g.shape)
    a = orc.periodic_fill(g, u.copy())
    (b,) = orc.exchange([g], [u.copy()], [0.0] * 3, [0.6] * 3, [1, 1, 1])
    np.testing.assert_allclose(b, a, rtol=1e-15, atol=1e-15)


def test_exchange_of_split_block_equals_whole_block():
    """A periodic 8^3-cell block split into 2 x 2 x 1 blocks: after the
    exchange every vertex of every part (ghosts included) equals the whole
    block's periodic fill at the same place."""
    h, lo, n = 0.125, -0.5, 8
    big = orc.Geom.block(n, h, lo)
    rng = np.random.default_rng(22)
    u = rng.standard_normal((2,) + big.shape)
    ref = orc.periodic_fill(big, u.copy())
    geoms, starts = _split(big, n, (2, 2, 1), h, lo)
    parts = []
    for g, s in zip(geoms, starts):
        v = rng.standard_normal((2,) + g.shape)  # ghosts: garbage to be replaced
        sl = (slice(None),) + tuple(slice(4 + s[2 - d], 4 + s[2 - d] + (g.ku - g.kl, g.ju - g.jl,
                                                                              g.iu - g.il)[d] + 1)
                                    for d in range(3))
        v[(slice(None),) + g.interior()] = ref[sl]
        parts.append(v)
    orc.exchange(geoms, parts, [lo] * 3, [lo + n * h] * 3, [1, 1, 1])
    for g, s, v in zip(geoms, starts, parts):
        sl = (slice(None),) + tuple(slice(s[2 - d], s[2 - d] + g.shape[d]) for d in range(3))
        # the part's padded box, cut from the whole block's periodic extension
        ext = np.pad(ref[(slice(None),) + big.interior()][:, :-1, :-1, :-1],
                     ((0, 0),) + ((4, 5),) * 3, mode="wrap")
        np.testing.assert_allclose(v, ext[sl], rtol=1e-14, atol=1e-15)


def test_exchange_brute_force_nonperiodic():
    """Two blocks side by side in x on a non-periodic domain, distinct data:
    every vertex is the mean over the interior vertices at the same place,
    found by coordinates; vertices outside the domain keep their values."""
    h = 0.25
    geoms = [orc.Geom.block((4, 4, 4), h, (0.0, 0.0, 0.0)),
             orc.Geom.block((4, 4, 4), h, (1.0, 0.0, 0.0))]
    rng = np.random.default_rng(23)
    us = [rng.standard_normal((1,) + g.shape) for g in geoms]
    before = [u.copy() for u in us]
    orc.exchange(geoms, us, [0.0] * 3, [2.0, 1.0, 1.0], [0, 0, 0])
    owners = {}
    for b, g in enumerate(geoms):
        X, Y, Z = g.mesh()
        for k in range(g.kl, g.ku + 1):
            for j in range(g.jl, g.ju + 1):
                for i in range(g.il, g.iu + 1):
                    key = tuple(np.round(np.array([X[k, j, i], Y[k, j, i], Z[k, j, i]]) / h).astype(int))
                    owners.setdefault(key, []).append(before[b][0, k, j, i])
    for b, g in enumerate(geoms):
        X, Y, Z = g.mesh()
        for k in range(g.nn3):
            for j in range(g.nn2):
                for i in range(g.nn1):
                    key = tuple(np.round(np.array([X[k, j, i], Y[k, j, i], Z[k, j, i]]) / h).astype(int))
                    exp = np.mean(owners[key]) if key in owners else before[b][0, k, j, i]
                    assert abs(us[b][0, k, j, i] - exp) <= 1e-15 * max(1.0, abs(exp)), (b, i, j, k)


def _fill_set(g, faces):
    """Vertices the reference's VC physical fill writes (cell-centred ranges
    is..ie = il..iu-1 on the vertex arrays, physical_bcs.cpp:506-676)."""
    ng = 4
    is_, ie, js, je, ks, ke = g.il, g.iu - 1, g.jl, g.ju - 1, g.kl, g.ku - 1
    bjs = js if faces[2] else js - ng
    bje = je if faces[3] else je + ng
    bks = ks if faces[4] else ks - ng
    bke = ke if faces[5] else ke + ng
    pts = set()
    for k in range(bks, bke + 1):
        for j in range(bjs, bje + 1):
            if faces[0]:
                pts |= {(i, j, k) for i in range(is_ - ng, is_)}
            if faces[1]:
                pts |= {(i, j, k) for i in range(ie + 1, ie + ng + 1)}
    for k in range(bks, bke + 1):
        for i in range(is_ - ng, ie + ng + 1):
            if faces[2]:
                pts |= {(i, j, k) for j in range(js - ng, js)}
            if faces[3]:
                pts |= {(i, j, k) for j in range(je + 1, je + ng + 1)}
    for j in range(js - ng, je + ng + 1):
        for i in range(is_ - ng, ie + ng + 1):
            if faces[4]:
                pts |= {(i, j, k) for k in range(ks - ng, ks)}
            if faces[5]:
                pts |= {(i, j, k) for k in range(ke + 1, ke + ng + 1)}
    return pts


@pytest.mark.parametrize("faces", [(1, 1, 1, 1, 1, 1), (0, 1, 0, 0, 1, 0), (1, 0, 1, 1, 0, 0)])
def test_physical_bc_extrapolates_quartics_over_reference_ranges(faces):
    """5-point extrapolation is exact for degree <= 4: a separable quartic
    set everywhere except the fill set (garbage there) comes back whole on
    the fill set; every other vertex keeps its value."""
    g = orc.Geom.block((6, 7, 5), 0.1, (0.2, -0.1, 0.3))
    X, Y, Z = g.mesh()
    poly = (1 + X - 2 * X ** 2 + 0.5 * X ** 4) * (2 - Y + Y ** 3) * (1 + 0.3 * Z ** 2 - Z ** 4)
    fill = _fill_set(g, faces)
    u = np.stack([poly, 2 * poly])
    rng = np.random.default_rng(5)
    for (i, j, k) in fill:
        u[:, k, j, i] = rng.standard_normal(2) * 1e3
    before = u.copy()
    orc.physical_bc(g, u, faces)
    mask = np.zeros(g.shape, bool)
    for (i, j, k) in fill:
        mask[k, j, i] = True
    # exact up to rounding lifted by the weights (sum |c| = 31 per layer,
    # nested over up to four layers and three faces at the corners)
    ref = np.stack([poly, 2 * poly])
    np.testing.assert_allclose(u[:, mask], ref[:, mask], rtol=1e-6,
                               atol=1e-6 * np.abs(ref).max())
    np.testing.assert_array_equal(u[:, ~mask], before[:, ~mask])
