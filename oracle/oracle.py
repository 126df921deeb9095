"""ctypes binding of the CPU oracle (liboracle_z4c.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Imported by tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py -- never by
the product path.

Arrays are numpy float64, C-contiguous, shape (nvar, nn3, nn2, nn1) in the
reference's AthenaArray order u(n,k,j,i) (src/athena_arrays.hpp:192-193).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_z4c.so")

NZ4C, NADM, NCON, NMAT = 22, 17, 7, 10
ND1, NDDCHI, NAUX, NAUXEXT = 33, 6, 30, 7
NGHOST = 4

# Z4c variable indices (src/z4c/z4c.hpp:54-88)
I_CHI, I_GXX, I_KHAT, I_AXX, I_GAMX, I_THETA, I_ALPHA, I_BETAX = 0, 1, 7, 8, 14, 17, 18, 19
Z4C_NAMES = ["chi", "gxx", "gxy", "gxz", "gyy", "gyz", "gzz", "Khat",
             "Axx", "Axy", "Axz", "Ayy", "Ayz", "Azz", "Gamx", "Gamy", "Gamz",
             "Theta", "alpha", "betax", "betay", "betaz"]
# ADM indices (src/z4c/z4c.hpp:99-135)
I_ADM_ALPHA, I_ADM_BETAX, I_ADM_GXX, I_ADM_KXX, I_ADM_PSI4 = 0, 1, 4, 10, 16


class Params(C.Structure):
    """Z4c::opt, defaults of src/z4c/z4c.cpp:318-472."""
    _fields_ = [(n, C.c_double) for n in (
        "chi_psi_power", "chi_div_floor", "diss", "eps_floor", "damp_kappa1",
        "damp_kappa2", "lapse_oplog", "lapse_harmonicf", "lapse_harmonic",
        "lapse_advect", "lapse_K", "shift_Gamma", "shift_alpha2Gamma",
        "shift_H", "shift_advect", "shift_eta", "r_max_con")] + [
        ("cowling", C.c_int), ("tf_force_regularization", C.c_int),
        ("force_regularization", C.c_int), ("pad_", C.c_int)]

    def __init__(self, **kw):
        d = dict(chi_psi_power=-4.0, chi_div_floor=-1000.0, diss=0.0,
                 eps_floor=1e-12, damp_kappa1=0.0, damp_kappa2=0.0,
                 lapse_oplog=2.0, lapse_harmonicf=1.0, lapse_harmonic=0.0,
                 lapse_advect=1.0, lapse_K=0.0, shift_Gamma=1.0,
                 shift_alpha2Gamma=0.0, shift_H=0.0, shift_advect=1.0,
                 shift_eta=2.0, r_max_con=float("inf"), cowling=0,
                 tf_force_regularization=1, force_regularization=0, pad_=0)
        d.update(kw)
        super().__init__(**d)


class Geom(C.Structure):
    _fields_ = [("nn1", C.c_int), ("nn2", C.c_int), ("nn3", C.c_int),
                ("ng", C.c_int), ("il", C.c_int), ("iu", C.c_int),
                ("jl", C.c_int), ("ju", C.c_int), ("kl", C.c_int),
                ("ku", C.c_int), ("dx", C.c_double * 3), ("x0", C.c_double * 3)]

    @classmethod
    def block(cls, nx, dx, xmin):
        """VC block with meshblock/nx = (nx1,nx2,nx3) cells: nx+1 owned
        vertices per axis plus 4 ghosts each side (src/mesh/meshblock.cpp:869-889).
        ``xmin`` is the coordinate of the first interior vertex."""
        nx = np.broadcast_to(np.asarray(nx, dtype=int), (3,))
        dx = np.broadcast_to(np.asarray(dx, dtype=float), (3,))
        xmin = np.broadcast_to(np.asarray(xmin, dtype=float), (3,))
        nn = nx + 1 + 2 * NGHOST
        g = cls()
        g.nn1, g.nn2, g.nn3 = (int(v) for v in nn)
        g.ng = NGHOST
        g.il, g.iu = NGHOST, NGHOST + int(nx[0])
        g.jl, g.ju = NGHOST, NGHOST + int(nx[1])
        g.kl, g.ku = NGHOST, NGHOST + int(nx[2])
        for d in range(3):
            g.dx[d] = float(dx[d])
            g.x0[d] = float(xmin[d]) - NGHOST * float(dx[d])
        return g

    def coarse(self):
        """Coarse VC geometry of this block (mesh_refinement_vc.cpp with
        NCGHOST = 4): per axis nx/2 + 9 vertices, interior 4 .. 4 + nx/2,
        spacing 2 dx, coarse vertex 4 at the block's first interior vertex."""
        c = Geom()
        nx = [self.iu - self.il, self.ju - self.jl, self.ku - self.kl]
        nn = [n // 2 + 9 for n in nx]
        c.nn1, c.nn2, c.nn3 = nn
        c.ng = 4
        c.il, c.iu = 4, 4 + nx[0] // 2
        c.jl, c.ju = 4, 4 + nx[1] // 2
        c.kl, c.ku = 4, 4 + nx[2] // 2
        for d in range(3):
            c.dx[d] = 2.0 * self.dx[d]
            c.x0[d] = self.x0[d] + NGHOST * self.dx[d] - 4 * c.dx[d]
        return c

    @property
    def shape(self):
        return (self.nn3, self.nn2, self.nn1)

    def coords(self):
        """x, y, z arrays of the padded vertices (x = x0 + i*dx)."""
        return tuple(self.x0[d] + np.arange(n) * self.dx[d]
                     for d, n in enumerate((self.nn1, self.nn2, self.nn3)))

    def mesh(self):
        x, y, z = self.coords()
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        return X, Y, Z

    def interior(self):
        return (slice(self.kl, self.ku + 1), slice(self.jl, self.ju + 1),
                slice(self.il, self.iu + 1))


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH) or not os.path.exists(FLOPS_PATH):
        subprocess.run(["make", "-C", HERE, "-B" if force else "all"],
                       check=True, capture_output=True)
    return LIB_PATH


FLOPS_PATH = os.path.join(HERE, "liboracle_z4c_flops.so")
_lib = None
_plain = None
_counting = None


def lib():
    """The library every wrapper calls: the plain build, or the counting
    build inside ``counting()``."""
    global _lib, _plain
    if _lib is None:
        if _plain is None:
            _plain = _setup(C.CDLL(build()))
        _lib = _plain
    return _lib


class counting:
    """Context manager: wrapper calls inside it run the FLOP-counting build
    (oracle/flops/fdouble.hpp; same results) and ``counts`` then holds the
    additions, subtractions, multiplications, divisions and libm calls they
    performed, split into ``stencil`` (inside the finite-difference
    functions) and ``other``."""

    def __enter__(self):
        global _lib, _counting
        lib()
        if _counting is None:
            build()
            _counting = _setup(C.CDLL(FLOPS_PATH))
            _counting.oracle_flops_get.argtypes = [C.POINTER(C.c_ulonglong)]
        _counting.oracle_flops_reset()
        _lib = _counting
        return self

    def __exit__(self, *exc):
        global _lib
        out = (C.c_ulonglong * 10)()
        _counting.oracle_flops_get(out)
        ops = ("add", "sub", "mul", "div", "call")
        self.other = dict(zip(ops, (int(v) for v in out[:5])))
        self.stencil = dict(zip(ops, (int(v) for v in out[5:])))
        self.counts = {k: self.other[k] + self.stencil[k] for k in ops}
        _lib = _plain
        return False


def _setup(L):
    """Argument types of every entry point of a loaded oracle library."""
    if True:
        P = C.POINTER
        d = C.c_double
        pd = P(C.c_double)
        pg = P(Geom)
        pp = P(Params)
        L.oracle_fd_point.restype = d
        L.oracle_fd_point.argtypes = [C.c_int, pg, pd, C.c_long, C.c_int]
        L.oracle_lx_point.restype = d
        L.oracle_lx_point.argtypes = [pg, pd, C.c_long, C.c_int, d]
        L.oracle_prep_derivs.argtypes = [pp, pg, pd, pd, pd, pd]
        L.oracle_rhs.argtypes = [pp, pg, pd, pd, pd, pd, pd]
        L.oracle_boundary_rhs.argtypes = [pp, pg, pd, P(C.c_int), pd]
        L.oracle_algconstr.argtypes = [pp, pg, pd]
        L.oracle_algconstr_range.argtypes = [pp, pg, pd] + [C.c_int] * 7
        L.oracle_adm_to_z4c.argtypes = [pp, pg, pd, pd, pd]
        L.oracle_z4c_to_adm.argtypes = [pp, pg, pd, pd, pd]
        L.oracle_adm_constraints.argtypes = [pp, pg, pd, pd, pd, pd, pd, pd]
        L.oracle_bt_sum.argtypes = [pg, C.c_int, pd, pd, C.c_int, P(pd), pd, d]
        L.oracle_weighted_ave.argtypes = [pg, C.c_int, pd, pd, pd, pd]
        L.oracle_add_rhs.argtypes = [pg, C.c_int, pd, d, pd]
        L.oracle_new_dt.restype = d
        L.oracle_new_dt.argtypes = [pg, d]
        L.oracle_periodic_fill.argtypes = [pg, C.c_int, pd]
        L.oracle_history.argtypes = [pp, pg, pd, pd, pd]
        L.oracle_physical_bc.argtypes = [pg, C.c_int, pd, P(C.c_int)]
        L.oracle_exchange.argtypes = [C.c_int, pg, P(pd), C.c_int, pd, pd, P(C.c_int)]
        ub = P(C.c_ubyte)
        L.oracle_rst_header.restype = C.c_ulong
        L.oracle_rst_header.argtypes = [ub, C.c_int, C.c_int, pd, P(C.c_int), d, d, C.c_int,
                                        C.c_ulonglong]
        L.oracle_rst_idrec.restype = C.c_ulong
        L.oracle_rst_idrec.argtypes = [ub, P(C.c_longlong), C.c_int, d]
    return L


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def zeros(n, g):
    return np.zeros((n,) + g.shape)


def fd_point(kind, g, u, p, d):
    return lib().oracle_fd_point(kind, C.byref(g), _p(u), p, d)


def prep_derivs(o, g, u, want_aux=False):
    d1, dd = zeros(ND1, g), zeros(NDDCHI, g)
    aux = zeros(NAUX, g) if want_aux else None
    lib().oracle_prep_derivs(C.byref(o), C.byref(g), _p(u), _p(d1), _p(dd), _p(aux))
    return (d1, dd, aux) if want_aux else (d1, dd)


def rhs(o, g, u, mat=None, derivs=None):
    """Z4cRHS of ``u`` (prepares derivatives first unless given)."""
    d1, dd = derivs if derivs is not None else prep_derivs(o, g, u)
    out = zeros(NZ4C, g)
    lib().oracle_rhs(C.byref(o), C.byref(g), _p(u), _p(mat), _p(d1), _p(dd), _p(out))
    return out


def boundary_rhs(o, g, u, faces, out):
    f = (C.c_int * 6)(*[int(bool(v)) for v in faces])
    lib().oracle_boundary_rhs(C.byref(o), C.byref(g), _p(u), f, _p(out))
    return out


def algconstr(o, g, u):
    lib().oracle_algconstr(C.byref(o), C.byref(g), _p(u))
    return u


def algconstr_range(o, g, u, il, iu, jl, ju, kl, ku, skip_physical):
    lib().oracle_algconstr_range(C.byref(o), C.byref(g), _p(u), il, iu, jl, ju,
                                 kl, ku, int(skip_physical))
    return u


def adm_to_z4c(o, g, adm, want_aux_ext=False):
    u = zeros(NZ4C, g)
    ax = zeros(NAUXEXT, g) if want_aux_ext else None
    lib().oracle_adm_to_z4c(C.byref(o), C.byref(g), _p(adm), _p(u), _p(ax))
    return (u, ax) if want_aux_ext else u


def z4c_to_adm(o, g, u, want_aux_ext=False):
    adm = zeros(NADM, g)
    ax = zeros(NAUXEXT, g) if want_aux_ext else None
    lib().oracle_z4c_to_adm(C.byref(o), C.byref(g), _p(u), _p(adm), _p(ax))
    return (adm, ax) if want_aux_ext else adm


def adm_constraints(o, g, u, adm, mat=None, derivs=None):
    d1, dd = derivs if derivs is not None else prep_derivs(o, g, u)
    con = zeros(NCON, g)
    lib().oracle_adm_constraints(C.byref(o), C.byref(g), _p(u), _p(adm), _p(mat),
                                 _p(d1), _p(dd), _p(con))
    return con


def bt_sum(g, u0, ks, ws, dt):
    out = np.empty_like(u0)
    arr = (C.POINTER(C.c_double) * len(ks))(*[_p(k) for k in ks])
    w = (C.c_double * len(ws))(*ws)
    lib().oracle_bt_sum(C.byref(g), u0.shape[0], _p(out), _p(u0), len(ks), arr, w, dt)
    return out


def weighted_ave(g, u_out, u1, u2, w):
    ww = (C.c_double * 3)(*w)
    lib().oracle_weighted_ave(C.byref(g), u_out.shape[0], _p(u_out), _p(u1), _p(u2), ww)
    return u_out


def add_rhs(g, rhs_, wght, u_out):
    lib().oracle_add_rhs(C.byref(g), u_out.shape[0], _p(rhs_), wght, _p(u_out))
    return u_out


def new_dt(g, cfl):
    return lib().oracle_new_dt(C.byref(g), cfl)


def history(o, g, u, con, out=None):
    """Z4c .hst columns of one block, accumulated into ``out`` (8 values)."""
    acc = (C.c_double * 8)(*(out if out is not None else [0.0] * 8))
    lib().oracle_history(C.byref(o), C.byref(g), _p(u), _p(con), acc)
    return list(acc)


def physical_bc(g, u, faces):
    """Physical ghost fill on the flagged faces (ix1, ox1, ix2, ox2, ix3, ox3)."""
    f = (C.c_int * 6)(*[int(bool(v)) for v in faces])
    lib().oracle_physical_bc(C.byref(g), u.shape[0], _p(u), f)
    return u


def exchange(geoms, us, lo, hi, periodic):
    """Same-level VC exchange of the blocks ``geoms`` with arrays ``us``
    (modified in place)."""
    n = len(geoms)
    garr = (Geom * n)(*geoms)
    uarr = (C.POINTER(C.c_double) * n)(*[_p(u) for u in us])
    lo_ = (C.c_double * 3)(*lo)
    hi_ = (C.c_double * 3)(*hi)
    per = (C.c_int * 3)(*[int(bool(v)) for v in periodic])
    lib().oracle_exchange(n, garr, uarr, us[0].shape[0], lo_, hi_, per)
    return us


def periodic_fill(g, u):
    lib().oracle_periodic_fill(C.byref(g), u.shape[0], _p(u))
    return u


# Butcher tableau of bt_rk4 (src/task_list/task_list_integrators_butcher.cpp:37-49)
BT_RK4_A = [[0, 0, 0, 0], [0.5, 0, 0, 0], [0, 0.5, 0, 0], [0, 0, 1, 0]]
BT_RK4_B = [1. / 6., 1. / 3., 1. / 3., 1. / 6.]


def rk4_step(o, g, u, dt, mat=None, between_stages=None):
    """One bt_rk4 step in the GR_Z4c task order (task_list_gr_z4c.cpp:45-89):
    k_s = RHS(u); u = u_n + sum a dt k (SumBT_ak(stage+1) / SumBT_bk);
    optional ghost update hook; AlgConstr; derivatives re-prepared by the
    next RHS call.  Ghosts of k are zero, so without a hook ghosts stay u_n."""
    un = u.copy()
    ks = []
    cur = u
    for s in range(4):
        ks.append(rhs(o, g, cur, mat))
        if s < 3:
            cur = bt_sum(g, un, ks, BT_RK4_A[s + 1][:len(ks)], dt)
        else:
            cur = bt_sum(g, un, ks, BT_RK4_B, dt)
        if between_stages is not None:
            between_stages(cur)
        algconstr(o, g, cur)
    return cur, ks


# Low-storage 2S integrators (TaskLists::Integrators::LowStorage,
# src/task_list/task_list_integrators_low_storage.cpp:17-204): per stage
# (delta, gamma_1, gamma_2, gamma_3, beta).  ONE_3RD / TWO_3RD of
# src/defs.hpp.in:257-258 round to 1/3 and 2/3.  ssprk5_4 is left out: the
# GR_Z4c task list rejects it (task_list_gr_z4c.cpp:106-113).
LS_TABLEAUS = {
    "vl2": [(1.0, 0.0, 1.0, 0.0, 0.5), (0.0, 0.0, 1.0, 0.0, 1.0)],
    "rk1": [(1.0, 0.0, 1.0, 0.0, 1.0)],
    "rk2": [(1.0, 0.0, 1.0, 0.0, 1.0), (0.0, 0.5, 0.5, 0.0, 0.5)],
    "rk3": [(1.0, 0.0, 1.0, 0.0, 1.0), (0.0, 0.25, 0.75, 0.0, 0.25),
            (0.0, 2.0 / 3.0, 1.0 / 3.0, 0.0, 2.0 / 3.0)],
    "rk4": [(1.0, 0.0, 1.0, 0.0, 1.193743905974738),
            (0.217683334308543, 0.121098479554482, 0.721781678111411, 0.0, 0.099279895495783),
            (1.065841341361089, -3.843833699660025, 2.121209265338722, 0.0, 1.131678018054042),
            (0.0, 0.546370891121863, 0.198653035682705, 0.0, 0.310665766509336)],
}


def ls_step(o, g, u, dt, tableau, mat=None, between_stages=None, faces=None, rhs_fn=None,
            constrain=True):
    """One low-storage step in the GR_Z4c task order: u1 = 0 at the start
    (StartupTaskList, task_list_gr_z4c.cpp:115-119), then per stage
    rhs = Z4cRHS(u) (+ Z4cBoundaryRHS on ``faces``), and IntegrateZ4c
    (:170-194): u1 = WeightedAve(u1, u, -, {1, delta, 0}), u =
    WeightedAve(u, u1, u2, {gamma_1, gamma_2, gamma_3}), u += (beta dt) rhs
    (AddZ4cRHS); the optional ghost update hook; AlgConstr.  ``tableau`` is
    a name of LS_TABLEAUS or its list.  ``rhs_fn(u)`` replaces the
    right-hand side and ``constrain=False`` skips AlgConstr (for scalar ODE
    checks of the integrator).  Returns u (updated in place)."""
    st = LS_TABLEAUS[tableau] if isinstance(tableau, str) else tableau
    u1 = np.zeros_like(u)
    for delta, g1, g2, g3, beta in st:
        if g3 != 0.0:
            raise ValueError("3S* stages (gamma_3 != 0) need u2, which GR_Z4c never allocates")
        if rhs_fn is not None:
            r = rhs_fn(u)
        else:
            r = rhs(o, g, u, mat)
            if faces is not None:
                boundary_rhs(o, g, u, faces, r)
        weighted_ave(g, u1, u, None, [1.0, delta, 0.0])
        weighted_ave(g, u, u1, None, [g1, g2, g3])
        add_rhs(g, r, beta * dt, u)
        if between_stages is not None:
            between_stages(u)
        if constrain:
            algconstr(o, g, u)
    return u


def weyl(o, g, u, adm=None):
    """Z4c::Z4cWeyl: (2, nn3, nn2, nn1) array of Re, Im psi4 (interior; NaN
    elsewhere).  The ADM variables default to Z4cToADM of u."""
    if adm is None:
        adm = z4c_to_adm(o, g, u)
    d1, dd = prep_derivs(o, g, u)
    out = zeros(2, g)
    lib().oracle_weyl(C.byref(o), C.byref(g), _p(u), _p(adm), _p(d1), _p(dd), _p(out))
    return out


def restrict_vc(g, fine):
    """RestrictInterior for VC data (injection onto the coarse interior):
    returns the coarse buffer (NaN outside the coarse interior)."""
    cg = g.coarse()
    nvar = fine.shape[0]
    coarse = np.full((nvar,) + cg.shape, np.nan)
    lib().oracle_restrict_vc(C.byref(g), C.byref(cg), nvar, _p(fine), _p(coarse))
    return coarse


def prolongate_vc(g, coarse, fine, rng):
    """ProlongateVertexCenteredValues over the coarse index box
    rng = (si, ei, sj, ej, sk, ek), in place on ``fine``."""
    cg = g.coarse()
    lib().oracle_prolongate_vc(C.byref(g), C.byref(cg), fine.shape[0], _p(coarse), _p(fine),
                               *[int(v) for v in rng])
    return fine


def prolongation_range(g, ox, parity):
    """Coarse index box of the prolongation for a coarser neighbour at offset
    ox = (ox1, ox2, ox3), block logical-coordinate parities ``parity``."""
    cg = g.coarse()
    lo = (cg.il, cg.jl, cg.kl)
    hi = (cg.iu, cg.ju, cg.ku)
    out = []
    for a in range(3):
        s_, e_ = C.c_int(), C.c_int()
        lib().oracle_prolongation_range(int(parity[a]), int(ox[a]), lo[a], hi[a], C.byref(s_),
                                        C.byref(e_))
        out += [s_.value, e_.value]
    return tuple(out)


def rst_header(nbtotal, root_level, size9, nx, time, dt, ncycle, datasize):
    """Bytes of the restart mesh header as the reference copies its structs."""
    buf = (C.c_ubyte * 256)()
    n = lib().oracle_rst_header(buf, nbtotal, root_level, (C.c_double * 9)(*size9),
                                (C.c_int * 3)(*nx), time, dt, ncycle, datasize)
    return bytes(buf[:n])


def rst_idrec(lx, level, cost):
    """Bytes of one restart ID-list record (LogicalLocation, cost)."""
    buf = (C.c_ubyte * 64)()
    n = lib().oracle_rst_idrec(buf, (C.c_longlong * 3)(*lx), level, cost)
    return bytes(buf[:n])
