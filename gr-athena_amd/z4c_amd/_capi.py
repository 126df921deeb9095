"""ctypes binding of the C-ABI declared in include/z4c_mi355x.h.

The shared library is the product; this module only declares prototypes and
turns negative return codes into exceptions.  It never falls back to another
implementation: if the library or a GPU is missing, loading fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

from . import build as _build

LIB_PATH = _build.LIB

Z4C_NVAR, Z4C_NADM, Z4C_NCON, Z4C_NMAT = 22, 17, 7, 10
ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "ESTATE", -4: "ENOMEM"}

# Every symbol the header declares (checked by the CPU test-suite).
EXPORTS = [
    "z4c_last_error", "z4c_abi_version", "z4c_ctx_create", "z4c_ctx_destroy",
    "z4c_set_blocks", "z4c_prepare_derivatives", "z4c_rhs", "z4c_boundary_rhs",
    "z4c_algconstr", "z4c_algconstr_range", "z4c_adm_to_z4c", "z4c_z4c_to_adm",
    "z4c_adm_constraints", "z4c_bt_sum", "z4c_weighted_ave", "z4c_add_rhs",
    "z4c_new_dt", "z4c_periodic_fill", "z4c_rk4_stage", "z4c_rk4_step",
    "z4c_set_timing", "z4c_get_timing", "z4c_history", "z4c_set_exchange", "z4c_exchange",
    "z4c_autotune", "z4c_stencil_threads", "z4c_set_stencil_threads", "z4c_rk4_step_graph",
    "z4c_physical_bc", "z4c_rhs_vjp", "z4c_rk4_step_vjp", "z4c_hamiltonian_norm",
    "z4c_exchange_vjp", "z4c_rk4_step_ghosts_vjp", "z4c_physical_bc_vjp",
    "z4c_hamiltonian_norm_vjp", "z4c_weyl", "z4c_coarse_layout", "z4c_restrict",
    "z4c_prolongation_range", "z4c_set_prolongation", "z4c_prolongate", "z4c_rk4_step_ghosts",
    "z4c_ls_tableau_get", "z4c_ls_stage", "z4c_ls_step", "z4c_history_noalloc",
    "z4c_rk4_step_ghosts_graph",
]

# parameter gradients of the adjoints (Z4C_PBAR_* of the header)
PBAR = ["shift_eta", "damp_kappa1", "damp_kappa2"]

# kernel-timer passes (Z4C_PASS_* of the header)
PASSES = ["stencil", "curvature", "assembly", "ghosts", "prepare", "bt_sum", "algconstr"]


class Params(C.Structure):
    """z4c_params == Z4c::opt (src/z4c/z4c.hpp:374-430); defaults of
    src/z4c/z4c.cpp:318-472."""
    _fields_ = [(n, C.c_double) for n in (
        "chi_psi_power", "chi_div_floor", "diss", "eps_floor", "damp_kappa1",
        "damp_kappa2", "lapse_oplog", "lapse_harmonicf", "lapse_harmonic",
        "lapse_advect", "lapse_K", "shift_Gamma", "shift_alpha2Gamma",
        "shift_H", "shift_advect", "shift_eta", "r_max_con")] + [
        ("cowling", C.c_int), ("tf_force_regularization", C.c_int),
        ("force_regularization", C.c_int), ("pad_", C.c_int)]

    DEFAULTS = dict(chi_psi_power=-4.0, chi_div_floor=-1000.0, diss=0.0,
                    eps_floor=1e-12, damp_kappa1=0.0, damp_kappa2=0.0,
                    lapse_oplog=2.0, lapse_harmonicf=1.0, lapse_harmonic=0.0,
                    lapse_advect=1.0, lapse_K=0.0, shift_Gamma=1.0,
                    shift_alpha2Gamma=0.0, shift_H=0.0, shift_advect=1.0,
                    shift_eta=2.0, r_max_con=float("inf"), cowling=0,
                    tf_force_regularization=1, force_regularization=0, pad_=0)

    def __init__(self, **kw):
        d = dict(self.DEFAULTS)
        unknown = set(kw) - set(d)
        if unknown:
            raise TypeError(f"unknown Z4c parameters: {sorted(unknown)}")
        d.update(kw)
        super().__init__(**d)

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Block(C.Structure):
    """z4c_block: one vertex-centred MeshBlock."""
    _fields_ = [("offset", C.c_longlong), ("nn1", C.c_int), ("nn2", C.c_int),
                ("nn3", C.c_int), ("il", C.c_int), ("iu", C.c_int), ("jl", C.c_int),
                ("ju", C.c_int), ("kl", C.c_int), ("ku", C.c_int),
                ("dx", C.c_double * 3), ("x0", C.c_double * 3), ("level", C.c_int),
                ("sommerfeld_faces", C.c_int), ("outflow_faces", C.c_int), ("pad_", C.c_int)]


class ProlongBox(C.Structure):
    """z4c_prolong_box: one ProlongateVertexCenteredValues call."""
    _fields_ = [(n, C.c_int) for n in ("block", "si", "ei", "sj", "ej", "sk", "ek")]


Z4C_LS_MAX_STAGES = 5


class LsTableau(C.Structure):
    """z4c_ls_tableau: per-stage weights of a low-storage 2S integrator
    (TaskLists::Integrators::LowStorage)."""
    _fields_ = [("nstages", C.c_int), ("pad_", C.c_int)] + [
        (n, C.c_double * Z4C_LS_MAX_STAGES) for n in ("delta", "gamma_1", "gamma_2", "gamma_3",
                                                       "beta")]

    def stages(self):
        """[(delta, gamma_1, gamma_2, gamma_3, beta), ...]"""
        return [tuple(getattr(self, n)[q] for n in ("delta", "gamma_1", "gamma_2", "gamma_3",
                                                    "beta")) for q in range(self.nstages)]


class Z4cError(RuntimeError):
    pass


_lib = None


def load(path: str | None = None):
    """Load the library (no build, no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise Z4cError(f"{p} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(p)
    vp, i, d, ll = C.c_void_p, C.c_int, C.c_double, C.c_longlong
    pp = C.POINTER(Params)
    L.z4c_last_error.restype = C.c_char_p
    L.z4c_last_error.argtypes = []
    L.z4c_abi_version.restype = i
    sig = {
        "z4c_ctx_create": [C.POINTER(vp), i],
        "z4c_ctx_destroy": [vp],
        "z4c_set_blocks": [vp, C.POINTER(Block), i, C.POINTER(ll)],
        "z4c_prepare_derivatives": [vp, pp, vp, vp, vp],
        "z4c_rhs": [vp, pp, vp, vp, vp, vp],
        "z4c_boundary_rhs": [vp, pp, vp, vp, vp],
        "z4c_algconstr": [vp, pp, vp, vp],
        "z4c_algconstr_range": [vp, pp, vp, i, i, i, i, i, i, i, vp],
        "z4c_adm_to_z4c": [vp, pp, vp, vp, vp, vp],
        "z4c_z4c_to_adm": [vp, pp, vp, vp, vp, vp],
        "z4c_adm_constraints": [vp, pp, vp, vp, vp, vp, vp],
        "z4c_bt_sum": [vp, i, vp, vp, i, C.POINTER(vp), C.POINTER(d), d, vp],
        "z4c_weighted_ave": [vp, i, vp, vp, vp, C.POINTER(d), vp],
        "z4c_add_rhs": [vp, i, vp, d, vp, vp],
        "z4c_new_dt": [vp, d, C.POINTER(d)],
        "z4c_periodic_fill": [vp, i, vp, vp],
        "z4c_rk4_stage": [vp, pp, i, d, vp, vp, C.POINTER(vp), vp, i, vp],
        "z4c_rk4_step": [vp, pp, d, vp, vp, vp, vp],
        "z4c_rk4_step_ghosts": [vp, pp, d, vp, vp, vp, i, vp],
        "z4c_rk4_step_ghosts_graph": [vp, pp, d, vp, vp, i, vp],
        "z4c_set_timing": [vp, i],
        "z4c_history": [vp, pp, vp, vp, C.POINTER(d), vp],
        "z4c_history_noalloc": [vp, pp, vp, vp, vp, vp],
        "z4c_set_exchange": [vp, C.POINTER(d), C.POINTER(d), C.POINTER(i)],
        "z4c_exchange": [vp, i, vp, vp],
        "z4c_exchange_vjp": [vp, i, vp, vp],
        "z4c_physical_bc_vjp": [vp, i, vp, vp],
        "z4c_autotune": [vp, pp, vp, i, vp],
        "z4c_stencil_threads": [vp, i, C.POINTER(i)],
        "z4c_set_stencil_threads": [vp, i, i],
        "z4c_rk4_step_graph": [vp, pp, d, vp, vp, vp],
        "z4c_physical_bc": [vp, i, vp, vp],
        "z4c_get_timing": [vp, C.POINTER(d), C.POINTER(ll), i],
        "z4c_rhs_vjp": [vp, pp, vp, i, vp, vp, vp, vp],
        "z4c_rk4_step_vjp": [vp, pp, d, vp, i, vp, vp, vp, vp],
        "z4c_rk4_step_ghosts_vjp": [vp, pp, d, vp, i, vp, vp, vp, i, vp],
        "z4c_hamiltonian_norm": [vp, pp, vp, C.POINTER(d), vp],
        "z4c_hamiltonian_norm_vjp": [vp, pp, vp, vp, vp],
        "z4c_weyl": [vp, pp, vp, vp, vp, vp],
        "z4c_coarse_layout": [vp, C.POINTER(ll), C.POINTER(ll)],
        "z4c_restrict": [vp, i, vp, vp, vp],
        "z4c_prolongation_range": [vp, i, C.POINTER(i), C.POINTER(ll), C.POINTER(ProlongBox)],
        "z4c_set_prolongation": [vp, C.POINTER(ProlongBox), i],
        "z4c_prolongate": [vp, i, vp, vp, vp],
        "z4c_ls_tableau_get": [C.c_char_p, C.POINTER(LsTableau)],
        "z4c_ls_stage": [vp, pp, C.POINTER(LsTableau), i, d, vp, vp, vp, vp, vp, i, vp],
        "z4c_ls_step": [vp, pp, C.POINTER(LsTableau), d, vp, vp, vp, i, vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.restype = i
        f.argtypes = args
    _lib = L
    return L


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = load().z4c_last_error().decode(errors="replace")
        raise Z4cError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")
    return rc
