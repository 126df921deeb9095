"""Host-side mirror of the reference's Z4c class surface (src/z4c/z4c.hpp:448-610)
for a batch of vertex-centred MeshBlocks, on one MI355X.

``Z4c`` owns one device context (z4c_ctx) and the block list; the method
names follow the reference (Z4cRHS, AlgConstr, ADMToZ4c, Z4cToADM,
ADMConstraints, PrepareZ4cDerivatives, ...), each acting on every block in
one launch.  Device memory comes from torch (plain flat float64 tensors); the
arithmetic is all in the HIP library behind the C-ABI.

Buffer layout: a buffer of ``nvar`` fields is a flat tensor of
nvar*total_points doubles; block b occupies
[offset_b*nvar, (offset_b+npts_b)*nvar) as (nvar, nn3, nn2, nn1).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _capi
from ._capi import Params, Z4C_NADM, Z4C_NCON, Z4C_NMAT, Z4C_NVAR, check

NGHOST = 4

# Butcher tableau of bt_rk4 (src/task_list/task_list_integrators_butcher.cpp:37-49)
BT_RK4_A = [[0, 0, 0, 0], [0.5, 0, 0, 0], [0, 0.5, 0, 0], [0, 0, 1, 0]]
BT_RK4_B = [1. / 6., 1. / 3., 1. / 3., 1. / 6.]


@dataclass
class MeshBlockGeom:
    """Geometry of one VC MeshBlock: ``nx`` cells per axis (nx+1 owned
    vertices), spacing ``dx`` and coordinate ``xmin`` of the first interior
    vertex (src/mesh/meshblock.cpp:852-915)."""
    nx: tuple
    dx: tuple
    xmin: tuple
    level: int = 0
    sommerfeld_faces: int = 0
    outflow_faces: int = 0

    @property
    def nn(self):
        return tuple(int(n) + 1 + 2 * NGHOST for n in self.nx)

    @property
    def npts(self):
        a, b, c = self.nn
        return a * b * c

    @property
    def shape(self):
        return (self.nn[2], self.nn[1], self.nn[0])

    def coords(self):
        return tuple(self.xmin[d] - NGHOST * self.dx[d] + np.arange(self.nn[d]) * self.dx[d]
                     for d in range(3))

    @staticmethod
    def cube(nx, dx, xmin, level=0, faces=0, outflow=0):
        t = lambda v: tuple(float(x) for x in np.broadcast_to(v, (3,)))
        return MeshBlockGeom(tuple(int(x) for x in np.broadcast_to(nx, (3,))), t(dx), t(xmin),
                             level, faces, outflow)


def _ptr(t):
    if t is None:
        return None
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise _capi.Z4cError("device buffers must be contiguous float64 CUDA tensors")
    return C.c_void_p(t.data_ptr())


def _stream(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class Z4c:
    """Batched Z4c on a list of MeshBlocks (one device context)."""

    def __init__(self, blocks, params: Params | None = None, device: int = 0):
        if not torch.cuda.is_available():
            raise _capi.Z4cError("no GPU visible: the Z4c hot path runs only on the HIP device")
        self.L = _capi.load()
        self.device = torch.device("cuda", device)
        self.params = params if params is not None else Params()
        self.blocks = list(blocks)
        self._ctx = C.c_void_p()
        with torch.cuda.device(self.device):
            check(self.L.z4c_ctx_create(C.byref(self._ctx), device), "ctx_create")
        arr = (_capi.Block * len(self.blocks))()
        off = 0
        self.offsets = []
        for b, g in enumerate(self.blocks):
            nn = g.nn
            x = arr[b]
            x.offset = off
            x.nn1, x.nn2, x.nn3 = nn
            x.il, x.iu = NGHOST, NGHOST + g.nx[0]
            x.jl, x.ju = NGHOST, NGHOST + g.nx[1]
            x.kl, x.ku = NGHOST, NGHOST + g.nx[2]
            for d in range(3):
                x.dx[d] = g.dx[d]
                x.x0[d] = g.xmin[d] - NGHOST * g.dx[d]
            x.level = g.level
            x.sommerfeld_faces = g.sommerfeld_faces
            x.outflow_faces = g.outflow_faces
            self.offsets.append(off)
            off += g.npts
        total = C.c_longlong()
        check(self.L.z4c_set_blocks(self._ctx, arr, len(self.blocks), C.byref(total)),
              "set_blocks")
        self.total_points = int(total.value)
        assert self.total_points == off

    def close(self):
        if self._ctx:
            self.L.z4c_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- buffers ---------------------------------------------------------------
    def empty(self, nvar=Z4C_NVAR):
        return torch.empty(nvar * self.total_points, dtype=torch.float64, device=self.device)

    def zeros(self, nvar=Z4C_NVAR):
        return torch.zeros(nvar * self.total_points, dtype=torch.float64, device=self.device)

    def block_view(self, buf, b, nvar):
        g = self.blocks[b]
        o = self.offsets[b] * nvar
        return buf[o:o + nvar * g.npts].view((nvar,) + g.shape)

    def upload(self, arrays, nvar=None):
        """List of per-block numpy arrays (nvar, nn3, nn2, nn1) -> device buffer."""
        nvar = nvar or arrays[0].shape[0]
        host = np.concatenate([np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in arrays])
        if host.size != nvar * self.total_points:
            raise ValueError("arrays do not match the block list")
        return torch.from_numpy(host).to(self.device)

    def download(self, buf, nvar):
        host = buf.detach().cpu().numpy()
        return [host[self.offsets[b] * nvar:(self.offsets[b] + g.npts) * nvar].reshape((nvar,) + g.shape)
                for b, g in enumerate(self.blocks)]

    # -- the reference surface (src/z4c/z4c.hpp:448-610) ---------------------------
    def PrepareZ4cDerivatives(self, u, aux=None, stream=None):
        check(self.L.z4c_prepare_derivatives(self._ctx, C.byref(self.params), _ptr(u), _ptr(aux),
                                             _stream(stream)), "prepare_derivatives")

    def Z4cRHS(self, u, mat, rhs, stream=None):
        check(self.L.z4c_rhs(self._ctx, C.byref(self.params), _ptr(u), _ptr(mat), _ptr(rhs),
                             _stream(stream)), "rhs")

    def Z4cBoundaryRHS(self, u, mat, rhs, stream=None):
        check(self.L.z4c_boundary_rhs(self._ctx, C.byref(self.params), _ptr(u), _ptr(rhs),
                                      _stream(stream)), "boundary_rhs")

    def AlgConstr(self, u, rng=None, skip_physical=False, stream=None):
        if rng is None:
            check(self.L.z4c_algconstr(self._ctx, C.byref(self.params), _ptr(u), _stream(stream)),
                  "algconstr")
        else:
            check(self.L.z4c_algconstr_range(self._ctx, C.byref(self.params), _ptr(u), *rng,
                                             int(skip_physical), _stream(stream)), "algconstr_range")

    def ADMToZ4c(self, adm, u, aux_ext=None, stream=None):
        check(self.L.z4c_adm_to_z4c(self._ctx, C.byref(self.params), _ptr(adm), _ptr(u),
                                    _ptr(aux_ext), _stream(stream)), "adm_to_z4c")

    def Z4cToADM(self, u, adm, aux_ext=None, stream=None):
        check(self.L.z4c_z4c_to_adm(self._ctx, C.byref(self.params), _ptr(u), _ptr(adm),
                                    _ptr(aux_ext), _stream(stream)), "z4c_to_adm")

    def ADMConstraints(self, con, adm, mat, u, stream=None):
        check(self.L.z4c_adm_constraints(self._ctx, C.byref(self.params), _ptr(u), _ptr(adm),
                                         _ptr(mat), _ptr(con), _stream(stream)), "adm_constraints")

    def Z4cWeyl(self, u_adm, u_mat, u_weyl, u, stream=None):
        """Z4c::Z4cWeyl(u_adm, u_mat, u_weyl) (calculate_weyl_scalars.cpp:44)
        with the state ``u`` explicit (the reference reads storage.u).
        ``u_mat`` is not read: the reference sets its aliases
        (calculate_weyl_scalars.cpp:57-58) but no term uses them."""
        self.weyl(u, u_adm, u_weyl, stream)

    def weyl(self, u, adm, out=None, stream=None):
        """Re / Im psi4 at interior vertices (NaN elsewhere) of state ``u``
        with ADM variables ``adm`` (Z4cToADM of u); returns the 2-field buffer."""
        out = self.empty(2) if out is None else out
        check(self.L.z4c_weyl(self._ctx, C.byref(self.params), _ptr(u), _ptr(adm), _ptr(out),
                              _stream(stream)), "weyl")
        return out

    def AddZ4cRHS(self, rhs, wght, u_out, nvar=Z4C_NVAR, stream=None):
        check(self.L.z4c_add_rhs(self._ctx, nvar, _ptr(rhs), float(wght), _ptr(u_out),
                                 _stream(stream)), "add_rhs")

    def WeightedAve(self, u_out, u_in1, u_in2, wght, nvar=Z4C_NVAR, stream=None):
        w = (C.c_double * 3)(*wght)
        check(self.L.z4c_weighted_ave(self._ctx, nvar, _ptr(u_out), _ptr(u_in1), _ptr(u_in2), w,
                                      _stream(stream)), "weighted_ave")

    def NewBlockTimeStep(self, cfl):
        out = C.c_double()
        check(self.L.z4c_new_dt(self._ctx, float(cfl), C.byref(out)), "new_dt")
        return out.value

    # -- integrator pieces ------------------------------------------------------------
    def bt_sum(self, u_out, u0, ks, ws, dt, nvar=Z4C_NVAR, stream=None):
        """Butcher SumBT_ak / SumBT_bk."""
        kp = (C.c_void_p * max(1, len(ks)))(*[_ptr(k) for k in ks])
        w = (C.c_double * max(1, len(ws)))(*ws)
        check(self.L.z4c_bt_sum(self._ctx, nvar, _ptr(u_out), _ptr(u0), len(ks), kp, w, float(dt),
                                _stream(stream)), "bt_sum")

    def periodic_fill(self, u, nvar=Z4C_NVAR, stream=None):
        check(self.L.z4c_periodic_fill(self._ctx, nvar, _ptr(u), _stream(stream)), "periodic_fill")

    def set_exchange(self, lo, hi, periodic=(True, True, True)):
        """Build the same-level VC exchange of the block list on [lo, hi]."""
        lo_ = (C.c_double * 3)(*np.broadcast_to(np.asarray(lo, float), (3,)))
        hi_ = (C.c_double * 3)(*np.broadcast_to(np.asarray(hi, float), (3,)))
        per = (C.c_int * 3)(*[int(bool(v)) for v in np.broadcast_to(periodic, (3,))])
        check(self.L.z4c_set_exchange(self._ctx, lo_, hi_, per), "set_exchange")

    def exchange(self, u, nvar=Z4C_NVAR, stream=None):
        """Same-level VC ghost exchange (SEND/RECV/SETB of GR_Z4c)."""
        check(self.L.z4c_exchange(self._ctx, nvar, _ptr(u), _stream(stream)), "exchange")

    def physical_bc(self, u, nvar=Z4C_NVAR, stream=None):
        """Physical ghost fill on the flagged faces (PHY_BVAL of GR_Z4c)."""
        check(self.L.z4c_physical_bc(self._ctx, nvar, _ptr(u), _stream(stream)), "physical_bc")

    # -- VC refinement between levels (src/mesh/mesh_refinement_vc.cpp) ---------
    def coarse_shape(self, b):
        """(nn3, nn2, nn1) of block b's coarse buffer: nx/2 + 9 vertices per axis."""
        g = self.blocks[b]
        return tuple(int(g.nx[d]) // 2 + 2 * NGHOST + 1 for d in (2, 1, 0))

    def coarse_layout(self):
        """Per-block point offsets of the coarse buffers and points per field."""
        offs = (C.c_longlong * len(self.blocks))()
        total = C.c_longlong()
        check(self.L.z4c_coarse_layout(self._ctx, offs, C.byref(total)), "coarse_layout")
        return [int(o) for o in offs], int(total.value)

    def coarse_full(self, nvar=Z4C_NVAR, value=0.0):
        _, total = self.coarse_layout()
        return torch.full((nvar * total,), value, dtype=torch.float64, device=self.device)

    def coarse_view(self, buf, b, nvar):
        offs, _ = self.coarse_layout()
        shp = self.coarse_shape(b)
        n = shp[0] * shp[1] * shp[2]
        return buf[offs[b] * nvar:(offs[b] + n) * nvar].view((nvar,) + shp)

    def RestrictInterior(self, u, coarse, nvar=Z4C_NVAR, stream=None):
        """Injection of every block's fine vertices onto its coarse interior
        (src/comm/refinement_ops.cpp:178-205)."""
        check(self.L.z4c_restrict(self._ctx, nvar, _ptr(u), _ptr(coarse), _stream(stream)),
              "restrict")

    def ProlongationRange(self, block, ox, lx):
        """Coarse index box (block, si, ei, sj, ej, sk, ek) prolongated for a
        coarser neighbour at offset ``ox`` of a block at logical location
        ``lx`` (src/comm/refinement_ops.hpp:48-72)."""
        o = (C.c_int * 3)(*[int(v) for v in ox])
        l = (C.c_longlong * 3)(*[int(v) for v in lx])
        box = _capi.ProlongBox()
        check(self.L.z4c_prolongation_range(self._ctx, int(block), o, l, C.byref(box)),
              "prolongation_range")
        return tuple(getattr(box, n) for n, _ in box._fields_)

    def set_prolongation(self, boxes):
        """Boxes (block, si, ei, sj, ej, sk, ek) that ``prolongate`` applies."""
        arr = (_capi.ProlongBox * max(len(boxes), 1))()
        for q, b in enumerate(boxes):
            arr[q] = _capi.ProlongBox(*[int(v) for v in b])
        check(self.L.z4c_set_prolongation(self._ctx, arr, len(boxes)), "set_prolongation")

    def prolongate(self, coarse, u, nvar=Z4C_NVAR, stream=None):
        """ProlongateVertexCenteredValues over every box set by set_prolongation."""
        check(self.L.z4c_prolongate(self._ctx, nvar, _ptr(coarse), _ptr(u), _stream(stream)),
              "prolongate")

    def rk4_stage(self, stage, dt, u, un, k, mat=None, periodic=False, exchange=False,
                  physical=False, stream=None):
        """One GR_Z4c stage; ghosts after SumBT: the single-block periodic
        fill (``periodic``) or the block exchange (``exchange``), then the
        physical ghost fill (``physical``); none of them: ghosts stay u_n."""
        kp = (C.c_void_p * 4)(*[_ptr(x) for x in k])
        mode = (2 if exchange else (1 if periodic else 0)) + (4 if physical else 0)
        check(self.L.z4c_rk4_stage(self._ctx, C.byref(self.params), int(stage), float(dt), _ptr(u),
                                   _ptr(un), kp, _ptr(mat), mode, _stream(stream)),
              "rk4_stage")

    def rk4_step(self, dt, u, un, k, mat=None, periodic=False, exchange=False, physical=False,
                 stream=None):
        """One bt_rk4 step in the GR_Z4c stage order (each stage: Z4cRHS,
        SumBT, the ghost update, AlgConstr, PrepareZ4cDerivatives)."""
        un.copy_(u)  # Butcher::PutScratchBT_u at stage 1
        for s in range(1, 5):
            self.rk4_stage(s, dt, u, un, k, mat, periodic, exchange, physical, stream)

    def rk4_step_fused(self, dt, u, work, mat=None, stream=None, graph=False, periodic=False,
                       exchange=False, physical=False):
        """One bt_rk4 step, fused per stage (RHS assembly + SumBT + AlgConstr
        in one kernel).  Without a ghost update the ghosts are held at u_n
        and the state equals ``rk4_step`` with no ghost update; with
        ``periodic`` / ``exchange`` / ``physical`` every stage updates the
        ghosts as ``rk4_step`` with the same flags does, then applies
        AlgConstr to the whole block.  The derivative arrays are not prepared
        (ADMConstraints prepares its own).  ``work`` comes from
        ``make_step_work``.  ``graph``: replay the step from a HIP graph
        (captured on first use and whenever u, work, dt, the parameters or
        the ghost update change; vacuum only)."""
        mode = (2 if exchange else (1 if periodic else 0)) + (4 if physical else 0)
        if mode:
            if graph:
                if mat is not None:
                    raise _capi.Z4cError("the graph-replayed step takes no matter sources")
                check(self.L.z4c_rk4_step_ghosts_graph(self._ctx, C.byref(self.params), float(dt),
                                                       _ptr(u), _ptr(work), mode,
                                                       _stream(stream)), "rk4_step_ghosts_graph")
                return
            check(self.L.z4c_rk4_step_ghosts(self._ctx, C.byref(self.params), float(dt), _ptr(u),
                                             _ptr(work), _ptr(mat), mode, _stream(stream)),
                  "rk4_step_ghosts")
            return
        if graph:
            if mat is not None:
                raise _capi.Z4cError("rk4_step_graph takes no matter sources: replay a vacuum "
                                     "step, or call the eager step with mat")
            check(self.L.z4c_rk4_step_graph(self._ctx, C.byref(self.params), float(dt), _ptr(u),
                                            _ptr(work), _stream(stream)), "rk4_step_graph")
            return
        check(self.L.z4c_rk4_step(self._ctx, C.byref(self.params), float(dt), _ptr(u),
                                  _ptr(work), _ptr(mat), _stream(stream)), "rk4_step")

    # -- low-storage integrators (task_list_integrators_low_storage.cpp) -------
    def ls_tableau(self, name):
        """The reference's low-storage tableau of <time>/integrator ``name``."""
        t = _capi.LsTableau()
        check(self.L.z4c_ls_tableau_get(name.encode(), C.byref(t)), "ls_tableau_get")
        return t

    def ls_stage(self, tableau, stage, dt, u, u1, rhs, u2=None, mat=None, periodic=False,
                 exchange=False, physical=False, stream=None):
        """One low-storage GR_Z4c stage (u1 zeroed at stage 1; Z4cRHS,
        IntegrateZ4c's two WeightedAve and AddZ4cRHS, ghost update,
        AlgConstr, PrepareZ4cDerivatives); ``tableau`` is a name or an
        LsTableau."""
        t = self.ls_tableau(tableau) if isinstance(tableau, str) else tableau
        mode = (2 if exchange else (1 if periodic else 0)) + (4 if physical else 0)
        check(self.L.z4c_ls_stage(self._ctx, C.byref(self.params), C.byref(t), int(stage),
                                  float(dt), _ptr(u), _ptr(u1), _ptr(u2), _ptr(rhs), _ptr(mat),
                                  mode, _stream(stream)), "ls_stage")

    def ls_step_stages(self, tableau, dt, u, u1, rhs, mat=None, periodic=False, exchange=False,
                       physical=False, stream=None):
        """One low-storage step as the task list runs it: ``ls_stage`` for
        every stage."""
        t = self.ls_tableau(tableau) if isinstance(tableau, str) else tableau
        for st in range(1, t.nstages + 1):
            self.ls_stage(t, st, dt, u, u1, rhs, None, mat, periodic, exchange, physical, stream)

    def ls_step(self, tableau, dt, u, work, mat=None, periodic=False, exchange=False,
                physical=False, stream=None):
        """One low-storage step fused per stage (z4c_ls_step); ``work`` from
        ``make_ls_work``."""
        t = self.ls_tableau(tableau) if isinstance(tableau, str) else tableau
        mode = (2 if exchange else (1 if periodic else 0)) + (4 if physical else 0)
        check(self.L.z4c_ls_step(self._ctx, C.byref(self.params), C.byref(t), float(dt), _ptr(u),
                                 _ptr(work), _ptr(mat), mode, _stream(stream)), "ls_step")

    def make_ls_work(self):
        """Work registers of ``ls_step``: u1 and a second state buffer."""
        return self.empty(2 * Z4C_NVAR)

    HISTORY_COLUMNS = ["H-norm2", "M-norm2", "Mx-norm2", "My-norm2", "Mz-norm2",
                       "Z-norm2", "Theta-norm2", "C-norm2"]

    def history(self, u, con, stream=None):
        """The Z4c columns of the reference's .hst output (history.cpp:178-205)
        for state ``u`` and constraints ``con``; waits for the stream."""
        out = (C.c_double * 8)()
        check(self.L.z4c_history(self._ctx, C.byref(self.params), _ptr(u), _ptr(con), out,
                                 _stream(stream)), "history")
        return list(out)

    def history_noalloc(self, u, con, out, stream=None):
        """``history`` into ``out`` (8 doubles of device memory) without any
        allocation: the per-tile partials live in the library's fixed device
        scratch.  Stream-ordered (does not wait); equals ``history`` bit for
        bit.  Calls that overlap on one GPU race."""
        check(self.L.z4c_history_noalloc(self._ctx, C.byref(self.params), _ptr(u), _ptr(con),
                                         _ptr(out), _stream(stream)), "history_noalloc")

    def autotune(self, u, reps=3, stream=None):
        """Time the stencil-pass workgroup configurations per block size on
        state ``u`` and keep the fastest; returns {block size: threads}."""
        check(self.L.z4c_autotune(self._ctx, C.byref(self.params), _ptr(u), int(reps),
                                  _stream(stream)), "autotune")
        return self.stencil_plan()

    def set_stencil_threads(self, threads, block=-1):
        """Fix the stencil workgroup size of ``block`` (-1: all blocks)."""
        check(self.L.z4c_set_stencil_threads(self._ctx, int(block), int(threads)),
              "set_stencil_threads")

    def stencil_plan(self):
        """{interior vertex counts (ni, nj, nk): stencil workgroup size}."""
        out = {}
        t = C.c_int()
        for b, g in enumerate(self.blocks):
            check(self.L.z4c_stencil_threads(self._ctx, b, C.byref(t)), "stencil_threads")
            out[tuple(n + 1 for n in g.nx)] = t.value
        return out

    def set_timing(self, enable=True):
        """Bracket every pass launch with HIP events on its stream."""
        check(self.L.z4c_set_timing(self._ctx, int(enable)), "set_timing")

    def get_timing(self):
        """{pass: (total ms, launches)} since the last call; waits for the
        recorded events."""
        from ._capi import PASSES
        n = len(PASSES)
        ms = (C.c_double * n)()
        cnt = (C.c_longlong * n)()
        check(self.L.z4c_get_timing(self._ctx, ms, cnt, n), "get_timing")
        return {name: (ms[q], cnt[q]) for q, name in enumerate(PASSES)}

    # -- adjoints (no counterpart in the reference) -------------------------------
    def rhs_vjp(self, u, w, nw=1, params=None, stream=None):
        """VJP of Z4cRHS at state ``u`` for ``nw`` cotangent registers ``w``
        (flat, nw*22*total_points): returns (ubar, pbar) with ubar the nw
        cotangents of u (every vertex) and pbar (nw, 3) the gradients with
        respect to shift_eta, damp_kappa1, damp_kappa2."""
        p = params if params is not None else self.params
        ubar = torch.empty(nw * Z4C_NVAR * self.total_points, dtype=torch.float64, device=self.device)
        pbar = torch.empty(nw * len(_capi.PBAR), dtype=torch.float64, device=self.device)
        check(self.L.z4c_rhs_vjp(self._ctx, C.byref(p), _ptr(u), int(nw), _ptr(w), _ptr(ubar),
                                 _ptr(pbar), _stream(stream)), "rhs_vjp")
        return ubar, pbar.view(nw, len(_capi.PBAR))

    def rk4_step_vjp(self, dt, u, ubar_out, nw=1, params=None, stream=None):
        """VJP of one fused bt_rk4 step (``rk4_step_fused``) from u_n = ``u``:
        ``ubar_out`` holds nw cotangents of u_{n+1}; returns (ubar_in, pbar)."""
        p = params if params is not None else self.params
        ubar = torch.empty(nw * Z4C_NVAR * self.total_points, dtype=torch.float64, device=self.device)
        pbar = torch.empty(nw * len(_capi.PBAR), dtype=torch.float64, device=self.device)
        check(self.L.z4c_rk4_step_vjp(self._ctx, C.byref(p), float(dt), _ptr(u), int(nw),
                                      _ptr(ubar_out), _ptr(ubar), _ptr(pbar), _stream(stream)),
              "rk4_step_vjp")
        return ubar, pbar.view(nw, len(_capi.PBAR))

    def exchange_vjp(self, ubar, nvar=Z4C_NVAR, stream=None):
        """ubar <- (the exchange)^T ubar, in place (z4c_exchange_vjp; one
        spacing or across levels)."""
        check(self.L.z4c_exchange_vjp(self._ctx, int(nvar), _ptr(ubar), _stream(stream)),
              "exchange_vjp")

    def physical_bc_vjp(self, ubar, nvar=Z4C_NVAR, stream=None):
        """ubar <- (the physical fill)^T ubar, in place (z4c_physical_bc_vjp)."""
        check(self.L.z4c_physical_bc_vjp(self._ctx, int(nvar), _ptr(ubar), _stream(stream)),
              "physical_bc_vjp")

    def rk4_step_ghosts_vjp(self, dt, u, ubar_out, nw=1, params=None, stream=None,
                            physical=False):
        """VJP of one step with the exchange after every stage
        (``rk4_step_fused(..., exchange=True, physical=physical)``) from
        u_n = ``u``; returns (ubar_in, pbar) as ``rk4_step_vjp``."""
        p = params if params is not None else self.params
        ubar = torch.empty(nw * Z4C_NVAR * self.total_points, dtype=torch.float64, device=self.device)
        pbar = torch.empty(nw * len(_capi.PBAR), dtype=torch.float64, device=self.device)
        check(self.L.z4c_rk4_step_ghosts_vjp(self._ctx, C.byref(p), float(dt), _ptr(u), int(nw),
                                             _ptr(ubar_out), _ptr(ubar), _ptr(pbar),
                                             6 if physical else 2, _stream(stream)),
              "rk4_step_ghosts_vjp")
        return ubar, pbar.view(nw, len(_capi.PBAR))

    def hamiltonian_norm(self, u, params=None, stream=None):
        """sum over blocks and the history's cell loop of dx dy dz Ht^2, with
        Ht the Hamiltonian constraint of the right-hand side (waits)."""
        p = params if params is not None else self.params
        J = C.c_double()
        check(self.L.z4c_hamiltonian_norm(self._ctx, C.byref(p), _ptr(u), C.byref(J),
                                          _stream(stream)), "hamiltonian_norm")
        return J.value

    def hamiltonian_norm_vjp(self, u, params=None, stream=None):
        """dJ/du of ``hamiltonian_norm`` at every vertex (one register)."""
        p = params if params is not None else self.params
        ubar = torch.empty(Z4C_NVAR * self.total_points, dtype=torch.float64, device=self.device)
        check(self.L.z4c_hamiltonian_norm_vjp(self._ctx, C.byref(p), _ptr(u), _ptr(ubar),
                                              _stream(stream)), "hamiltonian_norm_vjp")
        return ubar

    def make_step_work(self):
        """Work registers of ``rk4_step_fused``: two stage states and the
        running sum of the final combination."""
        return self.empty(3 * Z4C_NVAR)

    def make_rk_registers(self):
        """u_n and the four k registers, zero-initialised (k ghosts stay 0)."""
        return self.zeros(), [self.zeros() for _ in range(4)]
