"""Multi-level ghost fill of vertex-centred Z4c data, restated from the
reference (TEST INFRASTRUCTURE: the checker of the device exchange; nothing
in the product imports it).

One ghost update of the GR_Z4c stage (SEND_Z4C, RECV_Z4C, SETB_Z4C, PROLONG;
src/task_list/gr/task_list_gr_z4c.cpp:47-58) over a 2:1-balanced set of
equally sized blocks on a box with periodic or open axes (no neighbour
beyond an open face: its ghosts stay zero here, for the physical fill):

1. every block restricts its fine interior into its coarse buffer
   (CommRegistry::SendBoundaryBuffers -> RestrictInterior,
   src/comm/comm_registry.cpp:278-305, src/comm/refinement_ops.cpp:178-212);
2. every block packs, from these pre-unpack values, one message per
   neighbour (CommChannel::PackAndSend, src/comm/comm_channel.cpp:616-817):
   same level -> its fine slab (LoadSameLevel) plus its coarse slab
   (LoadSameLevel on the coarse buffer) when the receiver has a neighbour
   on another level; to a coarser neighbour -> its coarse slab
   (LoadToCoarser); to a finer one -> a half-block fine slab (LoadToFiner);
3. every block zeroes its fine and coarse ghosts, unpacks the messages in
   its neighbour order (CommChannel::Unpack, :990-1472): same level and
   from finer add into the fine array (SetSameLevel / SetFromFiner), the
   coarse payload adds into the coarse buffer, from coarser overwrites the
   coarse buffer (SetFromCoarser); then divides both arrays by the node
   multiplicity (CommRegistry::SetBoundaries, src/comm/comm_registry.cpp:
   336-382; NodeMultiplicity, src/comm/node_multiplicity.cpp:104-440);
4. every block with coarser neighbours fills the ghosts of its coarse buffer
   at physical faces (ApplyPhysicalBCsOnCoarseLevel,
   src/comm/physical_bcs.cpp:687-876: the fine level's face fill on the
   coarse buffer with the vertex-centred coarse ranges civs..cive, so the
   5-point extrapolation fills cive+1..cive+4 and keeps the boundary vertex)
   and prolongates its coarse buffer over each coarser neighbour's
   ProlongationRange box (ProlongateBoundaries,
   src/comm/refinement_ops.cpp:143-176).

Neighbours are found and ordered as mesh_topology::SearchAndSetNeighbors
(src/mesh/mesh_topology.cpp:278-960) does on the block tree (x1, x2, x3
faces, then x1x2, x1x3, x2x3 edges, then corners; coarser edge and corner
neighbours only where the block sits in that corner of its parent); index
ranges follow src/comm/index_utilities.{hpp,cpp} for Sampling::VC.

Pinned by tests/test_oracle_ml_exchange.py: on one level it equals the
same-level exchange; constant and linear fields stay exact; a tensor
product of quintics (which the prolongation reproduces) is exact at every
ghost and coarse-buffer vertex of a two- and a three-level mesh.
"""
from __future__ import annotations

import numpy as np

import oracle as orc

NG = 4   # NGHOST
CNG = 4  # NCGHOST


class Block:
    """Vertex-centred block: nx cells per axis (equal for all blocks, as
    the reference's single meshblock/nx), spacing dx, first interior vertex
    at xmin."""

    def __init__(self, nx, dx, xmin):
        self.nx = tuple(int(v) for v in nx)
        self.dx = tuple(float(v) for v in dx)
        self.xmin = tuple(float(v) for v in xmin)
        self.geom = orc.Geom.block(self.nx, self.dx, self.xmin)


def _axis_fine(n):
    """Fine VC index set of one axis (MeshBlock::SetIndicialParameters,
    src/mesh/meshblock.cpp:852-898) with n cells: vs, ve, ms, pe, gs, ge."""
    vs = NG
    ve = vs + n
    return dict(vs=vs, ve=ve, ms=0, me=NG - 1, ps=n + NG + 1, pe=n + 2 * NG,
                gs=vs + NG, ge=ve - NG)


def _axis_coarse(n):
    """Coarse VC index set of one axis (nx/2 cells, NCGHOST ghosts)."""
    h = n // 2
    vs = CNG
    ve = vs + h
    return dict(vs=vs, ve=ve, ms=0, me=CNG - 1, ps=h + CNG + 1, pe=h + 2 * CNG,
                gs=vs + CNG, ge=ve - CNG)


# ---------------------------------------------------------------------------
# per-axis index ranges (src/comm/index_utilities.hpp:116-210, VC)
# ---------------------------------------------------------------------------
def load_same(ox, a):
    return (a["ge"] if ox > 0 else a["vs"], a["gs"] if ox < 0 else a["ve"])


def set_same(ox, a):
    if ox == 0:
        return a["vs"], a["ve"]
    return (a["ve"], a["pe"]) if ox > 0 else (a["ms"], a["vs"])


def load_to_coarser(ox, c):
    if ox > 0:
        return c["ve"] - NG, c["ve"]
    if ox < 0:
        return c["vs"], c["vs"] + NG
    return c["vs"], c["ve"]


def set_from_finer(ox, a, fi, half):
    if ox == 0:
        s, e = a["vs"], a["ve"]
        if fi == 1:
            s += half
        else:
            e -= half
        return s, e
    return (a["ve"], a["pe"]) if ox > 0 else (a["ms"], a["vs"])


def load_to_finer(ox, a, fi, half):
    if ox > 0:
        return a["ve"] - CNG, a["ve"] - 1
    if ox < 0:
        return a["vs"] + 1, a["vs"] + CNG
    s, e = a["vs"], a["ve"]
    if fi == 1:
        s += half
    else:
        e -= half
    return s, e


def set_from_coarser(ox, c, lx):
    if ox == 0:
        s, e = c["vs"], c["ve"]
        if lx & 1 == 0:
            e = c["pe"]
        else:
            s = c["ms"]
        return s, e
    return (c["ps"], c["pe"]) if ox > 0 else (c["ms"], c["me"])


def _fi_axis(ni, d):
    """fi1 / fi2 of axis d for an ox == 0 axis (index_utilities.cpp:769-772)."""
    ox = ni["ox"]
    use_fi1 = True if d == 0 else (ox[0] != 0 if d == 1 else (ox[0] != 0 and ox[1] != 0))
    return ni["fi1"] if use_fi1 else ni["fi2"]


# ---------------------------------------------------------------------------
# node multiplicity (src/comm/node_multiplicity.cpp:104-183, 343-436)
# ---------------------------------------------------------------------------
def _compact(kind, ox, lx=0, fi=0):
    if kind == "same":
        if ox == 0:
            return 1, 5
        return (5, 6) if ox > 0 else (0, 1)
    if kind == "coarser":
        if ox == 0:
            return (1, 6) if lx & 1 == 0 else (0, 5)
        return (6, 6) if ox > 0 else (0, 0)
    if ox == 0:
        return (3, 5) if fi == 1 else (1, 3)
    return (5, 6) if ox > 0 else (0, 1)


def _phys_range(c, t):
    if c & 1:
        return t[c], t[c]
    if c <= 3:
        return t[c], t[c + 1] - 1
    return t[c - 1] + 1, t[c]


def _divide(arr, mult, axes):
    """ApplyDivisionImpl on (nvar, k, j, i) with per-axis index sets."""
    piv = []
    for a in axes:
        half = (a["ve"] - a["vs"]) // 2
        piv.append([a["ms"], a["vs"], a["vs"] + 1, a["vs"] + half, a["ve"] - 1, a["ve"], a["pe"]])
    for K in range(7):
        kl, ku = _phys_range(K, piv[2])
        for J in range(7):
            jl, ju = _phys_range(J, piv[1])
            for I in range(7):
                m = mult[K, J, I]
                if m <= 1:
                    continue
                il, iu = _phys_range(I, piv[0])
                arr[:, kl:ku + 1, jl:ju + 1, il:iu + 1] *= 1.0 / m


def _zero_ghosts(arr, axes):
    ax, ay, az = axes
    keep = arr[:, az["vs"]:az["ve"] + 1, ay["vs"]:ay["ve"] + 1, ax["vs"]:ax["ve"] + 1].copy()
    arr[...] = 0.0
    arr[:, az["vs"]:az["ve"] + 1, ay["vs"]:ay["ve"] + 1, ax["vs"]:ax["ve"] + 1] = keep


# ---------------------------------------------------------------------------
# the block tree
# ---------------------------------------------------------------------------
class Mesh:
    """Blocks on the box [lo, hi] (``periodic[d]``: axis d wraps); levels
    follow from the spacing (level 0 = the coarsest spacing present)."""

    def __init__(self, blocks, lo, hi, periodic=(True, True, True)):
        self.blocks = blocks
        self.periodic = tuple(bool(p) for p in periodic)
        self.lo = np.asarray(lo, float)
        self.hi = np.asarray(hi, float)
        nx = blocks[0].nx
        if any(b.nx != nx for b in blocks):
            raise ValueError("all blocks need the same nx")
        if any(n % 2 for n in nx):
            raise ValueError("nx must be even")
        self.nx = nx
        dmax = max(b.dx[0] for b in blocks)
        self.level, self.loc = [], []
        self.leaf = {}
        for q, b in enumerate(blocks):
            L = int(round(np.log2(dmax / b.dx[0])))
            if not np.allclose(np.asarray(b.dx), dmax * 2.0 ** -L):
                raise ValueError("spacing is not the coarsest halved L times")
            lx = tuple(int(round((b.xmin[d] - self.lo[d]) / (nx[d] * b.dx[d]))) for d in range(3))
            self.level.append(L)
            self.loc.append(lx)
            self.leaf[(L,) + lx] = q
        self.nrb0 = [int(round((self.hi[d] - self.lo[d]) / (nx[d] * dmax))) for d in range(3)]
        self.nbrs = [self._neighbors(q) for q in range(len(blocks))]
        self.all_same = [all(n["level"] == self.level[q] for n in self.nbrs[q])
                         for q in range(len(blocks))]

    def _wrap(self, L, t):
        out = []
        for d in range(3):
            n = self.nrb0[d] << L
            if not 0 <= t[d] < n:
                if not self.periodic[d]:
                    return None
            out.append(t[d] % n)
        return tuple(out)

    def _covered(self, L, t):
        """'same' (a leaf at L, t), 'finer' (t refined) or 'coarser'."""
        if (L,) + t in self.leaf:
            return "same"
        kids = [(L + 1, 2 * t[0] + a, 2 * t[1] + b, 2 * t[2] + c)
                for c in (0, 1) for b in (0, 1) for a in (0, 1)]
        if any(k in self.leaf for k in kids):
            return "finer"
        par = (L - 1, t[0] >> 1, t[1] >> 1, t[2] >> 1)
        if par in self.leaf:
            return "coarser"
        raise ValueError("mesh is not 2:1 balanced or not covering the domain")

    def _neighbors(self, q):
        L = self.level[q]
        lx = self.loc[q]
        myox = [2 * (lx[d] & 1) - 1 for d in range(3)]
        out = []

        def add(ox, kind, t, fi1=0, fi2=0, child=None):
            if kind == "same":
                nid, nl = self.leaf[(L,) + t], L
            elif kind == "coarser":
                nid, nl = self.leaf[(L - 1, t[0] >> 1, t[1] >> 1, t[2] >> 1)], L - 1
            else:
                nid, nl = self.leaf[(L + 1,) + child], L + 1
            out.append(dict(ox=ox, fi1=fi1, fi2=fi2, level=nl, nid=nid))

        def child(t, c):
            return (2 * t[0] + c[0], 2 * t[1] + c[1], 2 * t[2] + c[2])

        # faces: x1, x2, x3; finer children f2 outer, f1 inner
        for axis in range(3):
            for n in (-1, 1):
                ox = [0, 0, 0]
                ox[axis] = n
                t = self._wrap(L, tuple(lx[d] + ox[d] for d in range(3)))
                if t is None:
                    continue
                kind = self._covered(L, t)
                if kind == "finer":
                    ff = 1 - (n + 1) // 2
                    others = [d for d in range(3) if d != axis]
                    for f2 in (0, 1):
                        for f1 in (0, 1):
                            c = [0, 0, 0]
                            c[axis] = ff
                            c[others[0]] = f1
                            c[others[1]] = f2
                            add(tuple(ox), kind, t, f1, f2, child(t, c))
                else:
                    add(tuple(ox), kind, t)
        # edges: (a, b) = (x1, x2), (x1, x3), (x2, x3); m over b, n over a
        for a, b in ((0, 1), (0, 2), (1, 2)):
            free = 3 - a - b
            for m in (-1, 1):
                for n in (-1, 1):
                    ox = [0, 0, 0]
                    ox[a], ox[b] = n, m
                    t = self._wrap(L, tuple(lx[d] + ox[d] for d in range(3)))
                    if t is None:
                        continue
                    kind = self._covered(L, t)
                    if kind == "finer":
                        for f1 in (0, 1):
                            c = [0, 0, 0]
                            c[a] = 1 - (n + 1) // 2
                            c[b] = 1 - (m + 1) // 2
                            c[free] = f1
                            add(tuple(ox), kind, t, f1, 0, child(t, c))
                    elif kind == "same" or (myox[a] == n and myox[b] == m):
                        add(tuple(ox), kind, t)
        # corners: l over x3, m over x2, n over x1
        for l in (-1, 1):
            for m in (-1, 1):
                for n in (-1, 1):
                    ox = (n, m, l)
                    t = self._wrap(L, tuple(lx[d] + ox[d] for d in range(3)))
                    if t is None:
                        continue
                    kind = self._covered(L, t)
                    if kind == "finer":
                        c = tuple(1 - (o + 1) // 2 for o in ox)
                        add(ox, kind, t, 0, 0, child(t, c))
                    elif kind == "same" or (myox[0] == n and myox[1] == m and myox[2] == l):
                        add(ox, kind, t)
        return out

    # -----------------------------------------------------------------------
    def coarse_shape(self):
        return tuple(self.nx[d] // 2 + 2 * CNG + 1 for d in (2, 1, 0))

    def coarse_geom(self, q):
        """oracle.Geom of block q's coarse buffer with iu = cive + 1, so that
        oracle.physical_bc's cell-centred end iu - 1 is the coarse vertex
        range's end cive, as ApplyPhysicalBCsOnCoarseLevel passes it."""
        b = self.blocks[q]
        g = orc.Geom()
        g.nn1, g.nn2, g.nn3 = (self.nx[d] // 2 + 2 * CNG + 1 for d in range(3))
        g.ng = CNG
        g.il, g.iu = CNG, CNG + self.nx[0] // 2 + 1
        g.jl, g.ju = CNG, CNG + self.nx[1] // 2 + 1
        g.kl, g.ku = CNG, CNG + self.nx[2] // 2 + 1
        for d in range(3):
            g.dx[d] = 2 * b.dx[d]
            g.x0[d] = b.xmin[d] - CNG * 2 * b.dx[d]
        return g

    def ghost_fill(self, us, coarse=None, faces=None):
        """One ghost update, in place on the fine arrays ``us`` (list of
        (nvar, nn3, nn2, nn1)); returns the coarse buffers (list of (nvar,
        cn3, cn2, cn1)).  ``coarse`` supplies the coarse buffers' values
        before the update (their ghosts are zeroed; default NaN).  ``faces``
        (per block six flags ix1, ox1, ix2, ox2, ix3, ox3, or None) marks
        the faces with the physical fill, applied to the coarse buffer of
        every block that prolongates."""
        nb = len(self.blocks)
        nvar = us[0].shape[0]
        fa = [_axis_fine(self.nx[d]) for d in range(3)]
        ca = [_axis_coarse(self.nx[d]) for d in range(3)]
        half_f = [self.nx[d] // 2 for d in range(3)]
        half_l = [self.nx[d] // 2 - CNG for d in range(3)]
        cs = [np.full((nvar,) + self.coarse_shape(), np.nan) if coarse is None else coarse[q].copy()
              for q in range(nb)]
        # 1. restriction of every block's interior
        for q in range(nb):
            r = orc.restrict_vc(self.blocks[q].geom, us[q])
            ci = tuple(slice(ca[d]["vs"], ca[d]["ve"] + 1) for d in (2, 1, 0))
            cs[q][(slice(None),) + ci] = r[(slice(None),) + ci]
        # 2. messages from pre-unpack values: msg[(receiver, index in its list)]
        msgs = {}
        for q in range(nb):
            L = self.level[q]
            for j, nb_ in enumerate(self.nbrs[q]):
                src = nb_["nid"]
                # the receiver q as seen from the sender: reversed offset
                rev = tuple(-o for o in nb_["ox"])
                if nb_["level"] == L:
                    sl = tuple(slice(*_incl(load_same(rev[d], fa[d]))) for d in (2, 1, 0))
                    fine = us[src][(slice(None),) + sl].copy()
                    cpay = None
                    if not self.all_same[q]:
                        sl = tuple(slice(*_incl(load_same(rev[d], ca[d]))) for d in (2, 1, 0))
                        cpay = cs[src][(slice(None),) + sl].copy()
                    msgs[(q, j)] = (fine, cpay)
                elif nb_["level"] < L:
                    # coarser sender: LoadToFiner with the sender's entry for q
                    ent = self._entry(src, q, rev)
                    sl = tuple(slice(*_incl(load_to_finer(rev[d], fa[d], _fi_axis(ent, d),
                                                          half_l[d]))) for d in (2, 1, 0))
                    msgs[(q, j)] = us[src][(slice(None),) + sl].copy()
                else:
                    sl = tuple(slice(*_incl(load_to_coarser(rev[d], ca[d]))) for d in (2, 1, 0))
                    msgs[(q, j)] = cs[src][(slice(None),) + sl].copy()
        # 3. unpack and divide
        for q in range(nb):
            L = self.level[q]
            lx = self.loc[q]
            u, c = us[q], cs[q]
            _zero_ghosts(u, fa)
            _zero_ghosts(c, ca)
            mult = np.zeros((7, 7, 7), dtype=int)
            mult[1:6, 1:6, 1:6] = 1
            for j, nb_ in enumerate(self.nbrs[q]):
                ox = nb_["ox"]
                if nb_["level"] == L:
                    fine, cpay = msgs[(q, j)]
                    sl = tuple(slice(*_incl(set_same(ox[d], fa[d]))) for d in (2, 1, 0))
                    u[(slice(None),) + sl] += fine
                    if not self.all_same[q]:
                        sl = tuple(slice(*_incl(set_same(ox[d], ca[d]))) for d in (2, 1, 0))
                        c[(slice(None),) + sl] += cpay
                    cr = [_compact("same", ox[d]) for d in range(3)]
                elif nb_["level"] < L:
                    sl = tuple(slice(*_incl(set_from_coarser(ox[d], ca[d], lx[d]))) for d in (2, 1, 0))
                    c[(slice(None),) + sl] = msgs[(q, j)]
                    cr = [_compact("coarser", ox[d], lx[d]) for d in range(3)]
                else:
                    sl = tuple(slice(*_incl(set_from_finer(ox[d], fa[d], _fi_axis(nb_, d),
                                                           half_f[d]))) for d in (2, 1, 0))
                    u[(slice(None),) + sl] += msgs[(q, j)]
                    cr = [_compact("finer", ox[d], fi=_fi_axis(nb_, d)) for d in range(3)]
                mult[cr[2][0]:cr[2][1] + 1, cr[1][0]:cr[1][1] + 1, cr[0][0]:cr[0][1] + 1] += 1
            _divide(u, mult, fa)
            _divide(c, mult, ca)
        # 4. the coarse level's physical fill, then prolongation from every
        # coarser neighbour
        for q in range(nb):
            L = self.level[q]
            if faces is not None and any(faces[q]) and any(n["level"] < L for n in self.nbrs[q]):
                orc.physical_bc(self.coarse_geom(q), cs[q], faces[q])
            for nb_ in self.nbrs[q]:
                if nb_["level"] >= L:
                    continue
                rng = orc.prolongation_range(self.blocks[q].geom, nb_["ox"], self.loc[q])
                orc.prolongate_vc(self.blocks[q].geom, cs[q], us[q], rng)
        return cs

    def _entry(self, q, r, ox):
        """Block r's entry in q's neighbour list at offset ox."""
        for e in self.nbrs[q]:
            if e["nid"] == r and e["ox"] == tuple(ox):
                return e
        raise KeyError((q, r, ox))


def _incl(se):
    return se[0], se[1] + 1
