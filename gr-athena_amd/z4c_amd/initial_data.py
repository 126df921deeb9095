"""Synthetic initial data and block lists for the Z4c workloads (numpy, host).

* perturbed flat space on one block (sum of seeded Gaussian bumps per
  variable) -- the single-block RHS workload;
* the 1-D gauge wave (src/z4c/awa_z4c.cpp:298-351);
* Brill-Lindquist two-puncture data with a pre-collapsed lapse
  alpha = psi^-2 (src/z4c/gauge.cpp:46-62), zero shift, K = 0;
* a seeded octree of MeshBlocks refined toward the punctures, with mixed
  block sizes -- the batched uneven-block workload.

Arrays are (nvar, nn3, nn2, nn1) in the reference's AthenaArray order.
"""
from __future__ import annotations

import math

import numpy as np

from .z4c import NGHOST, MeshBlockGeom

NADM = 17


def _mesh(geom):
    x, y, z = geom.coords()
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def perturbed_flat(geom: MeshBlockGeom, seed: int = 0, nbumps: int = 8,
                   amp=(1e-6, 1e-3), width=(0.1, 0.3)):
    """Flat space + 8 Gaussian bumps per Z4c variable: centres uniform in the
    block, widths uniform in [0.1,0.3] of the box, amplitudes uniform in
    [1e-6,1e-3] with random sign; chi = 1+d, alpha = 1+d, g~_ij = delta_ij +
    d_ij.  AlgConstr is NOT applied here (apply it on the device)."""
    rng = np.random.default_rng(seed)
    X, Y, Z = _mesh(geom)
    lo = np.array(geom.xmin)
    L = np.array(geom.nx) * np.array(geom.dx)
    u = np.zeros((22,) + geom.shape)
    for n in range(22):
        acc = np.zeros(geom.shape)
        for _ in range(nbumps):
            c = lo + rng.random(3) * L
            w = rng.uniform(*width) * L.max()
            a = rng.uniform(*amp) * rng.choice([-1.0, 1.0])
            acc += a * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (w * w))
        u[n] = acc
    u[0] += 1.0            # chi
    u[1] += 1.0            # g~xx
    u[4] += 1.0            # g~yy
    u[6] += 1.0            # g~zz
    u[18] += 1.0           # alpha
    return u


def gauge_wave_adm(geom: MeshBlockGeom, A: float = 0.01, d: float = 1.0):
    """ADM data of ds^2 = -H dt^2 + H dx^2 + dy^2 + dz^2, H = 1 - A sin(2 pi x/d)
    at t = 0 (ADMGaugeWave1 + GaugeGaugeWave1, src/z4c/awa_z4c.cpp:298-351)."""
    # the wave is formed on the x line with the C library's sin / cos and
    # broadcast, so that every (y, z) row holds the same bits and the data do
    # not depend on the host's SIMD code paths
    x, _, _ = geom.coords()
    sw = np.array([A * math.sin(2.0 * math.pi * (v / d)) for v in x])
    dsw = np.array([-A * 2.0 * math.pi * math.cos(2.0 * math.pi * (v / d)) for v in x])
    adm = np.zeros((NADM,) + geom.shape)
    adm[4] = 1.0 - sw                      # gxx
    adm[7] = 1.0                           # gyy
    adm[9] = 1.0                           # gzz
    adm[10] = 0.5 * dsw / np.sqrt(1.0 - sw)  # Kxx
    adm[16] = 1.0                          # psi4
    adm[0] = np.sqrt(1.0 - sw)             # alpha
    return adm


# Puncture offset: half the finest spacing of the octree (h0 = 1, level 7),
# so that no vertex of any level coincides with a puncture.
PUNCTURE_OFFSET = 0.5 * 2.0 ** -7


def brill_lindquist_adm(geom: MeshBlockGeom, m=(0.5, 0.5), sep: float = 10.0,
                        offset: float = PUNCTURE_OFFSET):
    """Two punctures at (+-sep/2, 0, 0) + offset*(1,1,1): psi = 1 + m1/(2 r1)
    + m2/(2 r2), gamma = psi^4 delta, K = 0, alpha = psi^-2, beta = 0."""
    X, Y, Z = _mesh(geom)
    X, Y, Z = X - offset, Y - offset, Z - offset
    r1 = np.sqrt((X - sep / 2) ** 2 + Y ** 2 + Z ** 2)
    r2 = np.sqrt((X + sep / 2) ** 2 + Y ** 2 + Z ** 2)
    psi = 1.0 + m[0] / (2.0 * r1) + m[1] / (2.0 * r2)
    psi4 = psi ** 4
    adm = np.zeros((NADM,) + geom.shape)
    adm[4] = adm[7] = adm[9] = psi4
    adm[16] = psi4
    adm[0] = psi4 ** -0.5
    return adm


def octree_blocks(nblocks: int = 2048, seed: int = 0, h0: float = 1.0,
                  max_level: int = 7, sizes=(16, 24, 32), mix=(0.7, 0.2, 0.1),
                  punctures=((5.0, 0.0, 0.0), (-5.0, 0.0, 0.0)), base_cells: int = 16):
    """Seeded octree of ``nblocks`` leaves refined toward the punctures.

    Start from a 4x4x4 root grid of level-0 cubes of side base_cells*h0 and
    repeatedly split (into 8) the leaf closest to a puncture, relative to its
    size, choosing among near-ties with the seed, until the leaf count is
    reached (levels 0..max_level).  Each leaf then gets a block size drawn
    from ``sizes`` with probabilities ``mix`` (seeded) and spacing
    h0*2^-level, so the block spans size*h of its cube from the cube corner.
    Blocks of different sizes do not tile the domain; the list is a batch of
    uneven, independent blocks."""
    rng = np.random.default_rng(seed)
    side0 = base_cells * h0
    leaves = []
    for i in range(4):
        for j in range(4):
            for k in range(4):
                leaves.append((0, np.array([i - 2.0, j - 2.0, k - 2.0]) * side0))
    P = np.array(punctures)
    while len(leaves) < nblocks:
        scores = []
        for lev, corner in leaves:
            if lev >= max_level:
                scores.append(np.inf)
                continue
            side = side0 * 0.5 ** lev
            c = corner + side / 2
            d = np.min(np.linalg.norm(P - c, axis=1))
            scores.append(d / side + 0.25 * rng.random())
        idx = int(np.argmin(scores))
        if not np.isfinite(scores[idx]):
            break
        lev, corner = leaves.pop(idx)
        half = side0 * 0.5 ** (lev + 1)
        for a in range(2):
            for b in range(2):
                for c in range(2):
                    leaves.append((lev + 1, corner + np.array([a, b, c]) * half))
    leaves = leaves[:nblocks]
    choice = rng.choice(len(sizes), size=len(leaves), p=np.array(mix) / np.sum(mix))
    out = []
    for (lev, corner), s in zip(leaves, choice):
        h = h0 * 0.5 ** lev
        out.append(MeshBlockGeom.cube(sizes[s], h, corner, level=lev))
    return out


def balanced_octree_blocks(nblocks: int = 2048, seed: int = 0, h0: float = 1.0,
                           max_level: int = 7, cells: int = 16,
                           punctures=((5.0, 0.0, 0.0), (-5.0, 0.0, 0.0)), nroot: int = 4):
    """Block tree of equal ``cells``^3 blocks refined toward the punctures,
    2:1 balanced across faces, edges and corners, covering the open box of
    nroot^3 level-0 blocks of side cells*h0 (centred on the origin), as the
    reference's mesh refinement builds it (a block is split only after its
    neighbours at its level exist, src/mesh/meshblock_tree.cpp).

    Leaves are split in the order of octree_blocks (closest to a puncture
    relative to size, near-ties broken by the seed) until at least
    ``nblocks`` exist; no leaf is dropped, so the blocks tile the box.
    Returns (blocks, lo, hi); faces on the box boundary carry the outflow
    physical fill (extrapolate_outflow)."""
    rng = np.random.default_rng(seed)
    side0 = cells * h0
    lo = -0.5 * nroot * side0
    P = np.array(punctures)
    leaves = {(0, i, j, k) for i in range(nroot) for j in range(nroot) for k in range(nroot)}

    def covering(L, t):
        """The leaf covering logical cell t at level L (or a finer region)."""
        for up in range(0, L + 1):
            key = (L - up,) + tuple(v >> up for v in t)
            if key in leaves:
                return key
        return None

    def split(key):
        L, i, j, k = key
        # 2:1 balance: every neighbour region at level L must be a leaf of
        # level >= L before the children (level L + 1) appear
        for o in np.ndindex(3, 3, 3):
            if o == (1, 1, 1):
                continue
            t = (i + o[0] - 1, j + o[1] - 1, k + o[2] - 1)
            if any(not 0 <= v < (nroot << L) for v in t):
                continue
            nb = covering(L, t)
            if nb is not None and nb[0] < L:
                split(nb)
        leaves.discard(key)
        for a in range(2):
            for b in range(2):
                for c in range(2):
                    leaves.add((L + 1, 2 * i + a, 2 * j + b, 2 * k + c))

    while len(leaves) < nblocks:
        best, best_key = np.inf, None
        for key in sorted(leaves):
            if key[0] >= max_level:
                continue
            side = side0 * 0.5 ** key[0]
            c = lo + (np.array(key[1:]) + 0.5) * side
            sc = np.min(np.linalg.norm(P - c, axis=1)) / side + 0.25 * rng.random()
            if sc < best:
                best, best_key = sc, key
        if best_key is None:
            break
        split(best_key)
    out = []
    for (L, i, j, k) in sorted(leaves):
        side = side0 * 0.5 ** L
        n = nroot << L
        faces = 0
        for d, t in enumerate((i, j, k)):
            faces |= (1 << (2 * d)) if t == 0 else 0
            faces |= (2 << (2 * d)) if t == n - 1 else 0
        out.append(MeshBlockGeom.cube(cells, side / cells, lo + np.array([i, j, k]) * side,
                                      level=L, outflow=faces))
    return out, (lo,) * 3, (-lo,) * 3


def owned_points(blocks):
    """Zone count of the reference's counter: nx1*nx2*nx3 per block per cycle
    (src/main.hpp:729-731)."""
    return int(sum(g.nx[0] * g.nx[1] * g.nx[2] for g in blocks))
