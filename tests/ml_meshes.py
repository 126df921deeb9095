"""Refined block meshes for the multi-level ghost-fill tests: a root grid of
equally sized blocks on [lo, hi]^3, blocks refined (each into 8 children)
wherever a predicate of the block's centre and level holds, then 2:1
balance enforced over the 26 neighbour directions (periodic axes wrap)."""
import itertools

import numpy as np


def refined_blocks(nrb, nx, lo, hi, refine, levels, periodic=(True, True, True)):
    """Leaves as (nx, dx, xmin) triples; ``refine(centre, level)`` marks the
    blocks to split, level by level up to ``levels`` - 1."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    size0 = (hi - lo) / nrb
    leaves = {(0, i, j, k) for i in range(nrb) for j in range(nrb) for k in range(nrb)}

    def centre(key):
        L, *t = key
        s = size0 / 2 ** L
        return lo + (np.asarray(t) + 0.5) * s

    def split(key):
        L, i, j, k = key
        leaves.discard(key)
        for a, b, c in itertools.product((0, 1), repeat=3):
            leaves.add((L + 1, 2 * i + a, 2 * j + b, 2 * k + c))

    for L in range(levels - 1):
        for key in sorted(k for k in leaves if k[0] == L):
            if refine(centre(key), L):
                split(key)
        _balance(leaves, nrb, periodic, split)
    out = []
    for (L, i, j, k) in sorted(leaves):
        s = size0 / 2 ** L
        out.append(((nx,) * 3, tuple(s / nx), tuple(lo + np.array([i, j, k]) * s)))
    return out


def _balance(leaves, nrb, periodic, split):
    """Refine leaves that touch a leaf two or more levels finer."""
    changed = True
    while changed:
        changed = False
        for key in sorted(leaves, key=lambda x: -x[0]):
            L, *t = key
            for o in itertools.product((-1, 0, 1), repeat=3):
                if o == (0, 0, 0):
                    continue
                n = [t[d] + o[d] for d in range(3)]
                ok = True
                for d in range(3):
                    m = nrb << L
                    if not 0 <= n[d] < m:
                        if not periodic[d]:
                            ok = False
                        n[d] %= m
                if not ok:
                    continue
                # the leaf covering n at level L - 2 or coarser must be split
                for up in range(2, L + 1):
                    anc = (L - up,) + tuple(v >> up for v in n)
                    if anc in leaves:
                        split(anc)
                        changed = True
                        break
                if changed:
                    break
            if changed:
                break


def quintic(x, y, z):
    """A tensor product of quintics (reproduced exactly by the VC
    prolongation) plus lower-order terms."""
    return (1.0 + 0.3 * x - 0.2 * y + 0.1 * z + 0.05 * x * y * z
            + 0.01 * (x - 0.3) ** 5 * (y + 0.2) ** 4 * (z - 0.1) ** 3)


def quartic(x, y, z):
    """Degree <= 4 along every axis: reproduced by the 5-point extrapolation
    of the physical fill as well as by the VC prolongation."""
    return (1.0 + 0.3 * x - 0.2 * y + 0.1 * z + 0.05 * x * y * z
            + 0.01 * (x - 0.3) ** 4 * (y + 0.2) ** 3 * (z - 0.1) ** 2)


def open_faces(mesh, q):
    """Six flags (ix1, ox1, ix2, ox2, ix3, ox3): block q's faces on an open
    face of the domain."""
    out = []
    for d in range(3):
        m = mesh.nrb0[d] << mesh.level[q]
        out += [(not mesh.periodic[d]) and mesh.loc[q][d] == 0,
                (not mesh.periodic[d]) and mesh.loc[q][d] == m - 1]
    return tuple(bool(v) for v in out)
