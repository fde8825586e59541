"""Space-group operations and the irreducible k-point set without spglib (FP64, CPU).

The reference takes both from spglib (``amaterasu/system/symmetry.py:135-206``,
``get_ir_reciprocal_mesh``), which is absent here. This module finds the operations
directly: every integer matrix R with entries in {-1, 0, 1} that preserves the lattice
metric, paired with a fractional translation t that maps the crystal onto itself
(x -> R x + t, spglib's convention). The irreducible mesh follows the semantics of
``symmetry.py:135-206``: k-points (address + shift/2)/mesh, orbits under k -> R^T k and
time reversal, the lowest grid address (x fastest) as representative, weights equal to
orbit sizes, coordinates folded into (-1/2, 1/2].

Parity: the operation set is pinned by its size (48 for diamond, 1 for a generic
displaced cell) and the k-point set by the closed form of the shifted 2x2x2 diamond
mesh (two points, weights 1/4 and 3/4); any other k-point list is parity unpinned
(spglib's own ordering cannot be checked here).
"""

from __future__ import annotations

import itertools

import numpy as np

from .system import Symmetry


def find_symmetry(cell, fractional, species, tol: float = 1e-5) -> Symmetry:
    cell = np.asarray(cell, dtype=float)
    frac = np.asarray(fractional, dtype=float) % 1.0
    species = list(species)
    metric = cell @ cell.T
    rots, trans = [], []
    first = [j for j, s in enumerate(species) if s == species[0]]
    for entries in itertools.product((-1, 0, 1), repeat=9):
        r = np.array(entries, dtype=int).reshape(3, 3)
        if abs(round(np.linalg.det(r))) != 1:
            continue
        if not np.allclose(r.T @ metric @ r, metric, atol=tol * max(1.0, np.abs(metric).max())):
            continue
        moved = (frac @ r.T) % 1.0
        for j in first:
            t = (frac[j] - moved[0]) % 1.0
            ok = True
            for i, si in enumerate(species):
                d = (moved[i] + t - frac) % 1.0
                d = np.minimum(d, 1.0 - d)
                hit = np.flatnonzero(np.all(d < tol, axis=1))
                if not any(species[h] == si for h in hit):
                    ok = False
                    break
            if ok:
                t = np.where(np.abs(t - 1.0) < tol, 0.0, t)
                rots.append(r)
                trans.append(t)
                break
    order = sorted(range(len(rots)), key=lambda i: (not np.array_equal(rots[i], np.eye(3, dtype=int)),))
    return Symmetry(np.array([rots[i] for i in order]), np.array([trans[i] for i in order]))


def ibz_kpoints(symmetry: Symmetry, mesh, shift=(0.0, 0.0, 0.0), time_reversal: bool = True):
    """Irreducible k-points and weights of a Monkhorst-Pack mesh (``symmetry.py:135-206``)."""
    mesh = np.asarray(mesh, dtype=int)
    is_shift = np.array([1 if abs(s) > 1e-12 else 0 for s in shift])
    addr = np.stack(np.meshgrid(*[np.arange(n) for n in mesh], indexing="ij"), axis=-1)
    addr = addr.transpose(2, 1, 0, 3).reshape(-1, 3)  # x fastest
    addr = addr - mesh * (addr > mesh // 2)  # fold into (-m/2, m/2]
    doubled = 2 * addr + is_shift  # k = doubled / (2 mesh)
    key = {tuple(int(v) % int(2 * m) for v, m in zip(d, mesh)): i for i, d in enumerate(doubled)}
    ops = [r.T for r in symmetry.rotations]
    mapping = np.arange(len(doubled))
    for i, d in enumerate(doubled):
        best = i
        for rt in ops:
            for sign in ((1, -1) if time_reversal else (1,)):
                img = sign * (rt @ d)
                j = key.get(tuple(int(v) % int(2 * m) for v, m in zip(img, mesh)))
                if j is not None and j < best:
                    best = j
        mapping[i] = best
    uniq, counts = np.unique(mapping, return_counts=True)
    k = doubled[uniq] / (2.0 * mesh)
    k -= np.rint(k)
    return k, counts / counts.sum()
