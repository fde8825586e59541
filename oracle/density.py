"""Electron density from the occupied orbitals (FP64, CPU).

Restates ``amaterasu/scf/density.py:30-84``.
"""

from __future__ import annotations

import numpy as np


def electron_count(system, density) -> float:
    """``density.py:30-32``."""
    return float(np.sum(density)) * system.grid.cell_volume_per_point


def compute_density(system, wavefunctions, occupations, *, symmetrize=True):
    """n_s(r) = (1/Omega) sum_k w_k sum_n (2/nspin) f |u(r)|^2, then symmetrise, clip at 0
    and rescale each channel to its pre-clip charge (``density.py:35-84``)."""
    grid = system.grid
    g = 2.0 / system.nspin
    rho = np.zeros((system.nspin, *grid.shape))
    for s in range(system.nspin):
        for ik, (sphere, w) in enumerate(zip(system.wave_spheres, system.kpoints.weights)):
            f = g * w * occupations[s, ik]
            act = f > 1e-12
            if not np.any(act):
                continue
            u = grid.wave_to_real(wavefunctions[s][ik][act], sphere)
            rho[s] += np.einsum("n,n...->...", f[act], np.abs(u) ** 2)
    rho /= grid.volume
    if symmetrize:
        sym = system.symmetrize_density(rho)
        before = sym.sum(axis=(-3, -2, -1))
        clipped = np.maximum(sym, 0.0)
        after = clipped.sum(axis=(-3, -2, -1))
        scale = np.where(after > 0, before / np.where(after > 0, after, 1.0), 1.0)
        rho = clipped * scale[:, None, None, None]
    return rho
