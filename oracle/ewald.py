"""Ion-ion Ewald energy (FP64, CPU). Restates the energy part of
``amaterasu/hamiltonian/ewald.py:52-228``."""

from __future__ import annotations

import numpy as np
from scipy.special import erfc

from .basis import miller_bounds, reciprocal_cell


def optimal_eta(volume: float) -> float:
    """``ewald.py:52-59``."""
    return 1.8 / volume ** (1.0 / 3.0)


def ewald_energy(cell, positions, charges, *, eta=None, tolerance=1e-14) -> float:
    """Real-space erfc sum over folded pair separations + reciprocal Gaussian sum + Gaussian
    self energy + neutralising background (``ewald.py:128-225``)."""
    cell = np.asarray(cell, dtype=float)
    pos = np.atleast_2d(np.asarray(positions, dtype=float))
    z = np.atleast_1d(np.asarray(charges, dtype=float))
    vol = abs(float(np.linalg.det(cell)))
    eta = optimal_eta(vol) if eta is None else eta
    ztot = float(z.sum())

    # real-space cutoff (ewald.py:62-69)
    rcut = 1.0 / eta
    tol_r = tolerance / max(1.0, ztot**2)
    while not erfc(eta * rcut) / rcut < tol_r:
        rcut *= 1.2
    sep = pos[:, None, :] - pos[None, :, :]
    frac = sep @ np.linalg.inv(cell)
    frac -= np.rint(frac)
    sep = frac @ cell
    recip = reciprocal_cell(cell)
    reps = np.ceil(rcut * np.linalg.norm(recip, axis=1) / (2.0 * np.pi)).astype(int) + 1
    offs = np.stack(np.meshgrid(*[np.arange(-n, n + 1) for n in reps], indexing="ij"), axis=-1).reshape(-1, 3) @ cell
    zz = z[:, None] * z[None, :]
    real = 0.0
    for off in offs:
        d = np.linalg.norm(sep + off, axis=-1)
        small = d < 1e-10
        safe = np.where(small, 1.0, d)
        pair = np.where(small, 0.0, zz * erfc(eta * safe) / safe)
        if np.max(np.abs(pair)) < tolerance and np.min(safe) > rcut:
            continue
        real += float(pair.sum())
    real *= 0.5

    # reciprocal cutoff (ewald.py:72-79)
    gcut = 2.0 * eta
    q2 = max(ztot**2, 1.0)
    while not 2.0 * np.pi / vol * q2 * np.exp(-(gcut**2) / (4.0 * eta**2)) / gcut**2 < tolerance:
        gcut *= 1.2
    b = miller_bounds(cell, gcut)
    m = np.stack(np.meshgrid(*[np.arange(-n, n + 1) for n in b], indexing="ij"), axis=-1).reshape(-1, 3)
    g = m @ recip
    g2 = np.einsum("ij,ij->i", g, g)
    keep = (g2 > 1e-12) & (g2 <= gcut**2)
    g, g2 = g[keep], g2[keep]
    s = np.exp(1j * (g @ pos.T)) @ z
    recip_sum = float(2.0 * np.pi / vol * np.sum(np.abs(s) ** 2 * np.exp(-g2 / (4.0 * eta**2)) / g2))
    self_e = -eta / np.sqrt(np.pi) * float(np.sum(z**2))
    background = -np.pi / (2.0 * vol * eta**2) * ztot**2
    return real + recip_sum + self_e + background
