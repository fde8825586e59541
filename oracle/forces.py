"""Forces on the ions, norm-conserving part (FP64 numpy restatement; test infrastructure).

Follows ``properties/forces.py:48-176`` (``compute_forces`` with ``_local_forces``,
``_core_forces``, ``_nonlocal_forces``, ``_real``) and the force half of
``hamiltonian/ewald.py:94-225``, plus ``symmetrize_vectors``
(``system/symmetry.py:238-307``). Every term is a Hellmann-Feynman derivative at fixed
density / wavefunctions, so each is pinned by a central finite difference of the energy
term it differentiates (``tests/test_oracle_forces.py``).
"""

from __future__ import annotations

import numpy as np
from scipy.special import erfc

from .hamiltonian import structure_factors
from .pseudo import density_form_factor, local_potential_form_factor


def ewald_forces(cell, positions, charges, eta=None, tolerance: float = 1e-14) -> np.ndarray:
    """``ewald.py:161-206``: real-space pair derivative and reciprocal structure-factor term."""
    cell = np.asarray(cell, dtype=float)
    pos = np.atleast_2d(np.asarray(positions, dtype=float))
    z = np.atleast_1d(np.asarray(charges, dtype=float))
    vol = abs(float(np.linalg.det(cell)))
    eta = 1.8 / vol ** (1.0 / 3.0) if eta is None else eta  # ewald.py optimal_eta
    ztot = float(z.sum())
    forces = np.zeros((len(z), 3))
    sep = pos[:, None, :] - pos[None, :, :]
    fr = sep @ np.linalg.inv(cell)
    sep = (fr - np.rint(fr)) @ cell
    recip = 2.0 * np.pi * np.linalg.inv(cell).T
    rc = 1.0 / eta
    while erfc(eta * rc) / rc >= tolerance / max(1.0, ztot**2):
        rc *= 1.2
    reps = np.ceil(rc * np.linalg.norm(recip, axis=1) / (2.0 * np.pi)).astype(int) + 1
    offs = np.stack(np.meshgrid(*[np.arange(-n, n + 1) for n in reps], indexing="ij"), -1).reshape(-1, 3) @ cell
    zz = z[:, None] * z[None, :]
    for o in offs:
        vec = sep + o
        d = np.linalg.norm(vec, axis=-1)
        small = d < 1e-10
        safe = np.where(small, 1.0, d)
        pair = np.where(small, 0.0, zz * erfc(eta * safe) / safe)
        if np.max(np.abs(pair)) < tolerance and np.min(safe) > rc:
            continue
        der = np.where(small, 0.0, -zz * (erfc(eta * safe) / safe**2
                                          + 2.0 * eta / np.sqrt(np.pi) * np.exp(-((eta * safe) ** 2)) / safe))
        forces -= np.einsum("ab,abi,ab->ai", der, vec, 1.0 / safe)
    gc = 2.0 * eta
    while 2.0 * np.pi / vol * max(ztot**2, 1.0) * np.exp(-(gc**2) / (4.0 * eta**2)) / gc**2 >= tolerance:
        gc *= 1.2
    b = np.ceil(gc * np.linalg.norm(cell, axis=1) / (2.0 * np.pi)).astype(int)
    m = np.stack(np.meshgrid(*[np.arange(-n, n + 1) for n in b], indexing="ij"), -1).reshape(-1, 3)
    g = m @ recip
    g2 = np.einsum("ij,ij->i", g, g)
    keep = (g2 > 1e-12) & (g2 <= gc**2)
    g, g2 = g[keep], g2[keep]
    ph = np.exp(1j * (g @ pos.T))
    s = ph @ z
    w = np.exp(-g2 / (4.0 * eta**2)) / g2
    forces += 4.0 * np.pi / vol * z[:, None] * np.einsum("g,ga,gi->ai", w, np.imag(s.conj()[:, None] * ph), g)
    return forces


def _real(vector, what):
    """``forces.py:178-186``."""
    residual = float(np.max(np.abs(vector.imag)))
    scale = max(float(np.max(np.abs(vector.real))), 1e-30)
    if residual > 1e-6 * scale and residual > 1e-12:
        raise ValueError(f"the {what} came out complex ({residual:.3e} against {scale:.3e})")
    return vector.real


def local_forces(system, density) -> np.ndarray:
    """``forces.py:111-126``: F_a = -sum_G i G v_a(|G|) e^{iG.tau_a} n(G)."""
    grid, sphere = system.grid, system.density_sphere
    total = np.asarray(density, dtype=float).sum(axis=0)
    ng = grid.gather(grid.to_reciprocal(total.astype(complex)), sphere)
    ph = structure_factors(sphere, system.positions)
    q = np.sqrt(sphere.g2)
    ff = {s: local_potential_form_factor(p, q) for s, p in system.pseudopotentials.items()}
    out = np.zeros((len(system.symbols), 3))
    for a, sym in enumerate(system.symbols):
        w = 1j * ff[sym] * ph[a].conj() * ng
        out[a] = -_real(np.einsum("g,gi->i", w, sphere.cartesian), "local force")
    return out


def core_forces(system, density) -> np.ndarray:
    """``forces.py:96-105,129-143``: the partial core moving with its atom, against the
    spin-averaged v_xc of n + n_core / nspin."""
    if system.core_density is None:
        return np.zeros((len(system.symbols), 3))
    grid, sphere = system.grid, system.density_sphere
    d = np.asarray(density, dtype=float)
    _, vxc = system.xc.evaluate_on_grid(d + system.core_density[None] / system.nspin, grid)
    vg = grid.gather(grid.to_reciprocal(vxc.mean(axis=0).astype(complex)), sphere)
    ph = structure_factors(sphere, system.positions)
    q = np.sqrt(sphere.g2)
    out = np.zeros((len(system.symbols), 3))
    for a, sym in enumerate(system.symbols):
        p = system.pseudopotentials[sym]
        if p.core_density is None:
            continue
        cg = density_form_factor(p.grid, p.core_density, q)
        w = 1j * vg.conj() * cg * ph[a]
        out[a] = _real(np.einsum("g,gi->i", w, sphere.cartesian), "core force")
    return out


def nonlocal_forces(system, wavefunctions, occupations) -> np.ndarray:
    """``forces.py:146-175``: derivative of the KB energy through the projector phases."""
    deg = 2.0 / system.nspin
    out = np.zeros((len(system.symbols), 3))
    for spin in range(system.nspin):
        for ik, sphere in enumerate(system.wave_spheres):
            pr = system.projectors[ik]
            if len(pr) == 0:
                continue
            atom_of_row = np.array([lab[0] for lab in pr.labels])
            block = np.asarray(wavefunctions[spin][ik])
            w = deg * system.kpoints.weights[ik] * np.asarray(occupations)[spin, ik]
            act = w > 1e-12
            if not np.any(act):
                continue
            block, w = block[act], w[act]
            coupled = pr.project(block) @ pr.dij
            for axis in range(3):
                der = (block * (1j * sphere.cartesian[:, axis])) @ pr.matrix.conj().T
                terms = 2.0 * np.real(coupled.conj() * der)
                np.add.at(out[:, axis], atom_of_row, -np.einsum("n,nj->j", w, terms))
    return out


def symmetrize_vectors(vectors, fractional, species, symmetry, cell) -> np.ndarray:
    """``symmetry.py:238-307``: F_a = (1/N) sum_S F_{S(a)} M_S^-1, M_S = cell^-1 R^T cell."""
    vectors = np.asarray(vectors, dtype=float)
    if symmetry is None or len(symmetry.rotations) <= 1:
        return vectors.copy()
    cell = np.asarray(cell, dtype=float)
    frac = np.asarray(fractional, dtype=float)
    numbers = np.array([hash(s) for s in species])
    ops = np.einsum("ij,sjk,kl->sil", np.linalg.inv(cell), np.asarray(symmetry.rotations, dtype=float)
                    .transpose(0, 2, 1), cell)
    out = np.zeros_like(vectors)
    for r, t, mat in zip(symmetry.rotations, symmetry.translations, ops):
        images = frac @ np.asarray(r, dtype=float).T + t
        diff = images[:, None, :] - frac[None, :, :]
        diff -= np.rint(diff)
        dist = np.linalg.norm(diff, axis=-1)
        dist[numbers[:, None] != numbers[None, :]] = np.inf
        target = np.argmin(dist, axis=1)
        out += vectors[target] @ np.linalg.inv(mat)
    return out / len(symmetry.rotations)


def compute_forces(system, density, wavefunctions, occupations, *, symmetrize=True, decompose=False):
    """``forces.py:48-108`` (norm-conserving branch). Ha/Bohr, shape (natoms, 3)."""
    charges = np.array([system.pseudopotentials[s].valence_charge for s in system.symbols])
    parts = {"ewald": ewald_forces(system.cell, system.positions, charges),
             "local": local_forces(system, density),
             "nonlocal": nonlocal_forces(system, wavefunctions, occupations)}
    if system.core_density is not None:
        parts["core"] = core_forces(system, density)
    if symmetrize and system.symmetry is not None:
        frac = system.positions @ np.linalg.inv(system.cell)
        parts = {k: symmetrize_vectors(v, frac, system.symbols, system.symmetry, system.cell)
                 for k, v in parts.items()}
    return parts if decompose else sum(parts.values())
