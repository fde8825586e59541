"""Stress tensor, norm-conserving part (FP64 numpy restatement; test infrastructure).

Follows ``properties/stress.py:63-290`` (``compute_stress`` with the kinetic, Hartree,
local, exchange-correlation, core and nonlocal terms and ``_projector_q_gradient``), the
stress half of ``hamiltonian/ewald.py:94-228`` and ``symmetrize_tensor``
(``system/symmetry.py:310-324``). Convention: sigma_uv = (1/Omega) dE/d eps_uv for the
strain r -> (1 + eps) r, in Ha/Bohr^3, with the plane-wave coefficients held fixed on
their Miller indices. Each term is pinned by a central finite difference of the energy
term it differentiates under a symmetric strain (``tests/test_oracle_stress.py``).
"""

from __future__ import annotations

import numpy as np
from scipy.special import erfc

from .basis import miller_bounds, reciprocal_cell
from .hamiltonian import structure_factors
from .pseudo import (
    density_form_factor_derivative,
    local_potential_form_factor,
    local_potential_form_factor_derivative,
    projector_form_factor,
    projector_form_factor_derivative,
    real_spherical_harmonics,
    solid_harmonics_gradient,
)
from .scf import hartree


def ewald_stress(cell, positions, charges, eta=None, tolerance: float = 1e-14) -> np.ndarray:
    """``ewald.py:128-228``: pair term dr/d eps_uv = r_u r_v / r, reciprocal term
    dG^2/d eps_uv = -2 G_u G_v with the structure factor invariant, and the 1/Omega
    scaling of the reciprocal sum and the neutralising background."""
    cell = np.asarray(cell, dtype=float)
    pos = np.atleast_2d(np.asarray(positions, dtype=float))
    z = np.atleast_1d(np.asarray(charges, dtype=float))
    vol = abs(float(np.linalg.det(cell)))
    eta = 1.8 / vol ** (1.0 / 3.0) if eta is None else eta  # ewald.py optimal_eta
    ztot = float(z.sum())
    out = np.zeros((3, 3))

    rcut = 1.0 / eta
    tol_r = tolerance / max(1.0, ztot**2)
    while not erfc(eta * rcut) / rcut < tol_r:
        rcut *= 1.2
    sep = pos[:, None, :] - pos[None, :, :]
    fr = sep @ np.linalg.inv(cell)
    fr -= np.rint(fr)
    sep = fr @ cell
    recip = reciprocal_cell(cell)
    reps = np.ceil(rcut * np.linalg.norm(recip, axis=1) / (2.0 * np.pi)).astype(int) + 1
    offs = np.stack(np.meshgrid(*[np.arange(-n, n + 1) for n in reps], indexing="ij"), axis=-1).reshape(-1, 3) @ cell
    zz = z[:, None] * z[None, :]
    for off in offs:
        v = sep + off
        d = np.linalg.norm(v, axis=-1)
        small = d < 1e-10
        safe = np.where(small, 1.0, d)
        pair = np.where(small, 0.0, zz * erfc(eta * safe) / safe)
        if np.max(np.abs(pair)) < tolerance and np.min(safe) > rcut:
            continue
        # d/dr [erfc(eta r)/r] times Z_a Z_b, over r for the r_u r_v / r chain factor
        slope = np.where(small, 0.0, -zz * (erfc(eta * safe) / safe**2
                                            + 2.0 * eta / np.sqrt(np.pi) * np.exp(-(eta * safe) ** 2) / safe))
        out += 0.5 * np.einsum("ab,abu,abv->uv", slope / safe, v, v)

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
    s2 = np.abs(np.exp(1j * (g @ pos.T)) @ z) ** 2
    w = np.exp(-g2 / (4.0 * eta**2)) / g2
    recip_sum = float(2.0 * np.pi / vol * np.sum(s2 * w))
    out += 2.0 * np.pi / vol * np.einsum("g,gu,gv->uv", s2 * w * (1.0 / (2.0 * eta**2) + 2.0 / g2), g, g)
    out -= np.eye(3) * recip_sum
    background = -np.pi / (2.0 * vol * eta**2) * ztot**2
    out -= np.eye(3) * background
    return out / vol


def kinetic_stress(system, wavefunctions, occupations) -> np.ndarray:
    """``stress.py:114-123``: -(1/Omega) sum w f |c|^2 (k+G)_u (k+G)_v."""
    deg = 2.0 / system.nspin
    out = np.zeros((3, 3))
    for spin in range(system.nspin):
        for ik, sphere in enumerate(system.wave_spheres):
            w = deg * system.kpoints.weights[ik] * occupations[spin, ik]
            amp = w @ (np.abs(np.asarray(wavefunctions[spin][ik])) ** 2)
            out -= np.einsum("g,gu,gv->uv", amp, sphere.cartesian, sphere.cartesian)
    return out / system.grid.volume


def hartree_stress(system, total_density, density_g) -> np.ndarray:
    """``stress.py:126-137``: -E_H delta/Omega + sum_G 4 pi |n(G)|^2 G_u G_v / G^4."""
    energy, _ = hartree(system.grid, system.density_sphere, total_density)
    sp = system.density_sphere
    nz = sp.g2 > 1e-12
    w = 4.0 * np.pi * np.abs(density_g[nz]) ** 2 / sp.g2[nz] ** 2
    vec = sp.cartesian[nz]
    return -np.eye(3) * energy / system.grid.volume + np.einsum("g,gu,gv->uv", w, vec, vec)


def _inverse_norm(sphere):
    q = np.sqrt(sphere.g2)
    inv = np.zeros_like(q)
    inv[q > 1e-12] = 1.0 / q[q > 1e-12]
    return q, inv


def local_stress(system, density_g) -> np.ndarray:
    """``stress.py:139-155``: -E_loc delta - sum_a sum_G Re(v_a'(q)/q S_a^* n(G)) G_u G_v,
    over Omega."""
    sp = system.density_sphere
    q, inv = _inverse_norm(sp)
    phases = structure_factors(sp, system.positions)
    energy = 0.0
    tensor = np.zeros((3, 3))
    for atom, sym in enumerate(system.symbols):
        ps = system.pseudopotentials[sym]
        form = local_potential_form_factor(ps, q)
        slope = local_potential_form_factor_derivative(ps, q)
        energy += float(np.sum((form * phases[atom].conj() * density_g).real))
        w = (slope * inv * phases[atom].conj() * density_g).real
        tensor -= np.einsum("g,gu,gv->uv", w, sp.cartesian, sp.cartesian)
    return (-np.eye(3) * energy + tensor) / system.grid.volume


def xc_stress(system, xc_density, energy, potential, gradients, vsigma) -> np.ndarray:
    """``stress.py:158-186``: delta (E_xc - sum_s int v_s n_s) minus, for a GGA, the
    gradient outer products weighted by vsigma (uu twice, ud both orders, dd twice)."""
    grid = system.grid
    dv = grid.cell_volume_per_point
    out = np.eye(3) * (energy - float(np.sum(potential * xc_density)) * dv)
    if vsigma is not None:
        def contract(w, a, b):
            return np.einsum("ijk,ijku,ijkv->uv", w, a, b)

        outer = (2.0 * contract(vsigma[0], gradients[0], gradients[0])
                 + contract(vsigma[1], gradients[0], gradients[1])
                 + contract(vsigma[1], gradients[1], gradients[0])
                 + 2.0 * contract(vsigma[2], gradients[1], gradients[1]))
        out -= outer * dv
    return out / grid.volume


def core_stress(system, potential) -> np.ndarray:
    """``stress.py:189-214``: the rigid partial core, -sum_a sum_G Re(V(G)^* rho_c'(q)/q S_a)
    G_u G_v / Omega with V the spin-average of v_xc."""
    grid, sp = system.grid, system.density_sphere
    avg = np.asarray(potential).mean(axis=0)
    vg = grid.gather(grid.to_reciprocal(avg.astype(complex)), sp)
    q, inv = _inverse_norm(sp)
    phases = structure_factors(sp, system.positions)
    tensor = np.zeros((3, 3))
    for atom, sym in enumerate(system.symbols):
        ps = system.pseudopotentials[sym]
        if ps.core_density is None:
            continue
        slope = density_form_factor_derivative(ps.grid, ps.core_density, q)
        w = (vg.conj() * slope * inv * phases[atom]).real
        tensor -= np.einsum("g,gu,gv->uv", w, sp.cartesian, sp.cartesian)
    return tensor / grid.volume


def projector_q_gradient(system, sphere) -> np.ndarray:
    """``stress.py:293-330``: dP_i/dq_alpha, shape (3, nproj, npw), rows in the order of
    ``build_projectors``. With P = (4 pi/sqrt Omega) (-i)^l C_lm(q) f_l(q) q^-l e^{-iq.tau}:
    dC/dq f/q^l - l Y f q/q^2 + Y f' q/q, times the same prefactor and phase."""
    qv = sphere.cartesian
    q, inv = _inverse_norm(sphere)
    pref = 4.0 * np.pi / np.sqrt(system.grid.volume)
    rows = []
    for atom, sym in enumerate(system.symbols):
        ps = system.pseudopotentials[sym]
        phase = np.exp(-1j * (qv @ system.positions[atom]))
        for p in ps.projectors:
            l = p.l
            f = projector_form_factor(ps.grid, p, q)
            fp = projector_form_factor_derivative(ps.grid, p, q)
            y = real_spherical_harmonics(l, qv)
            dc = solid_harmonics_gradient(l, qv)
            power = inv**l if l else np.ones_like(q)
            for mi in range(2 * l + 1):
                ang = (dc[mi] * (f * power)[:, None] - l * (y[mi] * f * inv**2)[:, None] * qv
                       + (y[mi] * fp * inv)[:, None] * qv)
                rows.append(pref * (-1j) ** l * ang * phase[:, None])
    return np.asarray(rows).transpose(2, 0, 1)


def nonlocal_stress(system, wavefunctions, occupations) -> np.ndarray:
    """``stress.py:217-255``: R^(mu nu) = sum_n w_n Re(sum_i conj((P D)_ni)
    sum_G c_nG conj(dP_i/dq_mu) q_nu); sigma = -(R + R^T + E_nl delta)/Omega."""
    deg = 2.0 / system.nspin
    out = np.zeros((3, 3))
    energy = 0.0
    for ik, sphere in enumerate(system.wave_spheres):
        pr = system.projectors[ik]
        if len(pr) == 0:
            continue
        grad = projector_q_gradient(system, sphere)
        qv = sphere.cartesian
        for spin in range(system.nspin):
            w = deg * system.kpoints.weights[ik] * occupations[spin, ik]
            act = w > 1e-12
            if not np.any(act):
                continue
            block = np.asarray(wavefunctions[spin][ik])[act]
            w = w[act]
            proj = pr.project(block)
            coupled = proj @ pr.dij
            energy += float(np.einsum("n,ni,ni->", w, coupled, proj.conj()).real)
            for mu in range(3):
                partial = np.einsum("ng,ig,gv->niv", block, grad[mu].conj(), qv)
                out[mu] -= np.einsum("n,ni,niv->v", w, coupled.conj(), partial).real
    out = out + out.T
    return (out - np.eye(3) * energy) / system.grid.volume


def symmetrize_tensor(tensor, symmetry, cell) -> np.ndarray:
    """``symmetry.py:310-324``: (1/N) sum_S M_S^T sigma M_S, M_S = cell^-1 R^T cell."""
    tensor = np.asarray(tensor, dtype=float)
    if symmetry is None or len(symmetry.rotations) <= 1:
        return tensor.copy()
    cell = np.asarray(cell, dtype=float)
    ops = np.einsum("ij,sjk,kl->sil", np.linalg.inv(cell),
                    np.asarray(symmetry.rotations, dtype=float).transpose(0, 2, 1), cell)
    return sum(m.T @ tensor @ m for m in ops) / len(ops)


def compute_stress(system, density, wavefunctions, occupations, *, symmetrize=True, decompose=False):
    """``stress.py:63-107`` (norm-conserving branch). Ha/Bohr^3, shape (3, 3)."""
    grid, sp = system.grid, system.density_sphere
    density = np.asarray(density, dtype=float)
    total = density.sum(axis=0)
    ng = grid.gather(grid.to_reciprocal(total.astype(complex)), sp)
    charges = np.array([system.pseudopotentials[s].valence_charge for s in system.symbols])
    parts = {"ewald": ewald_stress(system.cell, system.positions, charges),
             "kinetic": kinetic_stress(system, wavefunctions, occupations),
             "hartree": hartree_stress(system, total, ng),
             "local": local_stress(system, ng),
             "nonlocal": nonlocal_stress(system, wavefunctions, occupations)}
    xd = density
    if system.core_density is not None:
        xd = density + system.core_density[None] / system.nspin
    exc, vxc, grads, vsig = system.xc.evaluate_on_grid_details(xd, grid)
    parts["xc"] = xc_stress(system, xd, exc, vxc, grads, vsig)
    if system.core_density is not None:
        parts["core"] = core_stress(system, vxc)
    if symmetrize and system.symmetry is not None:
        parts = {k: symmetrize_tensor(v, system.symmetry, system.cell) for k, v in parts.items()}
    return parts if decompose else sum(parts.values())
