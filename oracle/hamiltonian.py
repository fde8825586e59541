"""The Kohn-Sham operator and the ionic fields on the grid (FP64, CPU).

Restates ``amaterasu/hamiltonian/operator.py``, ``kinetic.py`` and ``local.py``.
"""

from __future__ import annotations

import numpy as np

from .pseudo import density_form_factor, local_potential_form_factor


def apply_kinetic(coefficients, sphere):
    """``kinetic.py:22-24``."""
    return np.asarray(coefficients) * sphere.kinetic_energy


def band_kinetic_energies(coefficients, sphere):
    """``kinetic.py:27-30``."""
    c = np.atleast_2d(np.asarray(coefficients))
    return np.einsum("ng,g->n", np.abs(c) ** 2, sphere.kinetic_energy)


class Hamiltonian:
    """H = T + V_eff + V_NL at one fixed potential (``operator.py:37-95``)."""

    def __init__(self, system, effective_potential, coupling=None):
        v = np.asarray(effective_potential, dtype=float)
        if v.ndim == 3:
            v = v[None, ...]
        expected = (system.nspin, *system.grid.shape)
        if v.shape != expected:
            raise ValueError(f"effective potential has shape {v.shape}, expected {expected}")
        self.system = system
        self.effective_potential = v
        self.coupling = coupling

    def apply(self, coefficients, ik, spin=0):
        """``operator.py:60-69``: T c + gather(FFT(V . IFFT(scatter c))) + V_NL c."""
        sphere = self.system.wave_spheres[ik]
        grid = self.system.grid
        out = apply_kinetic(coefficients, sphere)
        real = grid.wave_to_real(coefficients, sphere)
        out = out + grid.real_to_wave(self.effective_potential[spin] * real, sphere)
        d = None if self.coupling is None else self.coupling[spin]
        return out + self.system.projectors[ik].apply(coefficients, d)

    def apply_overlap(self, coefficients, ik):
        return np.asarray(coefficients)

    @property
    def overlap_operator(self):
        return None

    def local_energy(self, density) -> float:
        """``operator.py:90-95``."""
        return float(np.sum(self.effective_potential * np.asarray(density, dtype=float))
                     * self.system.grid.cell_volume_per_point)


def structure_factors(sphere, positions):
    """exp(-i G.tau_a), shape (natoms, npw) (``local.py:37-39``)."""
    return np.exp(-1j * (np.atleast_2d(positions) @ sphere.cartesian.T))


def _superpose(grid, sphere, positions, symbols, radial):
    """``local.py:42-55``."""
    phases = structure_factors(sphere, positions)
    c = np.zeros(len(sphere), dtype=complex)
    for atom, sym in enumerate(symbols):
        c += radial[sym] * phases[atom]
    c /= grid.volume
    return grid.to_real(grid.scatter(c, sphere))


def _real(field, what):
    """``local.py:159-168``."""
    scale = float(np.max(np.abs(field.real)))
    resid = float(np.max(np.abs(field.imag)))
    if scale > 0 and resid > 1e-8 * scale:
        raise ValueError(f"the {what} came out complex ({resid:.3e} against {scale:.3e})")
    return np.ascontiguousarray(field.real)


def local_potential(grid, sphere, positions, symbols, pseudopotentials):
    """``local.py:58-74``."""
    q = np.sqrt(sphere.g2)
    radial = {s: local_potential_form_factor(p, q) for s, p in pseudopotentials.items()}
    return _real(_superpose(grid, sphere, positions, symbols, radial), "local potential")


def atomic_valence_density(grid, sphere, positions, symbols, pseudopotentials, *, total_electrons=None):
    """``local.py:77-106``: superposed atomic densities, floor 1e-12, rescaled to N_el."""
    q = np.sqrt(sphere.g2)
    radial = {}
    for s, p in pseudopotentials.items():
        if p.valence_density is None:
            raise ValueError(f"{s}: the dataset carries no atomic valence density")
        radial[s] = density_form_factor(p.grid, p.valence_density, q)
    rho = _real(_superpose(grid, sphere, positions, symbols, radial), "valence density")
    rho = np.maximum(rho, 1e-12)
    if total_electrons is not None:
        rho *= total_electrons / (float(np.sum(rho)) * grid.cell_volume_per_point)
    return rho


def atomic_core_density(grid, sphere, positions, symbols, pseudopotentials):
    """``local.py:109-143`` (not clipped)."""
    if not any(p.has_core_correction for p in pseudopotentials.values()):
        return None
    q = np.sqrt(sphere.g2)
    radial = {}
    for s, p in pseudopotentials.items():
        radial[s] = np.zeros_like(q) if p.core_density is None else density_form_factor(p.grid, p.core_density, q)
    return _real(_superpose(grid, sphere, positions, symbols, radial), "core density")
