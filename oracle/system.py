"""The fixed part of a plane-wave calculation, built without ase/spglib (FP64, CPU).

Restates the norm-conserving branch of ``amaterasu/hamiltonian/system.py:215-379``
(``build_system``), ``default_band_count`` (``system.py:48-59``),
``initial_density`` (``system.py:147-190``, nspin=1 path) and
``symmetrize_density`` (``system.py:192-212``). The reference obtains the k-point set
and the space group from spglib; spglib is absent, so the caller passes the k-points,
weights and (optionally) symmetry operations explicitly. A P1 cell needs none, and
``symmetrize_density`` is then the identity exactly as in the reference
(``system.py:198-199``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .basis import FFTGrid, GSphere, symmetry_compatible_multiple
from .hamiltonian import atomic_core_density, atomic_valence_density, local_potential
from .pseudo import build_projectors


def default_band_count(nelectrons: float, nspin: int, moment: float = 0.0) -> int:
    """``system.py:48-59``."""
    occ = 0.5 * nelectrons
    if nspin == 2:
        occ += 0.5 * abs(moment)
    return max(math.ceil(occ) + 4, math.ceil(1.2 * occ))


@dataclass
class KPointSet:
    points: np.ndarray
    weights: np.ndarray

    def __len__(self) -> int:
        return len(self.points)


@dataclass
class Symmetry:
    """Space-group operations in fractional coordinates: r -> R r + t."""

    rotations: np.ndarray
    translations: np.ndarray

    @property
    def n_operations(self) -> int:
        return len(self.rotations)


def symmetrize_sphere(coefficients, sphere, grid, rotations, translations):
    """n_sym(G) = (1/N_op) sum_S n(R_S^T G) exp(-2 pi i G.f_S) (``symmetry.py:327-370``)."""
    c = np.asarray(coefficients)
    rotations = np.asarray(rotations, dtype=int)
    translations = np.asarray(translations, dtype=float)
    if len(rotations) <= 1:
        return c.copy()
    lookup = np.full(grid.shape, -1, dtype=np.int64)
    lookup[grid.indices(sphere)] = np.arange(len(sphere))
    m = sphere.miller
    acc = np.zeros_like(c)
    for rot, tr in zip(rotations, translations):
        img = m @ rot
        idx = lookup[tuple(img[:, a] % grid.shape[a] for a in range(3))]
        if np.any(idx < 0):
            raise ValueError("a symmetry operation maps a G-vector outside the cutoff sphere")
        acc += c[idx] * np.exp(-2j * np.pi * (m @ tr))
    return acc / len(rotations)


@dataclass
class PlaneWaveSystem:
    cell: np.ndarray
    positions: np.ndarray
    symbols: list
    pseudopotentials: dict
    ecut: float
    ecut_rho: float
    grid: FFTGrid
    density_sphere: GSphere
    kpoints: KPointSet
    wave_spheres: list
    projectors: list
    nspin: int
    nbands: int
    nelectrons: float
    v_local: np.ndarray
    core_density: np.ndarray | None
    symmetry: Symmetry | None = None
    ewald: float = 0.0
    xc: object = None
    paw: object = None
    warnings: list = field(default_factory=list)

    @property
    def is_paw(self) -> bool:
        return False

    @property
    def volume(self) -> float:
        return self.grid.volume

    @property
    def nkpoints(self) -> int:
        return len(self.kpoints)

    def initial_density(self):
        """``system.py:147-161`` (spin-unpolarised)."""
        total = atomic_valence_density(self.grid, self.density_sphere, self.positions, self.symbols,
                                       self.pseudopotentials, total_electrons=self.nelectrons)
        if self.nspin != 1:
            raise NotImplementedError("the oracle restates the spin-unpolarised start only")
        return total[None, ...]

    def symmetrize_density(self, density):
        """``system.py:192-212``."""
        if self.symmetry is None or self.symmetry.n_operations <= 1:
            return density
        out = np.empty_like(density)
        for s in range(density.shape[0]):
            c = self.grid.gather(self.grid.to_reciprocal(density[s].astype(complex)), self.density_sphere)
            avg = symmetrize_sphere(c, self.density_sphere, self.grid, self.symmetry.rotations,
                                    self.symmetry.translations)
            out[s] = self.grid.to_real(self.grid.scatter(avg, self.density_sphere)).real
        return out


def build_system(cell, positions, symbols, pseudopotentials, ecut, ecut_rho, kpoints, weights=None, *,
                 nbands=None, nspin=1, symmetry: Symmetry | None = None, xc=None, wave_ecuts=None,
                 with_ewald=True):
    """Norm-conserving ``build_system`` (``system.py:215-379``) with explicit k-points.

    ``cell`` rows and ``positions`` are in Bohr. ``wave_ecuts`` optionally gives a
    per-k wavefunction cutoff (a synthetic-workload device used to make the plane-wave
    counts ragged; the reference always uses one ``ecut``).
    """
    cell = np.asarray(cell, dtype=float)
    positions = np.asarray(positions, dtype=float)
    kpts = np.atleast_2d(np.asarray(kpoints, dtype=float))
    w = np.full(len(kpts), 1.0 / len(kpts)) if weights is None else np.asarray(weights, dtype=float)
    nelectrons = float(sum(pseudopotentials[s].valence_charge for s in symbols))
    multiple = symmetry_compatible_multiple(symmetry.translations) if symmetry is not None else 1
    grid = FFTGrid(shape=FFTGrid.for_cutoff(cell, ecut_rho, multiple=multiple).shape, cell=cell)
    dsphere = GSphere.build(cell, ecut_rho)
    ecuts = [ecut] * len(kpts) if wave_ecuts is None else list(wave_ecuts)
    spheres = [GSphere.build(cell, e, kpoint=k) for e, k in zip(ecuts, kpts)]
    nb = nbands or default_band_count(nelectrons, nspin)
    nb = min(nb, min(len(s) for s in spheres))
    projectors = [build_projectors(s, positions, symbols, pseudopotentials, grid.volume) for s in spheres]
    vloc = local_potential(grid, dsphere, positions, symbols, pseudopotentials)
    ewald = 0.0
    if with_ewald:
        from .ewald import ewald_energy

        ewald = ewald_energy(cell, positions, np.array([pseudopotentials[s].valence_charge for s in symbols]))
    return PlaneWaveSystem(
        cell=cell, positions=positions, symbols=list(symbols), pseudopotentials=pseudopotentials,
        ecut=ecut, ecut_rho=ecut_rho, grid=grid, density_sphere=dsphere,
        kpoints=KPointSet(kpts, w), wave_spheres=spheres, projectors=projectors, nspin=nspin,
        nbands=nb, nelectrons=nelectrons, v_local=vloc,
        core_density=atomic_core_density(grid, dsphere, positions, symbols, pseudopotentials),
        symmetry=symmetry, ewald=ewald, xc=xc,
    )
