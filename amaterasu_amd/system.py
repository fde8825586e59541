"""A device-side counterpart of amaterasu's ``PlaneWaveSystem`` for the H.psi path.

Holds the attributes the operator, the eigensolver and the density read
(``hamiltonian/system.py:62-94``): the FFT grid, one sphere per k-point, the projector
matrices (device tensors), the k-point weights, ``nspin`` and ``nbands``. Built from
explicit k-points (the reference's spglib reduction is outside this path); a P1 cell
needs no density symmetrisation (``system.py:198-199``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .basis import FFTGrid, GSphere
from .pseudo import build_projectors, read_psp8

DATA = Path(__file__).resolve().parent / "data"


def default_band_count(nelectrons: float, nspin: int = 1, moment: float = 0.0) -> int:
    """``hamiltonian/system.py:48-59``: 20 % extra, never fewer than 4 extra."""
    occ = 0.5 * nelectrons + (0.5 * abs(moment) if nspin == 2 else 0.0)
    return max(math.ceil(occ) + 4, math.ceil(1.2 * occ))


@dataclass
class KPoints:
    points: np.ndarray
    weights: np.ndarray

    def __len__(self) -> int:
        return len(self.points)


@dataclass
class DeviceSystem:
    cell: np.ndarray
    positions: np.ndarray
    symbols: list
    grid: FFTGrid
    wave_spheres: list
    projectors: list
    kpoints: KPoints
    nspin: int = 1
    nbands: int = 0
    nelectrons: float = 0.0
    symmetry: object = None
    paw: object = None
    extra: dict = field(default_factory=dict)

    @property
    def volume(self) -> float:
        return self.grid.volume

    @property
    def nkpoints(self) -> int:
        return len(self.kpoints)


def load_dataset(name: str = "Si.pbe.psp8.gz"):
    return read_psp8(DATA / name)


def build_system(cell, positions, symbols, datasets: dict, ecut_rho: float, kpoints, ecuts, *, weights=None,
                 nbands=None, nspin: int = 1, with_projectors: bool = True, grid_shape=None) -> DeviceSystem:
    """Grid from the density cutoff (``system.py:262-271``), one sphere per k at its own
    cutoff, projector matrices on the device."""
    cell = np.asarray(cell, dtype=float)
    kpts = np.atleast_2d(np.asarray(kpoints, dtype=float))
    ecuts = np.broadcast_to(np.asarray(ecuts, dtype=float), (len(kpts),))
    grid = FFTGrid.for_cutoff(cell, ecut_rho) if grid_shape is None else FFTGrid(tuple(grid_shape), cell)
    spheres = [GSphere.build(cell, float(e), kpoint=k) for e, k in zip(ecuts, kpts)]
    vol = grid.volume
    projs = [build_projectors(s, positions, symbols, datasets, vol) for s in spheres] if with_projectors else None
    w = np.full(len(kpts), 1.0 / len(kpts)) if weights is None else np.asarray(weights, dtype=float)
    nel = float(sum(datasets[s].valence_charge for s in symbols))
    nb = nbands or default_band_count(nel, nspin)
    return DeviceSystem(cell=cell, positions=np.asarray(positions, dtype=float), symbols=list(symbols), grid=grid,
                        wave_spheres=spheres, projectors=projs, kpoints=KPoints(kpts, w), nspin=nspin, nbands=nb,
                        nelectrons=nel)


__all__ = ["DeviceSystem", "KPoints", "build_system", "default_band_count", "load_dataset"]
