"""Density accumulation on the GPU against the oracle's compute_density (scf/density.py:35-84)."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_density_matches_oracle(si2_system):
    from amaterasu_amd.density import compute_density, electron_count
    from oracle.density import compute_density as ocd
    from oracle.eigen import random_bands

    s = si2_system
    wf = [[random_bands(len(sp), s.nbands, sp.kinetic_energy, seed=7 + ik) for ik, sp in enumerate(s.wave_spheres)]]
    occ = np.zeros((1, 2, s.nbands))
    occ[:, :, :4] = 1.0
    occ[0, 1, 4] = 0.3  # a fractional occupation
    occ[0, 0, 5] = 1e-13  # below the 1e-12 activity cut (density.py:64-67)
    got, ref = compute_density(s, wf, occ), ocd(s, wf, occ)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12
    assert electron_count(s, got) == pytest.approx(8.0 + 2 * 0.75 * 0.3, abs=1e-9)


def test_density_fp32_orbitals(si2_system):
    import torch

    from amaterasu_amd.density import compute_density
    from oracle.density import compute_density as ocd
    from oracle.eigen import random_bands

    s = si2_system
    wf = [[random_bands(len(sp), s.nbands, sp.kinetic_energy, seed=ik) for ik, sp in enumerate(s.wave_spheres)]]
    occ = np.zeros((1, 2, s.nbands))
    occ[:, :, :4] = 1.0
    wf32 = [[torch.from_numpy(w).to("cuda", torch.complex64) for w in wf[0]]]
    got, ref = compute_density(s, wf32, occ), ocd(s, wf, occ)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-5


def test_density_generic_grid(si2_system):
    """Non-cubic grid: the generic density kernels against the oracle."""
    import copy

    from amaterasu_amd.density import compute_density
    from oracle.basis import FFTGrid
    from oracle.density import compute_density as ocd
    from oracle.eigen import random_bands

    s = copy.copy(si2_system)
    s.grid = FFTGrid((24, 20, 22), si2_system.cell)
    wf = [[random_bands(len(sp), s.nbands, sp.kinetic_energy, seed=3 + ik) for ik, sp in enumerate(s.wave_spheres)]]
    occ = np.zeros((1, 2, s.nbands))
    occ[:, :, :5] = 0.8
    got, ref = compute_density(s, wf, occ), ocd(s, wf, occ)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12
