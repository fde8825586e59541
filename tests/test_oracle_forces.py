"""Pins of the oracle's force terms: each is the Hellmann-Feynman derivative of one energy
term at fixed density / wavefunctions, so each must equal minus the central finite
difference of that term as one atom moves (the reference's own term-by-term test
strategy, ``properties/forces.py`` docstring). A dropped factor, a sign or a transposed
phase fails one of them."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import si_primitive_bohr

H = 2e-4  # Bohr


@pytest.fixture(scope="module")
def displaced(si_pseudo):
    from oracle.system import build_system
    from oracle.xc import XCFunctional

    cell, pos = si_primitive_bohr()
    pos = pos + np.array([[0.0, 0.0, 0.0], [0.21, -0.13, 0.08]])
    s = build_system(cell, pos, ["Si", "Si"], {"Si": si_pseudo}, 8.0, 32.0, [[0.25, 0.25, 0.25], [-0.25, 0.25, 0.25]],
                     [0.25, 0.75])
    s.xc = XCFunctional("PBE")
    return s


def _moved(positions, atom, axis, h):
    p = np.array(positions, dtype=float)
    p[atom, axis] += h
    return p


def test_ewald_force_is_minus_the_energy_gradient(displaced):
    from oracle.ewald import ewald_energy
    from oracle.forces import ewald_forces

    z = np.array([4.0, 4.0])
    f = ewald_forces(displaced.cell, displaced.positions, z)
    for atom, axis in ((1, 0), (1, 2), (0, 1)):
        ep = ewald_energy(displaced.cell, _moved(displaced.positions, atom, axis, H), z)
        em = ewald_energy(displaced.cell, _moved(displaced.positions, atom, axis, -H), z)
        assert f[atom, axis] == pytest.approx(-(ep - em) / (2 * H), abs=2e-7)
    assert np.abs(f.sum(axis=0)).max() < 1e-9  # Newton's third law


def test_local_force_is_minus_the_local_energy_gradient(displaced, si_pseudo):
    from oracle.forces import local_forces
    from oracle.hamiltonian import local_potential

    s = displaced
    rho = s.initial_density()
    f = local_forces(s, rho)
    dv = s.grid.cell_volume_per_point

    def energy(pos):
        v = local_potential(s.grid, s.density_sphere, pos, s.symbols, {"Si": si_pseudo})
        return float(np.sum(v * rho[0])) * dv

    for atom, axis in ((1, 0), (1, 1), (0, 2)):
        d = (energy(_moved(s.positions, atom, axis, H)) - energy(_moved(s.positions, atom, axis, -H))) / (2 * H)
        assert f[atom, axis] == pytest.approx(-d, abs=1e-6)


def test_core_force_is_minus_the_xc_energy_gradient(displaced, si_pseudo):
    from oracle.forces import core_forces
    from oracle.hamiltonian import atomic_core_density

    s = displaced
    if s.core_density is None:
        pytest.skip("dataset has no partial core")
    rho = s.initial_density()
    f = core_forces(s, rho)

    def energy(pos):
        core = atomic_core_density(s.grid, s.density_sphere, pos, s.symbols, {"Si": si_pseudo})
        e, _ = s.xc.evaluate_on_grid(rho + core[None], s.grid)
        return e

    for atom, axis in ((1, 0), (1, 2)):
        d = (energy(_moved(s.positions, atom, axis, H)) - energy(_moved(s.positions, atom, axis, -H))) / (2 * H)
        assert f[atom, axis] == pytest.approx(-d, abs=1e-6)


def test_nonlocal_force_is_minus_the_kb_energy_gradient(displaced, si_pseudo):
    from oracle.eigen import random_bands
    from oracle.forces import nonlocal_forces
    from oracle.pseudo import build_projectors

    s = displaced
    wfs = [[random_bands(len(sp.g2), 6, sp.g2 / 2, seed=ik) for ik, sp in enumerate(s.wave_spheres)]]
    occ = np.zeros((1, 2, 6))
    occ[0, :, :4] = 1.0
    occ[0, 1, 4] = 0.3
    f = nonlocal_forces(s, wfs, occ)

    def energy(pos):
        e = 0.0
        for ik, sp in enumerate(s.wave_spheres):
            pr = build_projectors(sp, pos, s.symbols, {"Si": si_pseudo}, s.volume)
            e += pr.energy(wfs[0][ik], 2.0 * s.kpoints.weights[ik] * occ[0, ik])
        return e

    for atom, axis in ((1, 0), (1, 1), (0, 2)):
        d = (energy(_moved(s.positions, atom, axis, H)) - energy(_moved(s.positions, atom, axis, -H))) / (2 * H)
        assert f[atom, axis] == pytest.approx(-d, abs=2e-7)


def test_symmetrized_vectors_vanish_on_diamond_sites(si2_system):
    """Both diamond sites have point group Td: the average of any vector over it is zero."""
    from oracle.forces import symmetrize_vectors
    from oracle.symmetry import find_symmetry

    s = si2_system
    frac = s.positions @ np.linalg.inv(s.cell)
    sym = find_symmetry(s.cell, frac, s.symbols)
    v = np.random.default_rng(0).normal(size=(2, 3))
    assert np.abs(symmetrize_vectors(v, frac, s.symbols, sym, s.cell)).max() < 1e-12
    assert np.array_equal(symmetrize_vectors(v, frac, s.symbols, None, s.cell), v)
