"""Oracle operator, local potential and density, pinned by the reference's closed forms
(tests/unit/test_hamiltonian.py, test_scf.py electron count)."""

from __future__ import annotations

import copy

import numpy as np
import pytest

from oracle.density import compute_density, electron_count
from oracle.eigen import davidson, random_bands
from oracle.hamiltonian import Hamiltonian, structure_factors
from oracle.pseudo import KBProjectors, local_potential_form_factor


def test_hamiltonian_is_hermitian(si2_system):
    rng = np.random.default_rng(0)
    h = Hamiltonian(si2_system, rng.normal(size=(1, *si2_system.grid.shape)))
    n = len(si2_system.wave_spheres[1])
    b = rng.normal(size=(4, n)) + 1j * rng.normal(size=(4, n))
    m = b.conj() @ h.apply(b, 1).T
    assert m == pytest.approx(m.conj().T, abs=1e-11)


def test_empty_lattice_gives_free_electron_bands(si2_system):
    # test_hamiltonian.py:221-244
    s = copy.copy(si2_system)
    s.projectors = [KBProjectors(np.zeros((0, len(sp)), dtype=complex), np.zeros((0, 0)), []) for sp in s.wave_spheres]
    h = Hamiltonian(s, np.zeros((1, *s.grid.shape)))
    for ik, sp in enumerate(s.wave_spheres):
        r = davidson(lambda b, ik=ik: h.apply(b, ik), random_bands(len(sp), 10, sp.kinetic_energy, seed=ik),
                     sp.kinetic_energy, tolerance=1e-12, max_iterations=200)
        assert r.eigenvalues == pytest.approx(np.sort(sp.kinetic_energy)[:10], abs=1e-11)


def test_wrong_potential_shape_rejected(si2_system):
    with pytest.raises(ValueError, match="effective potential has shape"):
        Hamiltonian(si2_system, np.zeros((2, *si2_system.grid.shape)))


def test_local_potential_mean_is_alpha(si2_system, si_pseudo):
    # test_hamiltonian.py:143-158
    alpha = 2.0 * local_potential_form_factor(si_pseudo, np.array([0.0]))[0]
    assert float(np.mean(si2_system.v_local)) == pytest.approx(alpha / si2_system.grid.volume, rel=1e-10)


def test_structure_factors(si2_system):
    f = structure_factors(si2_system.density_sphere, si2_system.positions)
    assert np.abs(np.abs(f) - 1).max() < 1e-12
    assert f[:, 0] == pytest.approx(np.ones(2), abs=1e-12)


def test_initial_density_and_density_charge(si2_system):
    rho0 = si2_system.initial_density()
    assert electron_count(si2_system, rho0) == pytest.approx(8.0, rel=1e-10)
    wf = [[random_bands(len(sp), si2_system.nbands, sp.kinetic_energy, seed=ik)
           for ik, sp in enumerate(si2_system.wave_spheres)]]
    occ = np.zeros((1, 2, si2_system.nbands))
    occ[:, :, :4] = 1.0
    rho = compute_density(si2_system, wf, occ)
    assert electron_count(si2_system, rho) == pytest.approx(8.0, abs=1e-9)
    assert rho.min() >= 0.0


def test_ewald_of_diamond_silicon(si2_system):
    """Madelung-type constant of the diamond lattice with Z=4 at a=5.4307 A: the Ewald sum
    is independent of the splitting parameter eta (ewald.py:112-114)."""
    from oracle.ewald import ewald_energy

    z = np.array([4.0, 4.0])
    e1 = ewald_energy(si2_system.cell, si2_system.positions, z)
    e2 = ewald_energy(si2_system.cell, si2_system.positions, z, eta=0.9)
    assert e1 == pytest.approx(e2, abs=1e-10)
    assert e1 == pytest.approx(si2_system.ewald)
