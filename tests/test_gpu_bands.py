"""The second caller of ``solve``: a non-self-consistent band structure
(``solvers/bands.py:93-150``) -- the converged density's effective potential on a system
rebuilt at k-points along a path, one ``solve("davidson", ...)`` per k from the reference's
seeded start (``7919 + 31 spin + ik``) with its tolerance (``tol_energy / nbands``) and
iteration limit (200). The device eigenvalues must equal the oracle's on the same operator."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PATH = [[0.0, 0.0, 0.0], [0.5, 0.0, 0.5], [0.5, 0.5, 0.5], [0.375, 0.375, 0.75], [0.25, 0.0, 0.25]]  # G X L K mid


def test_band_structure_along_a_path_matches_oracle(si2_system, si_pseudo):
    import copy

    from amaterasu_amd.eigen import solve
    from amaterasu_amd.hamiltonian import Hamiltonian
    from amaterasu_amd.scf import SCFSettings, run_scf
    from oracle.eigen import random_bands
    from oracle.eigen import solve as osolve
    from oracle.hamiltonian import Hamiltonian as OH
    from oracle.scf import effective_potential
    from oracle.symmetry import find_symmetry
    from oracle.system import build_system
    from oracle.xc import XCFunctional

    s = copy.copy(si2_system)
    s.xc = XCFunctional("PBE")
    # the density of the two irreducible k-points is only cubic once symmetrised (as the
    # reference does), which the degeneracy check at Gamma below relies on
    s.symmetry = find_symmetry(s.cell, s.positions @ np.linalg.inv(s.cell), s.symbols)
    tol_energy = 1e-9
    scf = run_scf(s, SCFSettings(tol_energy=tol_energy, tol_density=1e-8))
    assert scf.converged
    bands = build_system(s.cell, s.positions, s.symbols, {"Si": si_pseudo}, 8.0, 32.0, PATH,
                         np.full(len(PATH), 1.0 / len(PATH)), nbands=8)
    bands.xc = s.xc
    assert bands.grid.shape == s.grid.shape  # the potential is reused on the same grid
    v, _, _ = effective_potential(bands, scf.density)
    h, o = Hamiltonian(bands, v), OH(bands, v)
    for ik, sphere in enumerate(bands.wave_spheres):
        guess = random_bands(len(sphere), bands.nbands, sphere.kinetic_energy, seed=7919 + ik)
        kw = dict(tolerance=tol_energy / bands.nbands, max_iterations=200)
        got = solve("davidson", lambda b, ik=ik: h.apply(b, ik), guess, sphere.kinetic_energy, **kw)
        ref = osolve("davidson", lambda b, ik=ik: o.apply(b, ik), guess, sphere.kinetic_energy, **kw)
        assert got.converged and ref.converged
        assert np.abs(got.eigenvalues - ref.eigenvalues).max() < 1e-9
    gamma = solve("davidson", lambda b: h.apply(b, 0), random_bands(len(bands.wave_spheres[0]), 8,
                  bands.wave_spheres[0].kinetic_energy, seed=7919), bands.wave_spheres[0].kinetic_energy,
                  tolerance=1e-10, max_iterations=200)
    e = gamma.eigenvalues
    assert abs(e[1] - e[3]) < 1e-8 and e[3] < e[4]  # the triply degenerate valence top at Gamma
