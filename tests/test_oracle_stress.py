"""Pins of the oracle's stress terms against a frozen-basis finite difference.

Straining the cell with the plane-wave coefficients held on their Miller indices and
recomputing every energy term is the exact counterpart of the analytic strain derivative
(the reference's own strategy, ``tests/unit/test_stress.py``): each term of
``oracle.stress.compute_stress`` must equal (1/Omega) dE_term/d eps_uv by central
difference, on a triclinic two-atom cell with displaced atoms, a GGA and the partial core
(the core term joins the exchange-correlation one, which the difference cannot separate).
A dropped factor, a sign, a transposed index or a wrong form-factor derivative fails here.
"""

from __future__ import annotations

import numpy as np
import pytest

STEP = 1e-5


def _strained_sphere(sphere, cell):
    from oracle.basis import GSphere, reciprocal_cell

    rec = reciprocal_cell(cell)
    cart = sphere.miller @ rec + sphere.kpoint @ rec
    return GSphere(miller=sphere.miller, cartesian=cart, g2=np.einsum("gi,gi->g", cart, cart),
                   kpoint=sphere.kpoint, ecut=sphere.ecut, cell=cell)


@pytest.fixture(scope="module")
def frozen(si_pseudo):
    from oracle.density import compute_density
    from oracle.eigen import random_bands
    from oracle.system import build_system
    from oracle.xc import XCFunctional

    cell = np.array([[0.3, 5.0, 5.2], [5.1, -0.2, 4.9], [4.8, 5.3, 0.4]])
    pos = np.array([[0.05, -0.08, 0.02], [2.7, 2.45, 2.75]])
    s = build_system(cell, pos, ["Si", "Si"], {"Si": si_pseudo}, 6.0, 24.0,
                     [[0.25, 0.25, 0.25], [-0.25, 0.25, 0.25]], [0.25, 0.75])
    s.xc = XCFunctional("PBE")
    nb = 4
    wf = [[random_bands(len(sp), nb, sp.kinetic_energy, seed=31 + ik) for ik, sp in enumerate(s.wave_spheres)]]
    occ = np.ones((1, len(s.wave_spheres), nb))
    occ[0, 1, 3] = 0.4  # a fractional occupation exercises the band weights
    rho = compute_density(s, wf, occ)
    return s, wf, occ, rho


def _energy_terms(s, wf, occ, rho, strain, pseudos):
    from oracle.basis import FFTGrid
    from oracle.ewald import ewald_energy
    from oracle.hamiltonian import atomic_core_density, local_potential
    from oracle.pseudo import build_projectors
    from oracle.scf import hartree

    mat = np.eye(3) + strain
    cell = s.cell @ mat.T
    pos = s.positions @ mat.T
    grid = FFTGrid(shape=s.grid.shape, cell=cell)
    dsp = _strained_sphere(s.density_sphere, cell)
    scaled = rho * (s.grid.volume / grid.volume)  # fixed coefficients: the density dilutes
    total = scaled.sum(axis=0)
    kin = nl = 0.0
    for ik, sp in enumerate(s.wave_spheres):
        st = _strained_sphere(sp, cell)
        pr = build_projectors(st, pos, s.symbols, pseudos, grid.volume)
        w = 2.0 * s.kpoints.weights[ik] * occ[0, ik]
        c = wf[0][ik]
        kin += float(np.einsum("n,ng->", w, np.abs(c) ** 2 * (0.5 * st.g2)))
        nl += pr.energy(c, w)
    v = local_potential(grid, dsp, pos, s.symbols, pseudos)
    core = atomic_core_density(grid, dsp, pos, s.symbols, pseudos)
    charges = np.array([pseudos[x].valence_charge for x in s.symbols])
    return {"kinetic": kin, "hartree": hartree(grid, dsp, total)[0],
            "local": float(np.sum(v * total)) * grid.cell_volume_per_point, "nonlocal": nl,
            "xc": s.xc.evaluate_on_grid(scaled + core[None] / s.nspin, grid)[0],
            "ewald": ewald_energy(cell, pos, charges)}


@pytest.fixture(scope="module")
def comparison(frozen, si_pseudo):
    from oracle.stress import compute_stress

    s, wf, occ, rho = frozen
    analytic = compute_stress(s, rho, wf, occ, symmetrize=False, decompose=True)
    merged = dict(analytic)
    merged["xc"] = merged["xc"] + merged.pop("core")
    numerical = {k: np.zeros((3, 3)) for k in merged}
    for u in range(3):
        for v in range(u, 3):
            e = np.zeros((3, 3))
            e[u, v] += 0.5
            e[v, u] += 0.5
            plus = _energy_terms(s, wf, occ, rho, STEP * e, {"Si": si_pseudo})
            minus = _energy_terms(s, wf, occ, rho, -STEP * e, {"Si": si_pseudo})
            for k in merged:
                numerical[k][u, v] = numerical[k][v, u] = (plus[k] - minus[k]) / (2 * STEP) / s.grid.volume
    return merged, numerical


@pytest.mark.parametrize("term", ["ewald", "kinetic", "hartree", "local", "nonlocal", "xc"])
def test_each_term_is_the_strain_derivative_of_its_energy(comparison, term):
    analytic, numerical = comparison
    assert np.abs(analytic[term]).max() > 1e-6  # the term is really exercised
    assert analytic[term] == pytest.approx(numerical[term], abs=2e-9)


def test_total_and_symmetry(comparison):
    analytic, numerical = comparison
    total = sum(analytic.values())
    assert total == pytest.approx(sum(numerical.values()), abs=5e-9)
    assert total == pytest.approx(total.T, abs=1e-12)


def test_solid_harmonics_reduce_to_the_real_harmonics_and_differentiate():
    from oracle.pseudo import real_spherical_harmonics, solid_harmonics, solid_harmonics_gradient

    q = np.random.default_rng(5).normal(size=(40, 3))
    n = np.linalg.norm(q, axis=1)
    for l in range(4):
        assert np.abs(solid_harmonics(l, q) / n**l - real_spherical_harmonics(l, q)).max() < 1e-14
        h = 1e-6
        fd = np.stack([(solid_harmonics(l, q + h * np.eye(3)[a]) - solid_harmonics(l, q - h * np.eye(3)[a])) / (2 * h)
                       for a in range(3)], axis=-1)
        assert np.abs(solid_harmonics_gradient(l, q) - fd).max() < 1e-8


def test_form_factor_derivatives_match_finite_differences(si_pseudo):
    from oracle.pseudo import (density_form_factor, density_form_factor_derivative, local_potential_form_factor,
                               local_potential_form_factor_derivative, projector_form_factor,
                               projector_form_factor_derivative)

    q = np.linspace(0.3, 6.0, 12)
    h = 1e-5
    fd = (local_potential_form_factor(si_pseudo, q + h) - local_potential_form_factor(si_pseudo, q - h)) / (2 * h)
    assert local_potential_form_factor_derivative(si_pseudo, q) == pytest.approx(fd, rel=1e-6, abs=1e-8)
    assert local_potential_form_factor_derivative(si_pseudo, np.array([0.0]))[0] == 0.0
    g = si_pseudo.grid
    fd = (density_form_factor(g, si_pseudo.core_density, q + h) - density_form_factor(g, si_pseudo.core_density, q - h)) / (2 * h)
    assert density_form_factor_derivative(g, si_pseudo.core_density, q) == pytest.approx(fd, rel=1e-6, abs=1e-9)
    for p in si_pseudo.projectors:
        fd = (projector_form_factor(g, p, q + h) - projector_form_factor(g, p, q - h)) / (2 * h)
        assert projector_form_factor_derivative(g, p, q) == pytest.approx(fd, rel=1e-6, abs=1e-9)


def test_cubic_symmetrisation_makes_a_tensor_isotropic(si2_system):
    from oracle.stress import symmetrize_tensor
    from oracle.symmetry import find_symmetry

    s = si2_system
    sym = find_symmetry(s.cell, s.positions @ np.linalg.inv(s.cell), s.symbols)
    assert len(sym.rotations) == 48
    t = np.random.default_rng(2).normal(size=(3, 3))
    t = t + t.T
    out = symmetrize_tensor(t, sym, s.cell)
    assert np.abs(out - np.eye(3) * np.trace(t) / 3).max() < 1e-14
