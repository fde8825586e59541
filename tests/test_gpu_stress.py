"""Stress on the GPU against the oracle's restatement of ``properties/stress.py`` (whose
terms the frozen-basis finite-difference tests pin), term by term on the same inputs;
on converged SCF states: the same comparison, the symmetrised stress of cubic Si being
isotropic, the sign of the pressure under compression and expansion (the reference's
``tests/unit/test_stress.py`` checks), and the stress error of the mixed-precision SCF."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import si_primitive_bohr

pytestmark = pytest.mark.gpu

KPTS, WEIGHTS = [[0.25, 0.25, 0.25], [-0.25, 0.25, 0.25]], [0.25, 0.75]


def _pair(si_pseudo, cell, pos, symmetric=False):
    from amaterasu_amd.system import attach_ionic_fields, build_system, load_dataset
    from oracle.symmetry import find_symmetry
    from oracle.system import build_system as obuild
    from oracle.xc import XCFunctional

    o = obuild(cell, pos, ["Si", "Si"], {"Si": si_pseudo}, 8.0, 32.0, KPTS, WEIGHTS)
    o.xc = XCFunctional("PBE")
    ds = load_dataset()
    s = build_system(cell, pos, ["Si", "Si"], {"Si": ds}, 32.0, KPTS, 8.0, weights=WEIGHTS, grid_shape=o.grid.shape)
    attach_ionic_fields(s, {"Si": ds}, 32.0, xc="PBE")
    if symmetric:
        o.symmetry = s.symmetry = find_symmetry(cell, pos @ np.linalg.inv(cell), ["Si", "Si"])
    return o, s


def _distorted():
    cell, pos = si_primitive_bohr()
    cell = cell @ np.array([[1.01, 0.02, -0.01], [0.02, 0.98, 0.015], [-0.01, 0.015, 1.02]])
    pos = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]]) @ cell + np.array([[0.0, 0.0, 0.0], [0.21, -0.13, 0.08]])
    return cell, pos


def test_stress_terms_match_oracle(si_pseudo):
    from amaterasu_amd.stress import compute_stress
    from oracle.eigen import random_bands
    from oracle.stress import compute_stress as ostress

    o, s = _pair(si_pseudo, *_distorted())
    rho = o.initial_density()
    wfs = [[random_bands(len(sp.g2), 6, sp.g2 / 2, seed=7 + ik) for ik, sp in enumerate(o.wave_spheres)]]
    occ = np.zeros((1, 2, 6))
    occ[0, :, :4] = 1.0
    occ[0, 1, 4] = 0.3
    got = compute_stress(s, rho, wfs, occ, decompose=True)
    want = ostress(o, rho, wfs, occ, decompose=True)
    assert set(got) == set(want)
    for key in want:
        assert np.abs(want[key]).max() > 1e-6, key
        assert np.abs(got[key] - want[key]).max() < 1e-12, key


def test_scf_stress_fp64_and_mixed(si_pseudo):
    """Stress from the device SCF against the oracle on the same converged state, and the
    mixed-precision stress against the FP64 one."""
    from amaterasu_amd.scf import SCFSettings, run_scf
    from amaterasu_amd.stress import compute_stress
    from oracle.stress import compute_stress as ostress

    o, s = _pair(si_pseudo, *_distorted())
    r64 = run_scf(s, SCFSettings(tol_energy=1e-9, tol_density=1e-8))
    assert r64.converged
    s64 = compute_stress(s, r64.density, r64.wavefunctions, r64.occupations)
    so = ostress(o, r64.density, r64.wavefunctions, r64.occupations)
    assert np.abs(s64 - so).max() < 1e-12
    assert np.abs(s64 - s64.T).max() < 1e-12
    rmx = run_scf(s, SCFSettings(tol_energy=1e-9, tol_density=1e-8, precision="mixed"))
    assert rmx.converged
    smx = compute_stress(s, rmx.density, rmx.wavefunctions, rmx.occupations)
    print(f"mixed - fp64 stress: {np.abs(smx - s64).max():.3e} Ha/Bohr^3 (|sigma| {np.abs(s64).max():.3e})")
    assert np.abs(smx - s64).max() < 1e-8


def test_cubic_stress_is_isotropic_and_pressure_has_the_right_sign(si_pseudo):
    from amaterasu_amd.scf import SCFSettings, run_scf
    from amaterasu_amd.stress import compute_stress

    pressures = []
    for scale in (0.94, 1.06):
        cell, pos = si_primitive_bohr()
        _, s = _pair(si_pseudo, cell * scale, pos * scale, symmetric=True)
        r = run_scf(s, SCFSettings(tol_energy=1e-9, tol_density=1e-8))
        assert r.converged
        sigma = compute_stress(s, r.density, r.wavefunctions, r.occupations)
        d = np.diag(sigma)
        assert np.abs(d - d.mean()).max() < 1e-14
        assert np.abs(sigma - np.diag(d)).max() < 1e-14
        pressures.append(-np.trace(sigma) / 3)
    compressed, expanded = pressures
    assert compressed > expanded
    assert compressed > 0.0 > expanded
