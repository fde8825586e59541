"""Host-side ionic setup of the product against the oracle: the Ewald energy (numpy on both
sides, so it runs on the CPU) and the radial form factors the device superposition uses."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import si_primitive_bohr


def test_ewald_matches_oracle():
    from amaterasu_amd.system import ewald_energy
    from oracle.ewald import ewald_energy as oracle_ewald

    cell, pos = si_primitive_bohr()
    assert ewald_energy(cell, pos, [4.0, 4.0]) == pytest.approx(oracle_ewald(cell, pos, np.array([4.0, 4.0])),
                                                               abs=1e-10)


def test_ewald_independent_of_splitting():
    """The Ewald sum does not depend on the Gaussian width eta; a wrong self or background
    term shows up as an eta dependence."""
    from amaterasu_amd.system import ewald_energy

    cell, pos = si_primitive_bohr()
    shifted = pos + np.array([0.1, -0.2, 0.05])
    e = [ewald_energy(cell, shifted, [4.0, 4.0], eta=eta) for eta in (0.4, 0.7, 1.1)]
    assert max(e) - min(e) < 1e-9


def test_form_factors_match_oracle(si_pseudo):
    from amaterasu_amd.system import density_form_factor, load_dataset, local_form_factor
    from oracle.pseudo import density_form_factor as odens
    from oracle.pseudo import local_potential_form_factor as olocal

    ds = load_dataset()
    q = np.array([0.0, 1e-3, 0.5, 1.3, 4.0, 9.0])
    np.testing.assert_allclose(local_form_factor(ds, q), olocal(si_pseudo, q), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(density_form_factor(ds, ds.valence_density, q), odens(si_pseudo.grid, si_pseudo.valence_density, q),
                               rtol=1e-12, atol=1e-12)
    if ds.core_density is not None:
        np.testing.assert_allclose(density_form_factor(ds, ds.core_density, q), odens(si_pseudo.grid, si_pseudo.core_density, q),
                                   rtol=1e-12, atol=1e-12)
