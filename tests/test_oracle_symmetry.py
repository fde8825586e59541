"""Oracle space-group operations and irreducible k-points (the spglib-free restatement),
pinned by the textbook fcc counts of tests/unit/test_symmetry.py:58-74 and the closed
form of the shifted 2x2x2 diamond mesh."""

from __future__ import annotations

import numpy as np
import pytest

from oracle.basis import symmetry_compatible_multiple
from oracle.symmetry import find_symmetry, ibz_kpoints
from tests.conftest import si_primitive_bohr


@pytest.fixture(scope="module")
def diamond():
    cell, pos = si_primitive_bohr()
    return find_symmetry(cell, pos @ np.linalg.inv(cell), ["Si", "Si"])


def test_diamond_has_48_operations_and_grid_multiple_4(diamond):
    assert diamond.n_operations == 48
    assert np.array_equal(diamond.rotations[0], np.eye(3, dtype=int))
    assert symmetry_compatible_multiple(diamond.translations) == 4


@pytest.mark.parametrize(("mesh", "shift", "expected"), [((4, 4, 4), (0, 0, 0), 8), ((4, 4, 4), (0.5, 0.5, 0.5), 10),
                                                         ((6, 6, 6), (0, 0, 0), 16), ((8, 8, 8), (0, 0, 0), 29)])
def test_fcc_irreducible_counts(diamond, mesh, shift, expected):
    k, w = ibz_kpoints(diamond, mesh, shift)
    assert len(k) == expected and w.sum() == pytest.approx(1.0, abs=1e-12)


def test_shifted_2x2x2_closed_form(diamond):
    k, w = ibz_kpoints(diamond, (2, 2, 2), (0.5, 0.5, 0.5))
    assert np.allclose(k, [[0.25, 0.25, 0.25], [-0.25, 0.25, 0.25]])
    assert np.allclose(w, [0.25, 0.75])


def test_displaced_cell_is_p1():
    from amaterasu_amd.workloads import si_conventional

    c, p = si_conventional(1, sigma_ang=0.05, seed=3)
    assert find_symmetry(c, p @ np.linalg.inv(c), ["Si"] * 8).n_operations == 1


def test_symmetrised_density_keeps_the_charge(diamond, si2_system):
    """G = 0 is fixed by every operation (density.py:13-18): symmetrisation conserves charge."""
    import copy

    s = copy.copy(si2_system)
    s.symmetry = diamond
    rng = np.random.default_rng(0)
    rho = si2_system.initial_density() * (1.0 + 0.1 * rng.normal(size=si2_system.grid.shape))
    sym = s.symmetrize_density(rho)
    assert sym.sum() == pytest.approx(rho.sum(), rel=1e-12)
    assert np.abs(s.symmetrize_density(sym) - sym).max() < 1e-12  # idempotent
