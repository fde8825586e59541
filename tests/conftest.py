"""Shared fixtures and the ``gpu`` marker.

``-m "not gpu"`` runs the oracle checks, the host logic and the C-ABI symbol checks
on the CPU; ``-m gpu`` runs the HIP parity tests on an MI355X.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
DATA = REPO / "amaterasu_amd" / "data"
GOLDEN = Path(__file__).resolve().parent / "golden"

#: Si lattice constant of the reference's own fixtures, in Angstrom (tests/conftest.py:13).
SI_LATTICE_CONSTANT = 5.4307


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X and the built HIP library")


@pytest.fixture(scope="session")
def si_pseudo():
    from oracle.pseudo import read_psp8

    return read_psp8(DATA / "Si.pbe.psp8.gz")


def si_primitive_bohr(a_ang=SI_LATTICE_CONSTANT):
    """fcc primitive cell of diamond Si (ase ``bulk('Si','diamond')`` layout) and its 2 atoms."""
    from oracle.units import ANG_IN_BOHR

    a = a_ang * ANG_IN_BOHR
    cell = 0.5 * a * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
    positions = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]]) @ cell
    return cell, positions


@pytest.fixture
def si_cell_bohr():
    return si_primitive_bohr()[0]


@pytest.fixture(scope="session")
def si2_system(si_pseudo):
    """Si2 primitive, ecut 8 Ha, the two IBZ k of the shifted 2x2x2 mesh, no symmetrisation."""
    from oracle.system import build_system

    cell, pos = si_primitive_bohr()
    return build_system(cell, pos, ["Si", "Si"], {"Si": si_pseudo}, 8.0, 32.0,
                        [[0.25, 0.25, 0.25], [-0.25, 0.25, 0.25]], [0.25, 0.75])
