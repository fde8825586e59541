"""Oracle datasets, radial transforms, Y_lm and KB projectors, pinned by the reference's
closed forms (tests/unit/test_pseudo_readers.py, test_formfactors.py, test_kb.py)."""

from __future__ import annotations

import numpy as np
import pytest
from numpy.polynomial import legendre

from oracle.basis import GSphere
from oracle.pseudo import (
    MAX_L,
    RadialGrid,
    build_projectors,
    density_form_factor,
    local_potential_form_factor,
    projector_form_factor,
    radial_transform,
    read_psp8,
    real_spherical_harmonics,
)
from tests.conftest import DATA, si_primitive_bohr


def test_psp8_header_and_invariants(si_pseudo):
    # test_pseudo_readers.py:37-112
    p = si_pseudo
    assert (p.symbol, p.atomic_number, p.valence_charge, p.xc_functional) == ("Si", 14.0, 4.0, "PBE")
    assert [q.l for q in p.projectors] == [0, 0, 1, 1, 2, 2]
    assert p.local_potential_tail_error(3.0) < 1e-3
    far = p.grid.r >= 5.0
    assert p.local_potential[far] == pytest.approx(-4.0 / p.grid.r[far], rel=1e-5)
    for q in p.projectors:
        assert q.norm(p.grid) == pytest.approx(1.0, abs=1e-4)
    charge = p.integrated_valence_charge()
    assert charge == pytest.approx(4.0, rel=0.02) and charge < 4.0
    assert 0 < p.grid.integrate(p.core_density) < 10


def test_pbesol_dataset_reads():
    p = read_psp8(DATA / "Si.pbesol.psp8.gz")
    assert p.xc_functional == "PBEsol" and len(p.projectors) == 6


@pytest.mark.parametrize("l", range(MAX_L + 1))
def test_addition_theorem(l):
    # test_kb.py:20-32
    rng = np.random.default_rng(l)
    a = rng.normal(size=(40, 3))
    b = rng.normal(size=(40, 3))
    a /= np.linalg.norm(a, axis=1)[:, None]
    b /= np.linalg.norm(b, axis=1)[:, None]
    left = np.einsum("mi,mi->i", real_spherical_harmonics(l, a), real_spherical_harmonics(l, b))
    coef = np.zeros(l + 1)
    coef[l] = 1.0
    right = (2 * l + 1) / (4 * np.pi) * legendre.legval(np.einsum("ij,ij->i", a, b), coef)
    assert left == pytest.approx(right, abs=1e-14)


def test_l1_ordering_is_y_z_x():
    # kb.py:73-75: the m ordering is a convention that fixes the projector layout
    y = real_spherical_harmonics(1, np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0]]))
    assert np.allclose(y, np.sqrt(0.75 / np.pi) * np.eye(3))


def test_radial_transform_of_a_gaussian():
    # test_formfactors.py:95-106
    a = 1.7
    r = np.linspace(0.0, 12.0, 4001)
    q = np.array([0.0, 0.5, 1.0, 2.5, 5.0])
    got = radial_transform(RadialGrid(r), 4 * np.pi * r**2 * np.exp(-a * r**2), 0, q)
    assert got == pytest.approx((np.pi / a) ** 1.5 * np.exp(-(q**2) / (4 * a)), rel=1e-10)


def test_density_form_factor_at_zero_is_the_charge(si_pseudo):
    v = density_form_factor(si_pseudo.grid, si_pseudo.valence_density, np.array([0.0]))[0]
    assert v == pytest.approx(si_pseudo.integrated_valence_charge(), rel=1e-12)


def test_local_potential_branches_join(si_pseudo):
    # test_formfactors.py:36-49
    z = si_pseudo.valence_charge
    at0 = local_potential_form_factor(si_pseudo, np.array([0.0]))[0]
    q = np.array([1e-2, 3e-3, 1e-3])
    lim = local_potential_form_factor(si_pseudo, q) + 4 * np.pi * z / q**2
    assert lim[-1] == pytest.approx(at0, rel=1e-6)
    small = np.array([0.05, 0.1])
    assert local_potential_form_factor(si_pseudo, small) == pytest.approx(-4 * np.pi * z / small**2, rel=0.01)


@pytest.mark.parametrize("index", range(6))
def test_projector_sum_rule(si_pseudo, index):
    # test_formfactors.py:71-81
    q = np.linspace(0.0, 80.0, 6001)
    f = projector_form_factor(si_pseudo.grid, si_pseudo.projectors[index], q)
    assert 2 / np.pi * np.trapezoid(q**2 * f**2, q) == pytest.approx(1.0, abs=2e-5)


def test_projector_count_and_diagonal_d(si_pseudo):
    # test_kb.py:111-126
    cell, pos = si_primitive_bohr()
    s = GSphere.build(cell, 6.0)
    kb = build_projectors(s, pos, ["Si", "Si"], {"Si": si_pseudo}, abs(np.linalg.det(cell)))
    assert len(kb) == 36 and kb.dij.shape == (36, 36)
    assert np.count_nonzero(kb.dij - np.diag(np.diag(kb.dij))) == 0
    assert [lab[1] for lab in kb.labels[:18]] == [0, 0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2]


def test_projector_completeness_matches_continuum(si_pseudo):
    # test_kb.py:55-90
    cell, pos = si_primitive_bohr()
    vol = abs(np.linalg.det(cell))
    prev = 0.0
    for ecut in (12.0, 24.0):
        s = GSphere.build(cell, ecut, kpoint=(0.2, 0.1, 0.0))
        kb = build_projectors(s, pos, ["Si", "Si"], {"Si": si_pseudo}, vol)
        d = kb.completeness()
        assert np.all(d < 1 + 1e-9) and d.min() > prev
        prev = d.min()
        q = np.linspace(0.0, np.sqrt(2 * ecut), 4000)
        for row, (_, l, ch, _) in enumerate(kb.labels):
            f = projector_form_factor(si_pseudo.grid, si_pseudo.projectors[ch], q)
            assert d[row] == pytest.approx(2 / np.pi * np.trapezoid(q**2 * f**2, q), rel=0.05)


def test_projector_operator_is_hermitian(si_pseudo):
    # test_kb.py:93-108
    cell, pos = si_primitive_bohr()
    s = GSphere.build(cell, 8.0, kpoint=(0.1, -0.2, 0.3))
    kb = build_projectors(s, pos, ["Si", "Si"], {"Si": si_pseudo}, abs(np.linalg.det(cell)))
    rng = np.random.default_rng(0)
    b = rng.normal(size=(3, len(s))) + 1j * rng.normal(size=(3, len(s)))
    m = b.conj() @ kb.apply(b).T
    assert m == pytest.approx(m.conj().T, abs=1e-12)


def test_projector_occupancies_against_the_energy(si_pseudo):
    """occupancies (kb.py:243-256) through the energy it must reproduce: for a real
    symmetric D, sum_n w_n <psi_n|V_NL|psi_n> = trace(D occ); occ is symmetric and, for
    non-negative weights, positive semi-definite; a single band with weight 1 gives the
    outer product of its projections."""
    cell, pos = si_primitive_bohr()
    s = GSphere.build(cell, 8.0, kpoint=(0.1, -0.2, 0.3))
    kb = build_projectors(s, pos, ["Si", "Si"], {"Si": si_pseudo}, abs(np.linalg.det(cell)))
    rng = np.random.default_rng(4)
    b = rng.normal(size=(5, len(s))) + 1j * rng.normal(size=(5, len(s)))
    w = np.array([1.0, 0.5, 0.25, 0.0, 2.0])
    occ = kb.occupancies(b, w)
    assert occ.shape == (len(kb), len(kb))
    assert np.abs(occ - occ.T).max() < 1e-12 * np.abs(occ).max()
    assert np.linalg.eigvalsh(occ).min() > -1e-10 * np.abs(occ).max()
    assert np.trace(kb.dij @ occ) == pytest.approx(kb.energy(b, w), rel=1e-12)
    one = kb.occupancies(b[:1], np.ones(1))
    p = kb.project(b[0])
    assert np.abs(one - np.outer(p, p.conj()).real).max() < 1e-12 * np.abs(one).max()


def test_projector_element_closed_form(si_pseudo):
    """One matrix element recomputed from its definition
    (4 pi/sqrt Omega)(-i)^l Y_lm f_l e^{-i(k+G).tau} with an independent quadrature of f_l."""
    cell, pos = si_primitive_bohr()
    vol = abs(np.linalg.det(cell))
    s = GSphere.build(cell, 6.0, kpoint=(0.1, 0.2, -0.3))
    kb = build_projectors(s, pos, ["Si", "Si"], {"Si": si_pseudo}, vol)
    g = 7
    kg = s.cartesian[g]
    qn = np.linalg.norm(kg)
    row = 18 + 2 + 1  # atom 1: two s rows, then l=1 channel m=-1 (y), m=0 (z)
    _, l, ch, m = kb.labels[row]
    assert (l, m) == (1, 0)
    r = si_pseudo.grid.r
    x = qn * r
    xs = np.where(x > 1e-8, x, 1.0)  # the series branch below covers x -> 0
    j1 = np.where(x > 1e-8, np.sin(xs) / xs**2 - np.cos(xs) / xs, x / 3)
    f = np.trapezoid(r * si_pseudo.projectors[ch].values * j1, r)
    expected = 4 * np.pi / np.sqrt(vol) * (-1j) * np.sqrt(0.75 / np.pi) * kg[2] / qn * f * np.exp(-1j * kg @ pos[1])
    assert kb.matrix[row, g] == pytest.approx(expected, rel=2e-4)
