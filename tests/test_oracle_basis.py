"""Oracle spheres, grid rule and transform pair, pinned by the reference's closed forms
(re-expressed from tests/unit/test_gvectors.py and tests/unit/test_fft.py)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle.basis import FFTGrid, GSphere, estimate_count, fast_grid_shape, miller_bounds, reciprocal_cell
from oracle.units import ANG_IN_BOHR


def test_reciprocal_cell_is_biorthogonal(si_cell_bohr):
    # test_gvectors.py:17-20
    assert np.allclose(reciprocal_cell(si_cell_bohr) @ si_cell_bohr.T, 2 * np.pi * np.eye(3), atol=1e-12)


def test_singular_cell_is_rejected():
    with pytest.raises(ValueError, match="singular"):
        reciprocal_cell(np.array([[1.0, 0, 0], [2.0, 0, 0], [0, 0, 1.0]]))


@pytest.mark.parametrize("ecut_ry", [60, 120, 200])
def test_count_matches_the_asymptote(si_cell_bohr, ecut_ry):
    # test_gvectors.py:28-41
    ecut = ecut_ry / 2.0
    vol = abs(np.linalg.det(si_cell_bohr))
    assert len(GSphere.build(si_cell_bohr, ecut)) == pytest.approx(estimate_count(vol, ecut), rel=0.01)


def test_sorted_and_first_shell(si_cell_bohr):
    # test_gvectors.py:65-78: sorted by g2, G=0 first, first shell (2 pi sqrt3 / a)^2 / 2
    s = GSphere.build(si_cell_bohr, 30.0)
    assert np.all(np.diff(s.g2) >= -1e-12)
    assert np.array_equal(s.miller[0], [0, 0, 0])
    a = 5.4307 * ANG_IN_BOHR
    assert s.kinetic_energy[1] == pytest.approx((2 * np.pi * np.sqrt(3) / a) ** 2 / 2, rel=1e-9)
    # ties are broken by (h1, h2, h3) ascending
    g2, m = s.g2, s.miller
    tie = np.flatnonzero(g2[1:] == g2[:-1])
    for t in tie:
        assert tuple(m[t]) < tuple(m[t + 1])


def test_every_vector_inside_and_bounds_enclose(si_cell_bohr):
    s = GSphere.build(si_cell_bohr, 20.0)
    assert np.all(s.kinetic_energy <= 20.0 + 1e-12)
    assert np.all(np.abs(s.miller).max(axis=0) <= miller_bounds(si_cell_bohr, np.sqrt(2 * 30.0)))


def test_density_sphere_reaches_twice_g(si_cell_bohr):
    # test_gvectors.py:111-117
    w = GSphere.build(si_cell_bohr, 20.0)
    r = GSphere.build(si_cell_bohr, 80.0)
    assert np.all(np.abs(r.miller).max(axis=0) == 2 * np.abs(w.miller).max(axis=0))


def test_shifted_sphere_has_no_zero_kinetic(si_cell_bohr):
    s = GSphere.build(si_cell_bohr, 30.0, kpoint=(0.25, 0.25, 0.25))
    assert s.g2.min() > 1e-6


def test_fast_grid_shape():
    # fft.py:64-67 doctest and test_fft.py:26-33
    assert fast_grid_shape((17, 23, 31)) == (18, 24, 32)
    assert fast_grid_shape((17, 23, 31), multiple=4) == (20, 24, 32)
    for n in fast_grid_shape((13, 29, 61)):
        for p in (2, 3, 5, 7, 11):
            while n % p == 0:
                n //= p
        assert n == 1


def test_reference_grids_of_the_workloads():
    """The grid rule (fft.py:72-81 + system.py:265-270) on the workload cells: 28 (Si2
    primitive, 15 Ha, multiple 4), 42 (Si8 cubic, 20 Ha), 84 (Si64, 20 Ha), 108 (Si216, 15 Ha)."""
    from tests.conftest import si_primitive_bohr

    a = 5.431 * ANG_IN_BOHR
    cell2, _ = si_primitive_bohr(5.431)
    assert FFTGrid.for_cutoff(cell2, 60.0, multiple=4).shape == (28, 28, 28)
    assert FFTGrid.for_cutoff(a * np.eye(3), 80.0).shape == (42, 42, 42)
    assert FFTGrid.for_cutoff(2 * a * np.eye(3), 80.0).shape == (84, 84, 84)
    assert len(GSphere.build(2 * a * np.eye(3), 20.0)) == 37073


@pytest.fixture
def spheres_and_grid(si_cell_bohr):
    wfc = GSphere.build(si_cell_bohr, 20.0)
    rho = GSphere.build(si_cell_bohr, 80.0)
    return wfc, rho, FFTGrid.for_sphere(rho)


def test_scatter_gather_round_trip_is_exact(spheres_and_grid):
    wfc, _, grid = spheres_and_grid
    c = np.random.default_rng(0).normal(size=(3, len(wfc))) + 0j
    assert np.array_equal(grid.gather(grid.scatter(c, wfc), wfc), c)


def test_fft_round_trip_and_parseval(spheres_and_grid):
    # test_fft.py:49-55, 76-84
    wfc, _, grid = spheres_and_grid
    rng = np.random.default_rng(1)
    c = rng.normal(size=len(wfc)) + 1j * rng.normal(size=len(wfc))
    u = grid.wave_to_real(c, wfc)
    assert np.abs(grid.real_to_wave(u, wfc) - c).max() < 1e-13
    assert float(np.real(grid.integrate(np.abs(u) ** 2))) == pytest.approx(grid.volume * np.sum(np.abs(c) ** 2), rel=1e-12)


def test_cosine_normalisation(spheres_and_grid):
    # test_fft.py:58-73: cos(G.r) -> 1/2 at +G and -G, nothing else
    wfc, _, grid = spheres_and_grid
    g = wfc.cartesian[5]
    c = grid.to_reciprocal(np.cos(grid.points_cartesian() @ g))
    shape = np.asarray(grid.shape)
    plus, minus = tuple(wfc.miller[5] % shape), tuple((-wfc.miller[5]) % shape)
    assert c[plus].real == pytest.approx(0.5, abs=1e-12)
    assert c[minus].real == pytest.approx(0.5, abs=1e-12)
    assert np.abs(c).sum() - abs(c[plus]) - abs(c[minus]) < 1e-11


def test_transform_is_the_plane_wave_sum(si_cell_bohr):
    """to_real is u(r_j) = sum_G c_G exp(i G.r_j) with no 1/N: brute force on a tiny grid."""
    wfc = GSphere.build(si_cell_bohr, 3.0)
    grid = FFTGrid.for_sphere(GSphere.build(si_cell_bohr, 12.0))
    rng = np.random.default_rng(3)
    c = rng.normal(size=len(wfc)) + 1j * rng.normal(size=len(wfc))
    r = grid.points_cartesian().reshape(-1, 3)
    brute = np.exp(1j * r @ wfc.cartesian.T) @ c
    assert np.abs(grid.wave_to_real(c, wfc).reshape(-1) - brute).max() < 1e-11


def test_alias_raises(si_cell_bohr, spheres_and_grid):
    _, rho, _ = spheres_and_grid
    with pytest.raises(ValueError, match="alias"):
        FFTGrid(shape=(8, 8, 8), cell=si_cell_bohr).scatter(np.ones(len(rho)), rho)


def test_gradient_and_divergence_of_a_plane_wave():
    """fft.py:288-304 on a triclinic cell: for u = cos(G.r) with G a grid vector below
    Nyquist, grad u = -G sin(G.r) and div grad u = -|G|^2 cos(G.r); a field with weight on
    the Nyquist plane of an even axis loses it in the gradient (zero_nyquist), while
    gvectors_cartesian without blanking keeps -n/2 there."""
    from oracle.basis import FFTGrid, reciprocal_cell

    cell = np.array([[5.1, 0.2, -0.3], [0.4, 4.7, 0.1], [-0.2, 0.3, 6.2]])
    grid = FFTGrid((12, 10, 15), cell)
    h = np.array([2, -3, 4])
    gvec = h @ reciprocal_cell(cell)
    r = grid.points_cartesian()
    phase = r @ gvec
    u = np.cos(phase)
    grad = grid.gradient(u)
    assert np.abs(grad + gvec * np.sin(phase)[..., None]).max() < 1e-11
    lap = grid.divergence(grad)
    assert np.abs(lap + (gvec @ gvec) * u).max() < 1e-10
    nyq = np.cos(np.pi * np.arange(12))[:, None, None] * np.ones((12, 10, 15))
    assert np.abs(grid.gradient(nyq)).max() < 1e-12
    g_full = grid.gvectors_cartesian()
    assert np.allclose(g_full[6, 0, 0], -6 * reciprocal_cell(cell)[0])
    assert np.allclose(grid.gvectors_cartesian(zero_nyquist=True)[6, 0, 0], 0.0)
