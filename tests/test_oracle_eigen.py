"""Oracle Davidson, subspace utilities and preconditioner, pinned by the reference's checks
(tests/unit/test_eigen.py) against dense LAPACK eigh on synthetic Hermitian operators."""

from __future__ import annotations

import numpy as np
import pytest
from scipy.linalg import eigh

from oracle.eigen import (
    davidson,
    orthogonalize_against,
    orthonormalize,
    random_bands,
    rayleigh_ritz,
    solve,
    teter_payne_allan,
)


def hermitian_operator(size=240, spread=40.0, coupling=0.02, seed=0):
    """Kinetic-like diagonal plus a random Hermitian perturbation (test_eigen.py:20-25)."""
    rng = np.random.default_rng(seed)
    diag = np.sort(rng.uniform(0.0, spread, size))
    off = coupling * (rng.normal(size=(size, size)) + 1j * rng.normal(size=(size, size)))
    return np.diag(diag) + off + off.conj().T, diag


def test_davidson_matches_dense_eigh():
    m, d = hermitian_operator()
    ref = eigh(m, eigvals_only=True)[:8]
    r = davidson(lambda b: b @ m.T, random_bands(len(d), 8, d, seed=1), d, tolerance=1e-10)
    assert r.converged
    assert r.eigenvalues == pytest.approx(ref, abs=1e-11)
    assert r.coefficients @ r.coefficients.conj().T == pytest.approx(np.eye(8), abs=1e-12)


def test_residual_bounds_the_error():
    m, d = hermitian_operator(seed=4)
    ref = eigh(m, eigvals_only=True)[:5]
    r = davidson(lambda b: b @ m.T, random_bands(len(d), 5, d, seed=2), d, tolerance=1e-5, max_iterations=6)
    assert np.all(np.abs(r.eigenvalues - ref) <= r.residual_norms + 1e-14)


def test_warm_start_is_faster():
    m, d = hermitian_operator(seed=7)
    cold = davidson(lambda b: b @ m.T, random_bands(len(d), 6, d, seed=3), d, tolerance=1e-10)
    warm = davidson(lambda b: b @ m.T, cold.coefficients, d, tolerance=1e-10)
    assert warm.iterations < cold.iterations
    assert warm.eigenvalues == pytest.approx(cold.eigenvalues, abs=1e-12)


def test_preconditioner_limits_and_closed_form():
    t = np.array([0.0, 1.0, 1000.0])
    f = teter_payne_allan(np.ones((1, 3)), t, np.array([1.0]))[0]
    assert f[0] == pytest.approx(1.0)
    assert f[2] == pytest.approx(1 / 2000.0, rel=2e-3)
    assert np.all(np.diff(f) < 0)
    x = 1.0
    assert f[1] == pytest.approx(65.0 / 81.0, rel=1e-15)  # (27+18+12+8)/(65+16)
    assert np.all(np.isfinite(teter_payne_allan(np.ones((1, 2)), np.array([1.0, 2.0]), np.array([0.0]))))
    del x


def test_orthonormalize_and_project():
    rng = np.random.default_rng(0)
    b = orthonormalize(rng.normal(size=(5, 30)) + 1j * rng.normal(size=(5, 30)))
    assert b @ b.conj().T == pytest.approx(np.eye(5), abs=1e-12)
    new = orthogonalize_against(rng.normal(size=(3, 30)) + 1j * rng.normal(size=(3, 30)), b)
    assert new @ b.conj().T == pytest.approx(np.zeros((len(new), 5)), abs=1e-12)
    assert new @ new.conj().T == pytest.approx(np.eye(len(new)), abs=1e-12)


def test_dependent_directions_are_dropped():
    rng = np.random.default_rng(1)
    basis = orthonormalize(rng.normal(size=(4, 20)))
    assert len(orthogonalize_against(basis[:2] + 1e-14 * rng.normal(size=(2, 20)), basis)) == 0


def test_rayleigh_ritz_exact_in_complete_space():
    m, _ = hermitian_operator(size=30, seed=9)
    basis = orthonormalize(np.eye(30, dtype=complex))
    w, v, hv = rayleigh_ritz(basis, basis @ m.T, 30)
    assert w == pytest.approx(eigh(m, eigvals_only=True), abs=1e-12)
    assert hv == pytest.approx(v @ m.T, abs=1e-11)


def test_random_bands_orthonormal_damped_reproducible():
    t = np.linspace(0.0, 50.0, 200)
    b = random_bands(200, 6, t, seed=0)
    assert b @ b.conj().T == pytest.approx(np.eye(6), abs=1e-12)
    w = np.mean(np.abs(b) ** 2, axis=0)
    assert w[:20].sum() > 5 * w[-20:].sum()
    assert np.array_equal(random_bands(50, 3, t[:50], seed=5), random_bands(50, 3, t[:50], seed=5))


def test_solve_rejects_unknown():
    t = np.linspace(0.1, 5.0, 20)
    with pytest.raises(NotImplementedError, match="lobpcg"):
        solve("lobpcg", lambda b: b, random_bands(20, 2, t), t)
