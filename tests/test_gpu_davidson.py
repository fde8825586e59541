"""Davidson on the GPU (scf/eigen/davidson.py:48-111) against dense LAPACK eigh and the
oracle, plus its fused residual and preconditioner kernels."""

from __future__ import annotations

import numpy as np
import pytest
from scipy.linalg import eigh

pytestmark = pytest.mark.gpu


def hermitian_operator(size=240, spread=40.0, coupling=0.02, seed=0):
    rng = np.random.default_rng(seed)
    d = np.sort(rng.uniform(0.0, spread, size))
    off = coupling * (rng.normal(size=(size, size)) + 1j * rng.normal(size=(size, size)))
    return np.diag(d) + off + off.conj().T, d


def test_dense_operator_matches_eigh():
    from amaterasu_amd.eigen import davidson, random_bands

    m, d = hermitian_operator()
    ref = eigh(m, eigvals_only=True)[:8]
    r = davidson(lambda b: b @ m.T, random_bands(len(d), 8, d, seed=1), d, tolerance=1e-10)
    assert r.converged
    assert np.abs(r.eigenvalues - ref).max() < 1e-11
    assert np.abs(r.coefficients @ r.coefficients.conj().T - np.eye(8)).max() < 1e-12


def test_residuals_bound_the_error():
    from amaterasu_amd.eigen import davidson, random_bands

    m, d = hermitian_operator(seed=4)
    ref = eigh(m, eigvals_only=True)[:5]
    r = davidson(lambda b: b @ m.T, random_bands(len(d), 5, d, seed=2), d, tolerance=1e-5, max_iterations=6)
    assert np.all(np.abs(r.eigenvalues - ref) <= r.residual_norms + 1e-14)


def test_random_bands_are_the_reference_draws():
    """Same default_rng stream as subspace.py:79-88: equal to the oracle up to the
    orthonormalisation's rounding."""
    from amaterasu_amd.eigen import random_bands
    from oracle.eigen import random_bands as orb

    t = np.linspace(0.0, 50.0, 300)
    a, b = random_bands(300, 6, t, seed=12345), orb(300, 6, t, seed=12345)
    assert np.abs(a - b).max() < 1e-12


def test_subspace_utilities():
    from amaterasu_amd.eigen import orthogonalize_against, orthonormalize, rayleigh_ritz

    rng = np.random.default_rng(0)
    b = orthonormalize(rng.normal(size=(5, 30)) + 1j * rng.normal(size=(5, 30)))
    assert np.abs(b @ b.conj().T - np.eye(5)).max() < 1e-12
    new = orthogonalize_against(rng.normal(size=(3, 30)) + 1j * rng.normal(size=(3, 30)), b)
    assert np.abs(new @ b.conj().T).max() < 1e-12
    assert np.abs(new @ new.conj().T - np.eye(len(new))).max() < 1e-12
    basis = orthonormalize(rng.normal(size=(4, 20)) + 0j)
    assert len(orthogonalize_against(basis[:2] + 1e-14 * rng.normal(size=(2, 20)), basis)) == 0
    m, _ = hermitian_operator(size=30, seed=9)
    q = orthonormalize(np.eye(30, dtype=complex))
    w, v, hv = rayleigh_ritz(q, q @ m.T, 30)
    assert np.abs(w - eigh(m, eigvals_only=True)).max() < 1e-12
    assert np.abs(hv - v @ m.T).max() < 1e-11


def test_residual_and_precondition_kernels():
    import torch

    from amaterasu_amd import ops
    from oracle.eigen import teter_payne_allan

    rng = np.random.default_rng(3)
    v = rng.normal(size=(7, 500)) + 1j * rng.normal(size=(7, 500))
    hv = rng.normal(size=(7, 500)) + 1j * rng.normal(size=(7, 500))
    eps = rng.normal(size=7)
    kin = np.linspace(0.0, 20.0, 500)
    res, norms = ops.residual(torch.from_numpy(hv).cuda(), torch.from_numpy(v).cuda(), torch.from_numpy(eps))
    ref = hv - eps[:, None] * v
    assert np.abs(res.cpu().numpy() - ref).max() < 1e-14
    assert np.abs(norms.cpu().numpy() - np.linalg.norm(ref, axis=1)).max() < 1e-12
    rows = np.array([0, 3, 6])
    out = ops.precondition(res, torch.from_numpy(v).cuda(), torch.from_numpy(kin), torch.from_numpy(rows))
    tb = np.einsum("ng,g->n", np.abs(v[rows]) ** 2, kin)
    c = teter_payne_allan(ref[rows], kin, tb)
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    assert np.abs(out.cpu().numpy() - c).max() < 1e-13


def test_davidson_with_gpu_hamiltonian_matches_oracle(si2_system):
    """The SCF's own call pattern (driver.py:345-353) on both paths from the same start."""
    from amaterasu_amd.eigen import solve
    from amaterasu_amd.hamiltonian import Hamiltonian
    from oracle.eigen import random_bands as orb
    from oracle.eigen import solve as osolve
    from oracle.hamiltonian import Hamiltonian as OH

    v = si2_system.v_local[None]
    h, o = Hamiltonian(si2_system, v), OH(si2_system, v)
    for ik, sp in enumerate(si2_system.wave_spheres):
        guess = orb(len(sp), 8, sp.kinetic_energy, seed=12345 + ik)
        a = solve("davidson", lambda b, ik=ik: h.apply(b, ik), guess, sp.kinetic_energy, tolerance=1e-9,
                  max_iterations=60)
        b = osolve("davidson", lambda x, ik=ik: o.apply(x, ik), guess, sp.kinetic_energy, tolerance=1e-9,
                   max_iterations=60)
        assert a.converged and b.converged
        assert np.abs(a.eigenvalues - b.eigenvalues).max() < 1e-10
        # same occupied subspace: projector difference
        pa, pb = a.coefficients[:4].T @ a.coefficients[:4].conj(), b.coefficients[:4].T @ b.coefficients[:4].conj()
        assert np.abs(pa - pb).max() < 1e-7


@pytest.mark.parametrize("dtype,cond_hard,orth", [("complex64", 1e3, 1e-5), ("complex128", 1e6, 1e-12)])
def test_orthonormalisation_refinement_and_its_fallback(monkeypatch, dtype, cond_hard, orth):
    """After the first Cholesky pass the rows of a well-conditioned block are orthonormal to
    within the Loewdin bound and take the refinement (no second factorisation); with the
    bound at zero the same block takes the second Cholesky pass; a block whose singular
    values spread over cond_hard goes whichever way its |E| sends it. Every result is
    orthonormal and spans the input rows."""
    import torch

    from amaterasu_amd import eigen

    calls = []
    second = eigen._cholesky_orthonormalize
    monkeypatch.setattr(eigen, "_cholesky_orthonormalize", lambda v: calls.append(1) or second(v))
    rng = np.random.default_rng(3)
    n, m = 4000, 24
    tdt = getattr(torch, dtype)

    def block_of(cond):
        u, _ = np.linalg.qr(rng.normal(size=(n, m)) + 1j * rng.normal(size=(n, m)))
        w, _ = np.linalg.qr(rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m)))
        return (w * np.geomspace(1.0, 1.0 / cond, m)) @ u.conj().T  # singular values 1 .. 1/cond

    def check(block, expect_calls):
        calls.clear()
        dev = torch.from_numpy(block).to(tdt).cuda()
        out = eigen._gram_orthonormalize(dev, 1e-7)
        if expect_calls is not None:
            assert len(calls) == expect_calls
        o = out.to(torch.complex128).cpu().numpy()
        b = dev.to(torch.complex128).cpu().numpy()
        assert o.shape == (m, n)
        assert np.abs(o @ o.conj().T - np.eye(m)).max() < orth
        assert np.abs((b @ o.conj().T) @ o - b).max() < (1e-3 if dtype == "complex64" else 1e-9)

    easy = block_of(2.0)
    check(easy, 0)
    monkeypatch.setitem(eigen._LOWDIN_MAX, tdt, 0.0)
    check(easy, 1)
    monkeypatch.undo()
    monkeypatch.setattr(eigen, "_cholesky_orthonormalize", lambda v: calls.append(1) or second(v))
    check(block_of(cond_hard), None)
