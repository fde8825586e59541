"""Block Davidson, subspace utilities and the Teter-Payne-Allan preconditioner (FP64, CPU).

Restates ``amaterasu/scf/eigen/davidson.py``, ``subspace.py``, ``precond.py`` and
``__init__.py`` (the norm-conserving ``davidson`` branch). LAPACK calls go through
scipy.linalg exactly as in the reference, so the same factorisations are used.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.linalg import cholesky, eigh, solve_triangular


@dataclass
class DavidsonResult:
    """``davidson.py:33-45``."""

    eigenvalues: np.ndarray
    coefficients: np.ndarray
    residual_norms: np.ndarray
    iterations: int
    converged: bool

    @property
    def nbands(self) -> int:
        return len(self.eigenvalues)


def teter_payne_allan(residual, kinetic, reference):
    """K(x) = P/(P + 16 x^4), P = 27 + 18x + 12x^2 + 8x^3, x = T_G / max(T_band, 1e-8)
    (``precond.py:27-46``)."""
    ref = np.maximum(np.atleast_1d(np.asarray(reference, dtype=float)), 1e-8)
    x = kinetic[None, :] / ref[:, None]
    p = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x))
    return residual * (p / (p + 16.0 * x**4))


def _svd_orthonormalize(block, threshold=1e-10):
    """``subspace.py:35-38``."""
    _, s, vh = np.linalg.svd(block, full_matrices=False)
    return vh[s > threshold * max(float(s[0]), 1e-300)]


def orthonormalize(block):
    """Cholesky of block block^H, then a lower triangular solve (``subspace.py:17-32``)."""
    block = np.asarray(block, dtype=complex)
    s = block @ block.conj().T
    try:
        lower = cholesky(s, lower=True)
    except np.linalg.LinAlgError:
        return _svd_orthonormalize(block)
    out = solve_triangular(lower, block, lower=True)
    if not np.all(np.isfinite(out)):
        return _svd_orthonormalize(block)
    return out


def orthogonalize_against(block, basis, *, threshold=1e-7):
    """Two projections off ``basis``, drop norms <= threshold, SVD-orthonormalise and keep
    singular values > threshold * s0 (``subspace.py:41-60``)."""
    block = np.asarray(block, dtype=complex)
    for _ in range(2):
        block = block - (block @ basis.conj().T) @ basis
    block = block[np.linalg.norm(block, axis=1) > threshold]
    if len(block) == 0:
        return block
    _, s, vh = np.linalg.svd(block, full_matrices=False)
    return vh[s > threshold * float(s[0])]


def rayleigh_ritz(basis, h_basis, nbands):
    """Full projected matrix conj(V) HV^T, Hermitised, eigh, lowest nbands
    (``subspace.py:63-76``)."""
    proj = basis.conj() @ h_basis.T
    proj = 0.5 * (proj + proj.conj().T)
    w, v = eigh(proj)
    rot = v[:, :nbands].T
    return w[:nbands], rot @ basis, rot @ h_basis


def random_bands(npw, nbands, kinetic, seed=0):
    """default_rng(seed): normal real part, then normal imaginary part, damped by
    1/(1+T) and orthonormalised (``subspace.py:79-88``)."""
    rng = np.random.default_rng(seed)
    block = rng.normal(size=(nbands, npw)) + 1j * rng.normal(size=(nbands, npw))
    return orthonormalize(block / (1.0 + kinetic))


def davidson(apply_hamiltonian, guess, kinetic, *, tolerance=1e-8, max_iterations=40, subspace_factor=4):
    """``davidson.py:48-111``."""
    basis = orthonormalize(np.asarray(guess, dtype=complex))
    nbands, npw = basis.shape
    max_subspace = min(max(subspace_factor, 2) * nbands, npw)
    h_basis = apply_hamiltonian(basis)
    eps, vec, hvec = rayleigh_ritz(basis, h_basis, nbands)
    norms = np.full(nbands, np.inf)
    it = 0
    for step in range(1, max_iterations + 1):
        it = step
        res = hvec - eps[:, None] * vec
        norms = np.linalg.norm(res, axis=1)
        active = norms > tolerance
        if not np.any(active):
            break
        tband = np.einsum("ng,g->n", np.abs(vec[active]) ** 2, kinetic)
        corr = teter_payne_allan(res[active], kinetic, tband)
        corr /= np.maximum(np.linalg.norm(corr, axis=1, keepdims=True), 1e-300)
        new = orthogonalize_against(corr, basis)
        if len(new) == 0:
            break
        basis = np.vstack([basis, new])
        h_basis = np.vstack([h_basis, apply_hamiltonian(new)])
        eps, vec, hvec = rayleigh_ritz(basis, h_basis, nbands)
        if len(basis) + nbands > max_subspace:
            basis, h_basis = vec, hvec
    return DavidsonResult(eps, vec, norms, it, bool(np.all(norms <= tolerance)))


def solve(name, apply_hamiltonian, guess, kinetic, *, apply_overlap=None, **kwargs):
    """``eigen/__init__.py:31-46`` (norm-conserving branch only)."""
    if name != "davidson":
        raise NotImplementedError(f"eigensolver {name!r} is not implemented yet; use 'davidson' (scf.eigensolver)")
    if apply_overlap is not None:
        raise NotImplementedError("the oracle restates the norm-conserving path only")
    return davidson(apply_hamiltonian, guess, kinetic, **kwargs)
