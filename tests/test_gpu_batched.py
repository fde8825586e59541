"""The k-batched Davidson (one launch per kernel over all k-points) against the per-k
device Davidson, which the Davidson tests pin to the oracle and to dense eigh.

Each k-point follows its own iteration in both, so converged eigenpairs agree to the
eigensolver tolerance; ragged plane-wave counts (+-8 %) exercise the padding. The
reference iteration (collapse to the Ritz vectors, no thick restart) converges the top
band of a block slowly, so the cases use the 16 valence bands of Si8 (a gap above the
block) and tolerances it reaches in well under the iteration limit."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _ragged_system(precision="fp64"):
    from amaterasu_amd.hamiltonian import Hamiltonian
    from amaterasu_amd.system import build_system, load_dataset
    from amaterasu_amd.workloads import si8_kbatch, smooth_potential

    spec = si8_kbatch(nk=4, nbands=16, ecut=8.0, spread=0.08)
    s = build_system(spec.cell, spec.positions, spec.symbols, {"Si": load_dataset()}, spec.ecut_rho, spec.kpoints,
                     spec.ecuts)
    v = smooth_potential(s.grid.shape, seed=spec.seed, rms=0.05)
    return s, Hamiltonian(s, v, precision=precision)


@pytest.mark.parametrize("precision,tol,eig_tol", [("fp64", 1e-6, 1e-10), ("fp32", 2e-5, 5e-6)])
def test_batched_matches_per_k(precision, tol, eig_tol):
    from amaterasu_amd.batched import solve_batched
    from amaterasu_amd.eigen import davidson, random_bands

    s, ham = _ragged_system(precision)
    plan = ham.plan
    assert len(set(plan.npw)) == len(plan.npw)  # ragged
    cdt = torch.complex128 if precision == "fp64" else torch.complex64
    nb = 16
    guess = torch.zeros((plan.nk, nb, plan.npw_pad), dtype=cdt, device="cuda")
    for k, sph in enumerate(s.wave_spheres):
        guess[k, :, : len(sph.g2)] = random_bands(len(sph.g2), nb, sph.g2 / 2, seed=7 + k, device=True).to(cdt)
    kin = plan.kinetic[1 if precision == "fp64" else 0]
    res = solve_batched(lambda b: ham.apply_batched(b), guess, kin, plan.npw, tolerance=tol, max_iterations=100)
    assert res.converged.all()
    for k, sph in enumerate(s.wave_spheres):
        n = len(sph.g2)
        ref = davidson(lambda b, k=k: ham.apply(b, k), guess[k, :, :n].contiguous(), kin[k, :n], tolerance=tol,
                       max_iterations=100, precision=precision)
        assert ref.converged
        assert torch.abs(res.eigenvalues[k] - ref.eigenvalues).max().item() < eig_tol
        c = res.coefficients[k].to(torch.complex128)
        assert c[:, n:].numel() == 0 or float(c[:, n:].abs().max()) == 0.0  # padding stays zero
        ov = c[:, :n] @ c[:, :n].conj().T
        assert torch.abs(ov - torch.eye(nb, dtype=ov.dtype, device=ov.device)).max().item() < 1e-5 * (
            1 if precision == "fp32" else 1e-6)
        assert res.iterations[k] >= 1


def test_batched_frozen_kpoint_keeps_its_answer():
    """A k-point whose guess is already converged returns after one step, untouched by the
    other k-points' further iterations."""
    from amaterasu_amd.batched import solve_batched
    from amaterasu_amd.eigen import davidson, random_bands

    s, ham = _ragged_system()
    plan = ham.plan
    nb = 16
    guess = torch.zeros((plan.nk, nb, plan.npw_pad), dtype=torch.complex128, device="cuda")
    for k, sph in enumerate(s.wave_spheres):
        guess[k, :, : len(sph.g2)] = random_bands(len(sph.g2), nb, sph.g2 / 2, seed=3 + k, device=True)
    kin = plan.kinetic[1]
    n0 = plan.npw[0]
    pre = davidson(lambda b: ham.apply(b, 0), guess[0, :, :n0].contiguous(), kin[0, :n0], tolerance=1e-7,
                   max_iterations=300)
    assert pre.converged
    guess[0, :, :n0] = pre.coefficients
    res = solve_batched(lambda b: ham.apply_batched(b), guess, kin, plan.npw, tolerance=1e-6, max_iterations=100)
    assert res.iterations[0] == 1
    assert res.iterations[1:].min() > 1
    assert torch.abs(res.eigenvalues[0] - pre.eigenvalues).max().item() < 1e-12


def test_library_eigensolver_inside_davidson(monkeypatch):
    """The library's Hermitian eigensolver inside the Davidson loop (projected matrices of
    16..64 rows sent to it by lowering the size threshold): the per-k Davidson and the
    k-batched one, whose per-k solves then run concurrently on side streams with their own
    workspaces, against the per-k Davidson on torch.linalg.eigh. Eigenvalues to 1e-10,
    orthonormal coefficients, and residual norms below the tolerance."""
    from amaterasu_amd import eigen
    from amaterasu_amd.batched import solve_batched
    from amaterasu_amd.eigen import davidson, random_bands

    s, ham = _ragged_system()
    plan = ham.plan
    nb, tol = 16, 1e-6
    guess = torch.zeros((plan.nk, nb, plan.npw_pad), dtype=torch.complex128, device="cuda")
    for k, sph in enumerate(s.wave_spheres):
        guess[k, :, : len(sph.g2)] = random_bands(len(sph.g2), nb, sph.g2 / 2, seed=21 + k, device=True)
    kin = plan.kinetic[1]
    ref = []
    for k in range(plan.nk):
        n = plan.npw[k]
        r = davidson(lambda b, k=k: ham.apply(b, k), guess[k, :, :n].contiguous(), kin[k, :n], tolerance=tol,
                     max_iterations=100)
        assert r.converged
        ref.append(r)
    monkeypatch.setattr(eigen, "HEEV_MIN", 16)
    res = solve_batched(lambda b: ham.apply_batched(b), guess, kin, plan.npw, tolerance=tol, max_iterations=100)
    assert res.converged.all()
    for k in range(plan.nk):
        n = plan.npw[k]
        one = davidson(lambda b, k=k: ham.apply(b, k), guess[k, :, :n].contiguous(), kin[k, :n], tolerance=tol,
                       max_iterations=100)
        assert one.converged and float(one.residual_norms.max()) <= tol
        for got in (one.eigenvalues, res.eigenvalues[k]):
            assert torch.abs(got - ref[k].eigenvalues).max().item() < 1e-10
        for c in (one.coefficients, res.coefficients[k][:, :n]):
            ov = c @ c.conj().T
            assert torch.abs(ov - torch.eye(nb, dtype=ov.dtype, device=ov.device)).max().item() < 1e-11
