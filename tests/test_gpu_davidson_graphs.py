"""Davidson passes replayed from captured HIP graphs (eigen.StepGraphs) against the same
passes run eagerly.

A graph holds the launches between the two host reads of one pass (scf/eigen/davidson.py
lines 79-101 minus the active-set and orthogonalisation decisions): H on the new rows, the
projected matrix's new columns, the Rayleigh-Ritz eigensolve and rotations, the residuals.
Replays run the eager kernels in the eager order, so every number must come out the same;
the tolerances below only allow for the last bit. The projected-matrix threshold of the
library eigensolver is lowered so that every pass of these small cases goes through it
(torch.linalg.eigh reads its status on the host and cannot be captured). Several Davidson
calls replay the cached graphs, with eager work and synchronisations between them, and
each call is checked against the eager result."""

from __future__ import annotations

import pytest
import torch

pytestmark = pytest.mark.gpu


def _system(precision):
    from amaterasu_amd.hamiltonian import Hamiltonian
    from amaterasu_amd.system import build_system, load_dataset
    from amaterasu_amd.workloads import si8_kbatch, smooth_potential

    spec = si8_kbatch(nk=2, nbands=16, ecut=8.0, spread=0.08)
    s = build_system(spec.cell, spec.positions, spec.symbols, {"Si": load_dataset()}, spec.ecut_rho, spec.kpoints,
                     spec.ecuts)
    v = smooth_potential(s.grid.shape, seed=spec.seed, rms=0.05)
    return s, Hamiltonian(s, v, precision=precision)


@pytest.mark.parametrize("precision,tol,same", [("fp64", 1e-6, 1e-12), ("fp32", 2e-5, 1e-6)])
def test_graph_replayed_passes_equal_eager(monkeypatch, precision, tol, same):
    from amaterasu_amd import eigen
    from amaterasu_amd.eigen import clear_graphs, davidson, graph_stats, random_bands

    monkeypatch.setattr(eigen, "HEEV_MIN", 16)
    clear_graphs()
    s, ham = _system(precision)
    plan = ham.plan
    cdt = torch.complex128 if precision == "fp64" else torch.complex64
    kin = plan.kinetic[1]
    nb = 16
    try:
        for k, sph in enumerate(s.wave_spheres):
            n = plan.npw[k]
            guess = random_bands(n, nb, sph.g2 / 2, seed=31 + k, device=True).to(cdt)
            run = dict(tolerance=tol, max_iterations=60, precision=precision)
            ref = davidson(lambda b, k=k: ham.apply(b, k), guess, kin[k, :n], **run)
            assert ref.iterations > 3
            for call in range(3):
                got = davidson(lambda b, k=k: ham.apply(b, k), guess, kin[k, :n], graph_key=("test", id(ham), k),
                               **run)
                torch.cuda.synchronize()
                assert got.iterations == ref.iterations and got.converged == ref.converged
                assert torch.abs(got.eigenvalues - ref.eigenvalues).max().item() <= same
                assert torch.abs(got.coefficients - ref.coefficients).max().item() <= 1e3 * same
                assert torch.abs(got.residual_norms - ref.residual_norms).max().item() <= 1e3 * same
                # eager work between replays, on the same buffers' neighbours
                again = davidson(lambda b, k=k: ham.apply(b, k), guess, kin[k, :n], **run)
                assert torch.equal(again.eigenvalues, ref.eigenvalues)
        stats = graph_stats()
        assert stats["captures"] > 0
        assert stats["replays"] > stats["captures"]
    finally:
        clear_graphs()


def test_scf_with_graph_replayed_davidson_equals_eager(si2_system, monkeypatch):
    """The SCF's per-k Davidson (k-batching off) with the potential in the plan's shared
    buffers and graph-replayed passes: the same energies and eigenvalues as the eager SCF,
    in the mixed-precision policy (both FP32 and FP64 passes)."""
    import copy

    from amaterasu_amd import eigen
    from amaterasu_amd.eigen import clear_graphs, graph_stats
    from amaterasu_amd.scf import SCFSettings, run_scf

    monkeypatch.setattr(eigen, "HEEV_MIN", 8)
    s = copy.copy(si2_system)
    kw = dict(tol_energy=1e-9, tol_density=1e-8, precision="mixed", batched_k=False)
    clear_graphs()
    try:
        ref = run_scf(s, SCFSettings(davidson_graphs=False, **kw))
        got = run_scf(s, SCFSettings(davidson_graphs=True, **kw))
        assert ref.converged and got.converged
        assert got.iterations == ref.iterations
        assert abs(got.energies.total - ref.energies.total) < 1e-12
        assert abs(got.eigenvalues - ref.eigenvalues).max() < 1e-11
        assert graph_stats()["replays"] > 0
    finally:
        clear_graphs()
