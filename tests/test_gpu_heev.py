"""The library's Hermitian eigensolver (device tridiagonalisation, tridiagonal divide and
conquer, device back-transformation) against LAPACK (numpy.linalg.eigh, FP64) on random
Hermitian matrices and on matrices with exactly degenerate eigenvalues."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _hermitian(n, seed, degenerate=False):
    rng = np.random.default_rng(seed)
    if degenerate:  # eigenvalues 0..n/4 each four times, random unitary basis
        lam = np.repeat(np.arange(n // 4 + 1, dtype=float), 4)[:n]
        q, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
        return (q * lam) @ q.conj().T
    x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return (x + x.conj().T) / 2


@pytest.mark.parametrize("n,nev", [(2, 1), (3, 3), (17, 5), (64, 20), (65, 65), (66, 30), (160, 40), (333, 100),
                                   (640, 160), (1024, 256), (1025, 300), (1760, 440), (2048, 512)])
def test_heev_matches_lapack(n, nev):
    from amaterasu_amd import ops

    a = _hermitian(n, n)
    w, zt = ops.heev_lowest(torch.from_numpy(a).cuda(), nev)
    w, z = w.cpu().numpy(), zt.cpu().numpy().T
    ref = np.linalg.eigvalsh(a)[:nev]
    scale = np.abs(np.linalg.eigvalsh(a)).max()
    assert np.abs(w - ref).max() < 1e-13 * scale * max(1, n / 100)
    assert np.abs(a @ z - z * w).max() < 1e-12 * scale * max(1, n / 100)
    assert np.abs(z.conj().T @ z - np.eye(nev)).max() < 1e-12 * max(1, n / 100)


@pytest.mark.parametrize("n", [40, 130])
def test_heev_zero_columns_in_the_last_steps(n):
    """A Hermitian block over a diagonal block: the last columns of the reduction (the last 64
    run in one workgroup) have nothing below the diagonal, so their reflectors are H = I
    (tau = 0), and the eigenvalues of the diagonal block are exact."""
    from amaterasu_amd import ops

    rng = np.random.default_rng(n)
    a = np.zeros((n, n), dtype=complex)
    h = n - 30
    a[:h, :h] = _hermitian(h, n + 1)
    a[h:, h:] = np.diag(rng.uniform(-3.0, 3.0, 30))
    nev = n // 2
    w, zt = ops.heev_lowest(torch.from_numpy(a).cuda(), nev)
    w, z = w.cpu().numpy(), zt.cpu().numpy().T
    assert np.abs(w - np.linalg.eigvalsh(a)[:nev]).max() < 1e-13 * 10
    assert np.abs(a @ z - z * w).max() < 1e-12 * 10
    assert np.abs(z.conj().T @ z - np.eye(nev)).max() < 1e-12


def test_heev_degenerate_clusters():
    from amaterasu_amd import ops

    a = _hermitian(200, 7, degenerate=True)
    w, zt = ops.heev_lowest(torch.from_numpy(a).cuda(), 60)
    w, z = w.cpu().numpy(), zt.cpu().numpy().T
    assert np.abs(w - np.linalg.eigvalsh(a)[:60]).max() < 1e-12 * 50
    assert np.abs(a @ z - z * w).max() < 1e-11 * 50
    assert np.abs(z.conj().T @ z - np.eye(60)).max() < 1e-12


def test_heev_leaves_input_and_rejects_bad_sizes():
    from amaterasu_amd import ops
    from amaterasu_amd._lib import AmtError

    a = torch.from_numpy(_hermitian(50, 3)).cuda()
    keep = a.clone()
    ops.heev_lowest(a, 10)
    assert torch.equal(a, keep)
    with pytest.raises(AmtError):
        ops.heev_lowest(torch.zeros((2100, 2100), dtype=torch.complex128, device="cuda"), 4)
    with pytest.raises(AmtError):
        ops.heev_lowest(a, 51)


def test_heev_repeated_calls_replay_correctly():
    """Consecutive calls of one size on different matrices, with other sizes in between and
    the workspace reused, must each match LAPACK."""
    from amaterasu_amd import ops

    for seed in (11, 12, 13):
        for n in (200, 120):
            a = _hermitian(n, seed * n)
            w, zt = ops.heev_lowest(torch.from_numpy(a).cuda(), 30)
            assert np.abs(w.cpu().numpy() - np.linalg.eigvalsh(a)[:30]).max() < 1e-12 * n
            z = zt.cpu().numpy().T
            assert np.abs(a @ z - z * w.cpu().numpy()).max() < 1e-11 * n


def _chain_matrix(n, seed, chain, start, spacing, degenerate_at):
    """Hermitian matrix with a chain of `chain` eigenvalues whose neighbours are closer than
    the solver's cluster threshold (2e-4 ||T||), an exactly degenerate pair at chain
    positions degenerate_at and degenerate_at + 1, and well-separated eigenvalues around it."""
    rng = np.random.default_rng(seed)
    below = -1.0 + 0.01 * np.arange(start)
    steps = spacing(rng, chain - 1)
    steps[degenerate_at] = 0.0
    mid = below[-1] + 0.01 + np.concatenate([[0.0], np.cumsum(steps)])
    above = mid[-1] + 0.01 * (1 + np.arange(n - start - chain))
    lam = np.concatenate([below, mid, above])
    q, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    return (q * lam) @ q.conj().T


@pytest.mark.parametrize("spacing", ["irregular", "uniform", "all-degenerate"])
def test_heev_chain_longer_than_a_cluster(spacing):
    """A chain of 100 close eigenvalues (more than the 64 one workgroup solves together) with
    an exactly degenerate pair at chain positions 63 and 64: the eigenvectors stay orthonormal
    and accurate across the cut the solver has to make. Also an evenly spaced chain and a
    70-fold degenerate eigenvalue."""
    from amaterasu_amd import ops

    n, nev = 300, 200
    if spacing == "irregular":
        a = _chain_matrix(n, 5, 100, 50, lambda rng, k: rng.uniform(1e-5, 3e-5, size=k), 63)
    elif spacing == "uniform":
        a = _chain_matrix(n, 6, 100, 50, lambda rng, k: np.full(k, 2e-5), 63)
    else:
        a = _chain_matrix(n, 7, 70, 50, lambda rng, k: np.zeros(k), 0)
    w, zt = ops.heev_lowest(torch.from_numpy(a).cuda(), nev)
    w, z = w.cpu().numpy(), zt.cpu().numpy().T
    ref = np.linalg.eigvalsh(a)[:nev]
    assert np.abs(w - ref).max() < 1e-13 * 3
    assert np.abs(a @ z - z * w).max() < 1e-12 * 3
    assert np.abs(z.conj().T @ z - np.eye(nev)).max() < 1e-12 * 3
