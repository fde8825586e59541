"""Host-side parts of the product path against the oracle: the sphere order (which fixes
every coefficient layout, so it must be bit-identical), the grid rule, the psp8 reader,
the radial transforms and the projector assembly (torch on the CPU here)."""

from __future__ import annotations

import numpy as np
import pytest

from tests.conftest import DATA, si_primitive_bohr


@pytest.mark.parametrize("ecut,k", [(8.0, (0.0, 0.0, 0.0)), (15.0, (0.25, 0.25, 0.25)), (20.0, (0.1, -0.37, 0.42))])
def test_sphere_order_is_bit_identical(ecut, k):
    from amaterasu_amd.basis import GSphere as P
    from oracle.basis import GSphere as O

    cell, _ = si_primitive_bohr()
    a, b = P.build(cell, ecut, kpoint=k), O.build(cell, ecut, kpoint=k)
    assert np.array_equal(a.miller, b.miller)
    assert np.array_equal(a.g2, b.g2)


def test_workload_spheres_and_grids_match_the_oracle():
    from amaterasu_amd.basis import GSphere as P
    from amaterasu_amd.basis import grid_shape_for_cutoff
    from amaterasu_amd.workloads import si8_kbatch
    from oracle.basis import FFTGrid, GSphere as O

    spec = si8_kbatch(nk=4)
    assert grid_shape_for_cutoff(spec.cell, spec.ecut_rho) == FFTGrid.for_cutoff(spec.cell, spec.ecut_rho).shape == (42, 42, 42)
    for k, e in zip(spec.kpoints, spec.ecuts):
        assert np.array_equal(P.build(spec.cell, e, kpoint=k).miller, O.build(spec.cell, e, kpoint=k).miller)


def test_kbatch_plane_wave_counts_are_ragged():
    from amaterasu_amd.basis import GSphere
    from amaterasu_amd.workloads import si8_kbatch

    spec = si8_kbatch()
    npw = np.array([len(GSphere.build(spec.cell, e, kpoint=k)) for k, e in zip(spec.kpoints, spec.ecuts)])
    assert len(npw) == 16
    assert 0.08 < (npw.max() - npw.min()) / npw.mean() < 0.17
    assert 4400 < npw.mean() < 4800


def test_psp8_reader_matches_the_oracle():
    from amaterasu_amd.pseudo import read_psp8 as P
    from oracle.pseudo import read_psp8 as O

    a, b = P(DATA / "Si.pbe.psp8.gz"), O(DATA / "Si.pbe.psp8.gz")
    assert np.array_equal(a.r, b.grid.r)
    assert np.array_equal(a.local_potential, b.local_potential)
    assert np.array_equal(a.core_density, b.core_density)
    assert np.array_equal(a.valence_density, b.valence_density)
    assert [(c.l, c.index, c.energy) for c in a.projectors] == [(c.l, c.index, c.energy) for c in b.projectors]
    for c, d in zip(a.projectors, b.projectors):
        assert np.array_equal(c.values, d.values)


def test_projectors_match_the_oracle():
    import torch

    from amaterasu_amd.basis import GSphere as P
    from amaterasu_amd.pseudo import build_projectors as pb
    from amaterasu_amd.pseudo import read_psp8 as pr
    from oracle.basis import GSphere as O
    from oracle.pseudo import build_projectors as ob
    from oracle.pseudo import read_psp8 as orr

    cell, pos = si_primitive_bohr()
    vol = abs(np.linalg.det(cell))
    k = (0.1, 0.2, -0.3)
    a = pb(P.build(cell, 8.0, kpoint=k), pos, ["Si", "Si"], {"Si": pr(DATA / "Si.pbe.psp8.gz")}, vol, device="cpu")
    b = ob(O.build(cell, 8.0, kpoint=k), pos, ["Si", "Si"], {"Si": orr(DATA / "Si.pbe.psp8.gz")}, vol)
    assert a.labels == b.labels
    assert np.array_equal(a.dij, b.dij)
    m = a.matrix.numpy()
    assert isinstance(a.matrix, torch.Tensor)
    assert np.abs(m - b.matrix).max() <= 1e-13 * np.abs(b.matrix).max()


def test_default_band_count():
    from amaterasu_amd.system import default_band_count

    # test_hamiltonian.py:388-391
    assert default_band_count(8, 1) == 8
    assert default_band_count(100, 1) == 60
    assert default_band_count(16, 2, moment=2.2) > default_band_count(16, 2, moment=0.0)


def test_smooth_potential_is_seeded_and_normalised():
    from amaterasu_amd.workloads import smooth_potential

    v = smooth_potential((12, 12, 12), seed=3)
    assert np.array_equal(v, smooth_potential((12, 12, 12), seed=3))
    assert np.sqrt(np.mean(v**2)) == pytest.approx(1.0)
    assert abs(v.mean()) < 1e-12


def test_occupation_errors_follow_the_reference():
    """occupations.py:187-196: a non-(nspin, nk, nbands) array is a ValueError, the
    tetrahedron scheme a NotImplementedError (host tensors; no GPU needed)."""
    import torch

    from amaterasu_amd.scf import occupations

    eig = torch.zeros((1, 2, 4), dtype=torch.float64)
    w = torch.tensor([0.5, 0.5], dtype=torch.float64)
    with pytest.raises(ValueError):
        occupations(eig[0], w, 4.0, "gaussian", 0.01, 1)
    with pytest.raises(NotImplementedError):
        occupations(eig, w, 4.0, "tetrahedron", 0.01, 1)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and says so, before touching
    a device."""
    import subprocess
    import sys
    from pathlib import Path

    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=root)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
