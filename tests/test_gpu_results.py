"""Restart through a result file: a converged GPU SCF written in the reference's HDF5
layout and read back restarts from its density (``read_scf(...)["density"]`` as the
``density=`` argument of ``run_scf``): its first density residual is already below 1e-6 and
it reaches the same energy in fewer iterations (the wavefunctions restart from the
reference's seeded random bands, as they do in the reference)."""

from __future__ import annotations

import numpy as np
import pytest

from amaterasu_amd import h5

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not h5.available(), reason="no libhdf5 on this machine")]


def test_restart_from_a_result_file(tmp_path, si2_system):
    from amaterasu_amd.forces import compute_forces
    from amaterasu_amd.results import HARTREE_IN_EV, read_scf, write_scf
    from amaterasu_amd.scf import SCFSettings, run_scf
    from amaterasu_amd.stress import compute_stress
    from amaterasu_amd.system import attach_ionic_fields, build_system, load_dataset

    o = si2_system
    ds = load_dataset()
    s = build_system(o.cell, o.positions, o.symbols, {"Si": ds}, 32.0, o.kpoints.points, 8.0,
                     weights=o.kpoints.weights, grid_shape=o.grid.shape)
    attach_ionic_fields(s, {"Si": ds}, 32.0, xc="PBE")
    settings = SCFSettings(tol_energy=1e-9, tol_density=1e-8)
    r = run_scf(s, settings)
    assert r.converged
    f = compute_forces(s, r.density, r.wavefunctions, r.occupations)
    st = compute_stress(s, r.density, r.wavefunctions, r.occupations)
    path = write_scf(tmp_path / "si2.h5", s, r, forces=f, stress=st, save_wavefunctions=True)
    back = read_scf(path)
    assert np.allclose(back["density"], r.density, rtol=1e-15, atol=0)
    assert back["energies"]["total"] == pytest.approx(r.energies.total * HARTREE_IN_EV, rel=1e-15)
    assert np.array_equal(back["wavefunctions"][0][1], r.wavefunctions[0][1])
    again = run_scf(s, settings, density=back["density"])
    assert again.converged and again.iterations < r.iterations
    assert again.history[0][1] < 1e-6 < r.history[0][1]  # it starts from the converged density
    assert abs(again.energies.total - r.energies.total) < 1e-8
