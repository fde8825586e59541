"""SCF result files (the reference's HDF5 layout, format version 2) on the host: the
datasets and attributes the reference's ``read_scf`` reads (``io/results.py:211-240``) are
where it looks for them, in its output units, with h5py's type conventions; a file read
back gives the written values. The reference cannot be run here (h5py is absent), so the
layout is checked name by name against its writer and the unit factors against the published
CODATA 2014 Hartree energy and Bohr radius (what ASE, and so the reference, uses) and the
29 421 GPa per Ha/Bohr^3 the reference's units module quotes; stress positive under compression."""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest

from amaterasu_amd import h5

pytestmark = pytest.mark.skipif(not h5.available(), reason="no libhdf5 on this machine")


def _case(tmp_path, with_wavefunctions=True):
    from amaterasu_amd.basis import FFTGrid, GSphere
    from amaterasu_amd.scf import Energies, SCFResult
    from amaterasu_amd.system import KPoints, load_dataset

    from conftest import si_primitive_bohr

    cell, pos = si_primitive_bohr()
    kpts = np.array([[0.25, 0.25, 0.25], [-0.25, 0.25, 0.25]])
    spheres = [GSphere.build(cell, 4.0, kpoint=k) for k in kpts]
    grid = FFTGrid((12, 12, 12), cell)
    system = SimpleNamespace(cell=cell, positions=pos, symbols=["Si", "Si"], datasets={"Si": load_dataset()},
                             grid=grid, wave_spheres=spheres, kpoints=KPoints(kpts, np.array([0.25, 0.75])),
                             xc=SimpleNamespace(name="PBE"), ecut_rho=16.0, extra={})
    rng = np.random.default_rng(4)
    wfs = [[rng.normal(size=(4, len(s.g2))) + 1j * rng.normal(size=(4, len(s.g2))) for s in spheres]]
    result = SCFResult(converged=True, iterations=9,
                       energies=Energies(kinetic=3.1, local=-2.2, hartree=0.5, xc=-2.4, nonlocal_=1.1, ewald=-8.4,
                                         entropy=0.01),
                       eigenvalues=rng.normal(size=(1, 2, 4)), occupations=np.full((1, 2, 4), 1.0),
                       fermi_level=0.21, gap=0.02, density=rng.random((1, 12, 12, 12)), wavefunctions=wfs,
                       history=[(-7.9, 0.1, 1e-3), (-7.95, 0.01, 1e-4)], electron_count=8.0, seconds=1.5)
    forces = rng.normal(size=(2, 3)) * 1e-2
    stress = np.diag([1e-4, 2e-4, 3e-4]) + 1e-5
    path = tmp_path / "out" / "scf.h5"
    from amaterasu_amd.results import write_scf

    write_scf(path, system, result, config={"xc": {"functional": "PBE"}}, forces=forces, stress=stress,
              save_wavefunctions=with_wavefunctions)
    return path, system, result, forces, stress


def test_written_file_reads_back(tmp_path):
    from amaterasu_amd.results import BOHR_IN_ANG, HARTREE_IN_EV, read_scf

    path, system, result, forces, stress = _case(tmp_path)
    r = read_scf(path)
    assert np.allclose(r["eigenvalues"], result.eigenvalues * HARTREE_IN_EV, rtol=0, atol=1e-13)
    assert np.array_equal(r["occupations"], result.occupations)
    assert np.allclose(r["density"], result.density, rtol=1e-15, atol=0)  # electrons/Bohr^3 again
    assert np.allclose(r["cell"], system.cell * BOHR_IN_ANG, rtol=1e-15)
    assert r["symbols"] == ["Si", "Si"]
    assert np.array_equal(r["kpoints"], system.kpoints.points) and np.array_equal(r["weights"], system.kpoints.weights)
    assert r["energies"]["total"] == pytest.approx(result.energies.total * HARTREE_IN_EV, rel=1e-15)
    assert r["energies"]["extrapolated"] == pytest.approx((result.energies.total - 0.005) * HARTREE_IN_EV)
    assert r["energies"]["nonlocal"] == pytest.approx(1.1 * HARTREE_IN_EV)  # trailing underscore dropped
    assert len(r["wavefunctions"]) == 1 and len(r["wavefunctions"][0]) == 2
    for got, want in zip(r["wavefunctions"][0], result.wavefunctions[0]):
        assert got.dtype == np.complex128 and np.array_equal(got, want)


def test_units_and_signs(tmp_path):
    from amaterasu_amd.results import read_scf

    path, system, result, forces, stress = _case(tmp_path, with_wavefunctions=False)
    r = read_scf(path)
    ha_ev, a0 = 27.21138602, 0.52917721067  # CODATA 2014 E_h (eV) and a_0 (A), the values ASE uses
    assert np.allclose(r["forces"] / forces, ha_ev / a0, rtol=1e-8)
    assert np.allclose(r["stress"] / stress, -ha_ev / a0**3, rtol=1e-8)  # positive under compression
    with h5.open_read(path) as f:
        a = f.dataset_attrs("stress")
        assert a["convention"] == "positive_under_compression" and a["units"] == "eV/Angstrom^3"
        assert a["pressure_GPa"] == pytest.approx(-np.trace(stress) / 3 * 29421.0, rel=1e-4)
        assert "wavefunctions" not in f


def test_layout_matches_the_reference_writer(tmp_path):
    path, system, result, forces, stress = _case(tmp_path)
    with h5.open_read(path) as f:
        for name in ("structure/cell", "structure/positions", "structure/numbers", "structure/symbols",
                     "structure/initial_magmoms", "basis/grid_shape", "basis/plane_waves", "kpoints/points",
                     "kpoints/weights", "scf/eigenvalues", "scf/occupations", "scf/history", "forces", "stress",
                     "density", "wavefunctions/spin0/kpoint0", "wavefunctions/spin0/miller1"):
            assert name in f, name
        at = f.attrs.to_dict()
        assert at["format_version"] == 2 and at["stress_convention"] == "positive_under_compression"
        assert {at[f"unit_{q}"] for q in ("energy", "length", "force", "stress", "pressure", "time")} == \
            {"eV", "Angstrom", "eV/Angstrom", "eV/Angstrom^3", "GPa", "fs"}
        assert at["solver"] == "scf" and at["functional"] == "PBE" and at["converged"] is np.True_
        assert "xc:" in at["resolved_input"]
        assert list(f.read("structure/numbers")) == [14, 14]
        assert list(f.read("basis/grid_shape")) == [12, 12, 12]
        assert list(f.read("basis/plane_waves")) == [len(s.g2) for s in system.wave_spheres]
        with f.group("basis") as g:
            assert g.attrs["units"] == "eV" and g.attrs["ecut_rho"] == pytest.approx(16.0 * 27.211386024367243)
        with f.group("scf") as g:
            s = g.attrs.to_dict()
            assert s["iterations"] == 9 and s["fermi_level"] == pytest.approx(0.21 * 27.211386024367243)
            assert {"seconds", "electron_count", "magnetization", "gap"} <= set(s)
        with f.group("scf/energies") as g:
            assert set(g.attrs.keys()) == {"units", "kinetic", "local", "hartree", "xc", "nonlocal", "ewald",
                                          "entropy", "total", "free_energy", "extrapolated"}
        hist = f.read("scf/history")
        assert hist[0, 0] == pytest.approx(-7.9 * 27.211386024367243) and hist[0, 1] == 0.1
        assert f.dataset_attrs("density")["units"] == "electrons/Angstrom^3"
        assert f.read("wavefunctions/spin0/miller1").shape == (len(system.wave_spheres[1].g2), 3)


def test_h5py_type_conventions(tmp_path):
    """Bool as the int8 enum {FALSE, TRUE}, complex as the compound {r, i}, strings as
    variable-length UTF-8: the types h5py writes and reads back as numpy bool / complex / str."""
    lib = h5._load()
    path, *_ = _case(tmp_path)
    with h5.open_read(path) as f:
        aid = lib.H5Aopen(f.id, b"converged", 0)
        t = lib.H5Aget_type(aid)
        assert lib.H5Tget_class(t) == h5.H5T_ENUM and lib.H5Tget_size(t) == 1
        lib.H5Tclose(t)
        lib.H5Aclose(aid)
        did = lib.H5Dopen2(f.id, b"wavefunctions/spin0/kpoint0", 0)
        t = lib.H5Dget_type(did)
        assert lib.H5Tget_class(t) == h5.H5T_COMPOUND and lib.H5Tget_size(t) == 16
        lib.H5Tclose(t)
        lib.H5Dclose(did)
        did = lib.H5Dopen2(f.id, b"structure/symbols", 0)
        t = lib.H5Dget_type(did)
        assert lib.H5Tget_class(t) == h5.H5T_STRING and lib.H5Tis_variable_str(t) > 0
        lib.H5Tclose(t)
        lib.H5Dclose(did)
