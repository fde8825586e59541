"""SCF result files in the reference's HDF5 layout (format version 2).

Writes and reads what ``amaterasu/io/results.py:82-240`` does -- the same groups,
datasets, shapes, attribute names and output units -- so a GPU run leaves a file the
reference's ``read_scf`` reads, and a file the reference wrote restarts a GPU run
(``read_scf(path)["density"]`` is the ``density=`` argument of ``run_scf``)::

    /                  provenance, format_version = 2, stress_convention, unit_* attributes,
                       solver, functional, converged[, resolved_input]
    /structure         cell, positions (Angstrom), numbers, symbols, initial_magmoms
    /basis             ecut, ecut_rho (eV) as attributes; grid_shape, plane_waves
    /kpoints           points (fractional), weights; symmetry_reduced, n_full
    /scf               iterations, seconds, electron_count, magnetization, fermi_level[, gap];
                       energies (eV, attributes); eigenvalues (eV), occupations[, history]
    /forces            eV/Angstrom                                   (when given)
    /stress            eV/Angstrom^3, positive under compression       (when given)
    /density           electrons/Angstrom^3, (nspin, *grid), gzip
    /wavefunctions     spin<s>/kpoint<k> (gzip), spin<s>/miller<k>   (on request)

Units follow ``amaterasu/units.py`` (CODATA 2014 through ASE): the stress is converted
with ``stress_to_output`` (``units.py:229-235``), the pressure attribute with ASE's GPa.
The HDF5 side is :mod:`amaterasu_amd.h5` (the system's libhdf5).
"""

from __future__ import annotations

import hashlib
import platform
import sys
from datetime import datetime, timezone
from pathlib import Path

import numpy as np

from . import __version__, h5

FORMAT_VERSION = 2
STRESS_CONVENTION = "positive_under_compression"
OUTPUT_UNITS = {"energy": "eV", "length": "Angstrom", "force": "eV/Angstrom", "stress": "eV/Angstrom^3",
                "pressure": "GPa", "time": "fs"}
HARTREE_IN_EV = 27.211386024367243
BOHR_IN_ANG = 0.5291772105638411
_ELEMENTARY_CHARGE = 1.6021766208e-19  # C, CODATA 2014 (ase.units)
_GPA_IN_EV_PER_ANG3 = 1e9 / (_ELEMENTARY_CHARGE * 1e30)
HARTREE_PER_BOHR_IN_EV_PER_ANG = HARTREE_IN_EV / BOHR_IN_ANG
HARTREE_PER_BOHR3_IN_EV_PER_ANG3 = HARTREE_IN_EV / BOHR_IN_ANG**3
HARTREE_PER_BOHR3_IN_GPA = HARTREE_PER_BOHR3_IN_EV_PER_ANG3 / _GPA_IN_EV_PER_ANG3


def stress_to_output(stress_hartree_per_bohr3):
    """Internal stress (Ha/Bohr^3, (1/V) dE/d eps) to eV/Angstrom^3, positive under compression."""
    return -np.asarray(stress_hartree_per_bohr3, dtype=float) * HARTREE_PER_BOHR3_IN_EV_PER_ANG3


def provenance(pseudopotential_paths: dict | None = None) -> dict:
    """A flat, string-valued record of what wrote the file (``io/provenance.py:50-65``)."""
    record = {"amaterasu_amd_version": __version__,
              "timestamp_utc": datetime.now(timezone.utc).isoformat(timespec="seconds"),
              "python": sys.version.split()[0], "platform": platform.platform()}
    for symbol, path in (pseudopotential_paths or {}).items():
        record[f"pseudopotential_{symbol}"] = str(path)
        try:
            record[f"pseudopotential_{symbol}_md5"] = hashlib.md5(Path(path).read_bytes()).hexdigest()
        except OSError:
            record[f"pseudopotential_{symbol}_md5"] = "unavailable"
    return record


def _file_attrs(handle, pseudopotential_paths=None) -> None:
    for key, value in provenance(pseudopotential_paths).items():
        handle.attrs[key] = value
    handle.attrs["format_version"] = FORMAT_VERSION
    handle.attrs["stress_convention"] = STRESS_CONVENTION
    for quantity, unit in OUTPUT_UNITS.items():
        handle.attrs[f"unit_{quantity}"] = unit


def _described(units: str, description: str) -> dict:
    return {"units": units, "description": description}


def _numbers(system) -> np.ndarray:
    atoms = getattr(system, "atoms", None)
    if atoms is not None:
        return np.asarray(atoms.numbers)
    ds = getattr(system, "datasets", None) or getattr(system, "pseudopotentials", None) or {}
    return np.array([int(round(float(ds[s].atomic_number))) for s in system.symbols])


def _initial_magmoms(system) -> np.ndarray:
    m = getattr(system, "initial_magmoms", None)
    if m is None and isinstance(getattr(system, "extra", None), dict):
        m = system.extra.get("initial_magmoms")
    return np.zeros(len(system.symbols)) if m is None else np.asarray(m, dtype=float)


def _ecut(system) -> float:
    e = getattr(system, "ecut", None)
    return float(e) if e is not None else max(float(s.ecut) for s in system.wave_spheres)


def _ecut_rho(system) -> float:
    e = getattr(system, "ecut_rho", None)
    return float(e) if e is not None else float(system.density_sphere.ecut)


def _yaml(config) -> str:
    if isinstance(config, str):
        return config
    if hasattr(config, "to_yaml"):
        return config.to_yaml()
    import yaml

    return yaml.safe_dump(config, sort_keys=False)


def _block(w):
    return w.detach().cpu().numpy() if hasattr(w, "detach") else np.asarray(w)


def write_scf(path, system, result, *, config=None, pseudopotential_paths: dict | None = None,
              save_wavefunctions: bool = False, forces=None, stress=None) -> Path:
    """Write an SCF result (``io/results.py:82-210``); ``forces`` in Ha/Bohr and ``stress`` in
    Ha/Bohr^3 (internal sign) are optional and converted here. Returns the path."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with h5.create(path) as f:
        _file_attrs(f, pseudopotential_paths)
        f.attrs["solver"] = "scf"
        xc = getattr(system, "xc", None)
        f.attrs["functional"] = str(getattr(xc, "name", xc))
        f.attrs["converged"] = np.bool_(bool(result.converged))
        if config is not None:
            f.attrs["resolved_input"] = _yaml(config)

        f.create_dataset("structure/cell", np.asarray(system.cell, dtype=float) * BOHR_IN_ANG,
                         attrs=_described("Angstrom", "lattice vectors as rows"))
        f.create_dataset("structure/positions", np.asarray(system.positions, dtype=float) * BOHR_IN_ANG,
                         attrs=_described("Angstrom", "Cartesian atomic positions"))
        f.create_dataset("structure/numbers", _numbers(system))
        f.create_dataset("structure/symbols", np.array(list(system.symbols), dtype=object))
        f.create_dataset("structure/initial_magmoms", _initial_magmoms(system),
                         attrs=_described("Bohr magneton", "initial moments the SCF was started from"))

        basis = f.create_group("basis")
        basis.attrs["ecut"] = _ecut(system) * HARTREE_IN_EV
        basis.attrs["ecut_rho"] = _ecut_rho(system) * HARTREE_IN_EV
        basis.attrs["units"] = "eV"
        basis.create_dataset("grid_shape", np.asarray(system.grid.shape))
        basis.create_dataset("plane_waves", np.asarray([len(s.g2) for s in system.wave_spheres]))
        basis.close()

        kp = system.kpoints
        f.create_dataset("kpoints/points", np.asarray(kp.points, dtype=float),
                         attrs=_described("fractional", "coordinates in the reciprocal cell"))
        f.create_dataset("kpoints/weights", np.asarray(kp.weights, dtype=float))
        with f.group("kpoints") as g:
            g.attrs["symmetry_reduced"] = np.bool_(bool(getattr(kp, "symmetry_reduced", False)))
            g.attrs["n_full"] = int(getattr(kp, "n_full", len(kp.points)))

        scf = f.create_group("scf")
        scf.attrs["iterations"] = int(result.iterations)
        scf.attrs["seconds"] = float(getattr(result, "seconds", 0.0))
        scf.attrs["electron_count"] = float(result.electron_count)
        scf.attrs["magnetization"] = float(getattr(result, "magnetization", 0.0))
        en = result.energies
        energies = scf.create_group("energies")
        energies.attrs["units"] = "eV"
        for name in ("kinetic", "local", "hartree", "xc", "nonlocal_", "ewald", "entropy"):
            energies.attrs[name.rstrip("_")] = float(getattr(en, name)) * HARTREE_IN_EV
        energies.attrs["total"] = float(en.total) * HARTREE_IN_EV
        energies.attrs["free_energy"] = float(en.free_energy) * HARTREE_IN_EV
        energies.attrs["extrapolated"] = float(en.total - 0.5 * en.entropy) * HARTREE_IN_EV
        energies.close()
        scf.create_dataset("eigenvalues", np.asarray(result.eigenvalues, dtype=float) * HARTREE_IN_EV,
                           attrs=_described("eV", "(nspin, nkpoints, nbands)"))
        scf.create_dataset("occupations", np.asarray(result.occupations, dtype=float),
                           attrs=_described("dimensionless", "occupation of each state, between 0 and 1"))
        scf.attrs["fermi_level"] = float(result.fermi_level) * HARTREE_IN_EV
        if result.gap is not None:
            scf.attrs["gap"] = float(result.gap) * HARTREE_IN_EV
        history = np.asarray(result.history, dtype=float)
        if history.size:
            history = history.copy()
            history[:, 0] *= HARTREE_IN_EV
            history[:, 2] *= HARTREE_IN_EV
            scf.create_dataset("history", history,
                               attrs=_described("eV, electrons per electron, eV",
                                                "per iteration: free energy, density residual, "
                                                "eigenvalue-sum residual"))
        scf.close()

        if forces is not None:
            f.create_dataset("forces", np.asarray(forces, dtype=float) * HARTREE_PER_BOHR_IN_EV_PER_ANG,
                             attrs=_described("eV/Angstrom", "force on each atom, shape (natoms, 3), in the "
                                                             "order of /structure"))
        if stress is not None:
            s = np.asarray(stress, dtype=float)
            attrs = _described("eV/Angstrom^3", "stress tensor, positive under compression: sigma = -(1/V) dE/dT "
                                                "for lattice' = lattice @ T. The opposite sign to ASE's get_stress")
            attrs["convention"] = STRESS_CONVENTION
            attrs["pressure_GPa"] = float(-np.trace(s) / 3 * HARTREE_PER_BOHR3_IN_GPA)
            f.create_dataset("stress", stress_to_output(s), attrs=attrs)
        density = _block(result.density)
        f.create_dataset("density", np.asarray(density, dtype=float) / BOHR_IN_ANG**3, compression="gzip",
                         attrs=_described("electrons/Angstrom^3", "(nspin, *grid_shape) on the real-space grid"))
        if save_wavefunctions:
            for spin, per_spin in enumerate(result.wavefunctions):
                for ik, blk in enumerate(per_spin):
                    f.create_dataset(f"wavefunctions/spin{spin}/kpoint{ik}", _block(blk).astype(np.complex128),
                                     compression="gzip")
                    f.create_dataset(f"wavefunctions/spin{spin}/miller{ik}", np.asarray(system.wave_spheres[ik].miller))
            with f.group("wavefunctions") as g:
                g.attrs["description"] = "coefficients of exp(i(k+G).r)/sqrt(Omega); sum_G |c|^2 = 1"
    return path


def read_scf(path) -> dict:
    """The parts of an SCF result needed to restart or post-process it (``io/results.py:211-240``):
    the density comes back in electrons/Bohr^3, everything else in the file's output units."""
    with h5.open_read(path) as f:
        out = {
            "attrs": f.attrs.to_dict(),
            "eigenvalues": f.read("scf/eigenvalues"),
            "occupations": f.read("scf/occupations"),
            "density": f.read("density") * BOHR_IN_ANG**3,
            "kpoints": f.read("kpoints/points"),
            "weights": f.read("kpoints/weights"),
            "cell": f.read("structure/cell"),
            "positions": f.read("structure/positions"),
            "symbols": [str(s) for s in np.atleast_1d(f.read("structure/symbols"))],
        }
        with f.group("scf/energies") as g:
            out["energies"] = g.attrs.to_dict()
        for name in ("forces", "stress"):
            if name in f:
                out[name] = f.read(name)
        if "wavefunctions" in f:
            nk = len(out["kpoints"])
            nspin = 0
            while f"wavefunctions/spin{nspin}" in f:
                nspin += 1
            out["wavefunctions"] = [[f.read(f"wavefunctions/spin{s}/kpoint{ik}") for ik in range(nk)]
                                    for s in range(nspin)]
    return out


__all__ = ["FORMAT_VERSION", "HARTREE_PER_BOHR3_IN_GPA", "provenance", "read_scf", "stress_to_output", "write_scf"]
