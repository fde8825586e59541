"""Seeded synthetic inputs for the H.psi workloads (plain numpy, reproducible).

* ``si8_kbatch``: the kernel-level H.psi batch -- conventional 8-atom Si cell
  (a = 5.431 A), ecut 20 Ha (40 Ry) on the grid the reference's rule gives (42^3),
  16 seeded k-points whose plane-wave counts are spread by +-8 % through a per-k cutoff
  scale (1 + d_k)^(2/3), d_k ~ U[-0.08, 0.08], 64 bands per k with coefficients
  CN(0, 1) / (1 + |k+G|^2)^2, and V_loc(r) a seeded normal field smoothed by a periodic
  Gaussian of sigma = 2 grid cells, scaled to 1 Ha RMS.
* ``si64_gamma``: 2x2x2 conventional supercell (64 atoms) with seeded Gaussian
  displacements (sigma = 0.05 A), ecut 20 Ha, Gamma only (complex), 160 bands, 84^3.
* ``si64_scf_system``: the same cell as a device system ready for the SCF loop (PBE,
  norm-conserving Si, local potential / core / starting density superposed on the device).

Everything here is numpy; device objects are built from it by the callers.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

#: CODATA 2014 Bohr radius in Angstrom, derived as ase.units does (amaterasu/units.py:70).
BOHR_IN_ANG = 0.5291772105638411
ANG_IN_BOHR = 1.0 / BOHR_IN_ANG

DIAMOND_BASIS = np.array([[0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0],
                          [0.25, 0.25, 0.25], [0.25, 0.75, 0.75], [0.75, 0.25, 0.75], [0.75, 0.75, 0.25]])


def si_conventional(reps: int = 1, a_ang: float = 5.431, sigma_ang: float = 0.0, seed: int = 0):
    """Cell (rows, Bohr) and Cartesian positions (Bohr) of a reps^3 diamond supercell."""
    a = a_ang * ANG_IN_BOHR
    cell = reps * a * np.eye(3)
    shifts = np.stack(np.meshgrid(*[np.arange(reps)] * 3, indexing="ij"), axis=-1).reshape(-1, 3)
    frac = (DIAMOND_BASIS[None, :, :] + shifts[:, None, :]).reshape(-1, 3) / reps
    pos = frac @ cell
    if sigma_ang > 0:
        rng = np.random.default_rng(seed)
        pos = pos + rng.normal(scale=sigma_ang * ANG_IN_BOHR, size=pos.shape)
    return cell, pos


def smooth_potential(shape, seed: int, sigma_cells: float = 2.0, rms: float = 1.0) -> np.ndarray:
    """Seeded normal field, periodic Gaussian filter of sigma grid cells, scaled to ``rms`` Ha."""
    rng = np.random.default_rng(seed)
    field = rng.normal(size=shape)
    f = [np.fft.fftfreq(n) for n in shape]  # cycles per cell
    m2 = f[0][:, None, None] ** 2 + f[1][None, :, None] ** 2 + f[2][None, None, :] ** 2
    filt = np.exp(-2.0 * np.pi**2 * sigma_cells**2 * m2)
    v = np.fft.ifftn(np.fft.fftn(field) * filt).real
    v -= v.mean()
    v *= rms / np.sqrt(np.mean(v**2))
    return np.ascontiguousarray(v)


def seeded_coefficients(g2: np.ndarray, nbands: int, seed: int) -> np.ndarray:
    """CN(0,1) / (1 + |k+G|^2)^2, shape (nbands, npw)."""
    rng = np.random.default_rng(seed)
    c = (rng.normal(size=(nbands, len(g2))) + 1j * rng.normal(size=(nbands, len(g2)))) / np.sqrt(2.0)
    return c / (1.0 + g2[None, :]) ** 2


@dataclass
class KBatchSpec:
    cell: np.ndarray
    positions: np.ndarray
    symbols: list
    kpoints: np.ndarray
    ecuts: np.ndarray  # per-k wavefunction cutoff (Ha)
    ecut_rho: float
    nbands: int
    seed: int


def si8_kbatch(nk: int = 16, nbands: int = 64, ecut: float = 20.0, spread: float = 0.08, seed: int = 2024):
    cell, pos = si_conventional(1)
    rng = np.random.default_rng(seed)
    kpts = rng.uniform(-0.5, 0.5, size=(nk, 3))
    d = rng.uniform(-spread, spread, size=nk)
    ecuts = ecut * (1.0 + d) ** (2.0 / 3.0)
    return KBatchSpec(cell=cell, positions=pos, symbols=["Si"] * 8, kpoints=kpts, ecuts=ecuts,
                      ecut_rho=4.0 * ecut, nbands=nbands, seed=seed)


def si64_gamma(nbands: int = 160, ecut: float = 20.0, seed: int = 64):
    cell, pos = si_conventional(2, sigma_ang=0.05, seed=seed)
    return KBatchSpec(cell=cell, positions=pos, symbols=["Si"] * 64, kpoints=np.zeros((1, 3)),
                      ecuts=np.array([ecut]), ecut_rho=4.0 * ecut, nbands=nbands, seed=seed)


def si64_scf_system(nbands: int = 160, ecut: float = 20.0, seed: int = 64, xc: str = "PBE"):
    """Device system for the Si64 Gamma-point SCF (needs the GPU)."""
    from .system import attach_ionic_fields, build_system, load_dataset

    spec = si64_gamma(nbands=nbands, ecut=ecut, seed=seed)
    ds = load_dataset()
    system = build_system(spec.cell, spec.positions, spec.symbols, {"Si": ds}, spec.ecut_rho, spec.kpoints,
                          spec.ecuts, nbands=spec.nbands)
    return attach_ionic_fields(system, {"Si": ds}, spec.ecut_rho, xc=xc)


#: The 2x2x2 Monkhorst-Pack mesh shifted by half a step, reduced by time reversal (k ~ -k):
#: four k-points of weight 1/4.
KMESH_SHIFTED_222 = np.array([[0.25, 0.25, 0.25], [0.25, 0.25, -0.25], [0.25, -0.25, 0.25], [-0.25, 0.25, 0.25]])


def si214_cell(seed: int = 214, sigma_ang: float = 0.05):
    """3x3x3 conventional Si (a = 5.431 A, 216 sites) with 2 seeded vacancies and seeded
    Gaussian displacements of the remaining 214 atoms: (cell, positions) in Bohr."""
    cell, pos = si_conventional(3, sigma_ang=sigma_ang, seed=seed)
    rng = np.random.default_rng(seed + 1)
    drop = rng.choice(len(pos), size=2, replace=False)
    return cell, np.delete(pos, drop, axis=0)


def si214_kgrid_system(nbands: int = 440, ecut: float = 15.0, seed: int = 214, xc: str = "PBE"):
    """Device system for the largest combined case: Si214 (two vacancies), PBE, ecut 15 Ha
    (30 Ry, grid 108^3, ~81k plane waves per k), the time-reversal-reduced shifted 2x2x2 mesh
    (4 k-points, batched), 440 bands (needs the GPU)."""
    from .system import attach_ionic_fields, build_system, load_dataset

    cell, pos = si214_cell(seed)
    ds = load_dataset()
    system = build_system(cell, pos, ["Si"] * len(pos), {"Si": ds}, 4.0 * ecut, KMESH_SHIFTED_222, ecut,
                          weights=np.full(4, 0.25), nbands=nbands)
    return attach_ionic_fields(system, {"Si": ds}, 4.0 * ecut, xc=xc)


__all__ = ["KBatchSpec", "KMESH_SHIFTED_222", "si214_cell", "si214_kgrid_system", "si64_scf_system", "si8_kbatch", "si64_gamma", "si_conventional", "smooth_potential", "seeded_coefficients"]


def main(argv=None) -> int:
    """``python -m amaterasu_amd.workloads {si64-scf,si214-kgrid-scf} [--precision P] [--learned]``:
    one SCF to 1e-6 Ha/atom and density residual 1e-6, printing each iteration (the wall time
    counts from the start density on; building the system is not in it)."""
    import argparse
    import time

    ap = argparse.ArgumentParser(prog="python -m amaterasu_amd.workloads")
    ap.add_argument("workload", choices=["si64-scf", "si214-kgrid-scf"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "mixed"])
    ap.add_argument("--learned", action="store_true", help="start from the trained network's density")
    ap.add_argument("--max-steps", type=int, default=60)
    ap.add_argument("--fp32-tol-floor", type=float, default=None, help="eigensolver tolerance floor of FP32 blocks")
    ap.add_argument("--fp64-below", type=float, default=None, help="mixed: density residual of the switch to FP64")
    ap.add_argument("--runs", type=int, default=1, help="SCFs one after another in this process (each timed)")
    a = ap.parse_args(argv)
    import torch

    from .learned import DensityModel, learned_density
    from .scf import HARTREE_IN_EV, SCFSettings, run_scf

    if a.workload == "si64-scf":
        system, smearing = si64_scf_system(), {}
    else:  # Fermi-Dirac smearing of 0.01 Ry
        system, smearing = si214_kgrid_system(), {"scheme": "fermi-dirac", "smearing": 0.005}
    t_setup = time.perf_counter()
    rho = learned_density(system, DensityModel.load()) if a.learned else system.initial_density()
    extra = {}
    if a.fp32_tol_floor is not None:
        extra["fp32_tol_floor"] = a.fp32_tol_floor
    if a.fp64_below is not None:
        extra["fp64_below"] = a.fp64_below
    settings = SCFSettings(tol_energy=1e-6 * HARTREE_IN_EV, tol_density=1e-6, max_steps=a.max_steps,
                           precision=a.precision, **smearing, **extra)
    torch.cuda.synchronize()
    t = t_setup
    print(f"system ready: {len(system.symbols)} atoms, grid {system.grid.shape}, "
          f"npw {[len(sp.g2) for sp in system.wave_spheres]}, {system.nbands} bands", flush=True)
    ok = True
    for run in range(max(a.runs, 1)):
        if run:
            torch.cuda.synchronize()
            t = time.perf_counter()
        r = run_scf(system, settings, density=rho,
                    log=lambda line: print(f"{time.perf_counter() - t:8.3f}s {line}", flush=True))
        torch.cuda.synchronize()
        print(f"converged {r.converged} after {r.iterations} iterations, E {r.energies.total:.10f} Ha, "
              f"{time.perf_counter() - t:.3f} s", flush=True)
        if settings.davidson_graphs:
            from .eigen import graph_stats

            print(f"Davidson graphs: {graph_stats()}", flush=True)
        ok = ok and r.converged
    return 0 if ok else 1


if __name__ == "__main__":
    raise SystemExit(main())
