#!/usr/bin/env python3
"""Si64 SCF wall time to 1e-6 Ha/atom on one MI355X (mixed precision + learned start), with
the plain FP64 GPU form and H.psi band-applications per second beside it.

A step is one whole Si64 Gamma-point SCF (2x2x2 conventional Si with seeded displacements,
PBE, ecut 20 Ha, 84^3 grid, 160 bands) from the start density to convergence (1e-6 Ha/atom
and density residual 1e-6) in the mixed-precision form from the learned start; ``value`` is
seconds per SCF (lower is better). The plain FP64 form from the atomic start is timed beside
it, with the energy, eigenvalue and force differences between the two. The second half of
the metric is the H.psi rate of the Si8 kernel batch (42^3, 16 ragged k-points x 64 bands,
inputs resident in HBM) in FP32 and FP64. ``roofline`` covers the dominant kernel of that
H.psi step (the per-plane 2D FFT / V_eff kernel) and ``mfma_roofline`` the projector and
subspace products, each timed with HIP events on the stream it runs on. ``cpu_baseline``
times the FP64 numpy oracle on this host's cores on bounded samples.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--si214]
    python bench.py --hpsi-only      # the k-batch H.psi part alone (kernel profiles)
    python bench.py --pmc-probe      # short program for rocprofv3 --pmc runs
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

METRIC = "Si64 SCF s-to-1e-6 Ha/atom & H·ψ bands/s: mixed+learned vs FP64 GPU"
HBM_PEAK_GBS = 8000.0
#: MI355X FP32 vector peak without packed (dual) FP32 math: 256 CUs x 64 FMA lanes x 2 x 2.4 GHz.
FP32_VALU_PEAK_TFLOPS = 78.6
#: MI355X dense FP32 matrix-core peak (published).
FP32_MFMA_PEAK_TFLOPS = 157.3
REPO = Path(__file__).resolve().parent


def _traffic_from_profile():
    """HBM bytes per launch of the plane kernel, from the committed rocprofv3 PMC summary."""
    path = REPO / "profiles" / "hpsi_kernels_si8_kbatch_fp32_pmc.txt"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch"):
            try:
                return float(line.split("=")[1])
            except (IndexError, ValueError):
                return None
    return None


def build_workload(precision: str):
    import numpy as np
    import torch

    from amaterasu_amd.hamiltonian import Hamiltonian
    from amaterasu_amd.system import build_system, load_dataset
    from amaterasu_amd.workloads import seeded_coefficients, si8_kbatch, smooth_potential

    spec = si8_kbatch()
    system = build_system(spec.cell, spec.positions, spec.symbols, {"Si": load_dataset()}, spec.ecut_rho,
                          spec.kpoints, spec.ecuts)
    v = smooth_potential(system.grid.shape, seed=spec.seed)
    ham = Hamiltonian(system, v, precision=precision)
    plan = ham.plan
    cdt = torch.complex64 if precision == "fp32" else torch.complex128
    psi = torch.zeros((plan.nk, spec.nbands, plan.npw_pad), dtype=cdt, device="cuda")
    for k, s in enumerate(system.wave_spheres):
        psi[k, :, : len(s.g2)] = torch.from_numpy(seeded_coefficients(s.g2, spec.nbands, seed=1000 + k)).to(
            "cuda", cdt)
    return spec, system, v, ham, psi


def time_steps(ham, psi, steps: int, warmup: int) -> float:
    import torch

    out = torch.zeros_like(psi)
    for _ in range(warmup):
        ham.plan.hpsi(psi, ham.veff(0, 0 if psi.dtype == torch.complex64 else 1), out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ham.plan.hpsi(psi, ham.veff(0, 0 if psi.dtype == torch.complex64 else 1), out=out)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def time_kernels(ham, psi, reps: int = 20):
    """Average duration (ms) of each kernel of the step, HIP events on the launch stream."""
    import torch

    from amaterasu_amd import ops
    from amaterasu_amd._lib import check, lib

    plan = ham.plan
    prec = ops.precision_of(psi)
    nkk, nb, _ = psi.shape
    out = torch.zeros_like(psi)
    st = plan.sticks(prec, nkk * nb)
    v = ham.veff(0, prec)
    res = {}
    names = {1: "stick_inverse", 2: "plane_vloc", 3: "stick_forward"}
    def stage(i):
        check(lib().amt_hpsi_stage(plan.handle, prec, i, psi.data_ptr(), out.data_ptr(), nb, 0, nkk, v.data_ptr(),
                                   st.data_ptr(), ops.stream_handle()))

    # The plane kernel works in place on the sticks, so stage 1 refills them (outside the
    # events) before every timed launch.
    for i in (1, 2, 3):
        pairs = []
        for _ in range(reps):
            if i != 1:
                stage(1)
            if i == 3:
                stage(2)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            stage(i)
            e1.record()
            pairs.append((e0, e1))
        torch.cuda.synchronize()
        res[names[i]] = sum(a.elapsed_time(b) for a, b in pairs) / reps
    mat = plan.kb_tables(prec)[0]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        ops.zgemm(psi, mat, 1)
    e1.record()
    e1.synchronize()
    res["kb_project_gemm"] = e0.elapsed_time(e1) / reps
    proj = ops.zgemm(psi, mat, 1)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        ops.zgemm(proj, mat, 0, beta=1.0, out=out)
    e1.record()
    e1.synchronize()
    res["kb_apply_gemm"] = e0.elapsed_time(e1) / reps
    return res


def si64_headline(steps: int, warmup: int, log=None) -> dict:
    """The headline: wall time of the Si64 Gamma-point SCF to 1e-6 Ha/atom (and density
    residual 1e-6) in the mixed-precision form (FP32 FFTs, V_eff and projector / subspace
    GEMMs; FP64 Cholesky, subspace eigensolve, occupations, energies and mixing) from the
    learned start (the atomic start when no trained model is present), and the plain FP64
    form from the atomic superposition beside it. One step is one whole SCF from the start
    density to convergence; the learned start's own cost (features, network, clip and
    rescale) is inside it, building the system is not. ``warmup`` whole SCFs of the timed
    form (and a two-iteration FP64 one) run first, untimed. The FP64 form is timed over
    min(steps, 3) runs. Energy and eigenvalue errors are those of the timed runs. Forces
    respond to the density at first order, so at this tolerance they differ by the SCF
    stopping point (~1e-4 eV/A between any two starts); their precision error is stated from
    a second, untimed pair of runs converged to a density residual of 1e-9, with the mixed
    form finishing in FP64 blocks."""
    import dataclasses

    import numpy as np
    import torch

    from amaterasu_amd.forces import compute_forces
    from amaterasu_amd.learned import MODEL_PATH, DensityModel, learned_density
    from amaterasu_amd.scf import HARTREE_IN_EV, SCFSettings, run_scf
    from amaterasu_amd.workloads import si64_scf_system

    t0 = time.perf_counter()
    system = si64_scf_system()
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    base = SCFSettings(tol_energy=1e-6 * HARTREE_IN_EV, tol_density=1e-6, max_steps=60)
    learned = MODEL_PATH.exists()
    model = DensityModel.load() if learned else None
    mixed = dataclasses.replace(base, precision="fp32")
    fp64 = dataclasses.replace(base, precision="fp64")

    def start(use_learned):
        return learned_density(system, model) if (use_learned and learned) else system.initial_density()

    def one(settings, use_learned):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = run_scf(system, settings, density=start(use_learned))
        torch.cuda.synchronize()
        return time.perf_counter() - t, r

    for i in range(warmup):
        tw, rw = one(mixed, True)
        if log:
            log(f"si64 warm-up {i + 1}/{warmup}: {tw:.3f} s, {rw.iterations} iterations")
    run_scf(system, dataclasses.replace(fp64, max_steps=2), density=start(False))
    torch.cuda.synchronize()
    times, res_mixed = [], None
    t_all = time.perf_counter()
    for i in range(steps):
        tm, res_mixed = one(mixed, True)
        times.append(tm)
        if log:
            log(f"si64 mixed+learned step {i + 1}/{steps}: {tm:.3f} s, {res_mixed.iterations} iterations")
    torch.cuda.synchronize()
    timed_total = time.perf_counter() - t_all
    t64, res64 = [], None
    for i in range(max(1, min(steps, 3))):
        ta, res64 = one(fp64, False)
        t64.append(ta)
        if log:
            log(f"si64 fp64 run {i + 1}: {ta:.3f} s, {res64.iterations} iterations")
    a, b = res64, res_mixed
    nocc = int(round(system.nelectrons / 2))
    # force precision check at tight convergence (untimed)
    tight = dataclasses.replace(base, tol_energy=1e-9 * HARTREE_IN_EV, tol_density=1e-9, max_steps=80)
    fa = run_scf(system, dataclasses.replace(tight, precision="fp64"), density=start(False))
    fb = run_scf(system, dataclasses.replace(tight, precision="mixed"), density=start(True))
    ev_per_ang = HARTREE_IN_EV / 0.5291772105638411
    f64 = compute_forces(system, fa.density, fa.wavefunctions, fa.occupations)
    fmx = compute_forces(system, fb.density, fb.wavefunctions, fb.occupations)
    mean_mixed = timed_total / steps
    mean64 = float(np.mean(t64))
    return {
        "seconds_per_scf": mean_mixed,
        "timed_total_seconds": timed_total,
        "setup_seconds": round(setup, 3),
        "learned_start": learned,
        "mixed_learned": {"seconds": round(mean_mixed, 4), "min_seconds": round(min(times), 4),
                          "max_seconds": round(max(times), 4), "iterations": b.iterations,
                          "davidson_steps": b.davidson_steps, "converged": b.converged, "energy_ha": b.energies.total,
                          "first_density_residual": b.history[0][1]},
        "fp64": {"seconds": round(mean64, 4), "runs": len(t64), "iterations": a.iterations,
                 "davidson_steps": a.davidson_steps, "converged": a.converged, "energy_ha": a.energies.total,
                 "first_density_residual": a.history[0][1]},
        "speedup_vs_fp64": round(mean64 / mean_mixed, 3),
        "errors_vs_fp64": {
            "energy_ha_per_atom": abs(a.energies.total - b.energies.total) / len(system.symbols),
            "occupied_eigenvalue_max_ha": float(np.abs(a.eigenvalues - b.eigenvalues)[..., :nocc].max()),
            "force_max_ev_per_ang_tight_convergence": float(np.abs(fmx - f64).max() * ev_per_ang),
            "largest_force_component_ev_per_ang": float(np.abs(f64).max() * ev_per_ang),  # scale, not an error
            "bounds": "energy 1e-6 Ha/atom, eigenvalues 1e-6 Ha, forces 1e-5 eV/A",
        },
    }


def mfma_roofline(reps: int = 10) -> dict:
    """The complex projector and subspace products of the H.psi path on the matrix cores, timed
    with HIP events on the launch stream (operands random: a GEMM's time does not depend on
    the values). Flops: 8 real per complex multiply-add. Peak: MI355X dense FP32 matrix rate."""
    import torch

    from amaterasu_amd import ops

    shapes = {  # name: (batch, m, n, k, op) for C[m,n] = A[m,k] op(B)
        "kb_project_si8_kbatch": (16, 64, 144, 4954, 1),
        "kb_apply_si8_kbatch": (16, 64, 4954, 144, 0),
        "kb_project_si64": (1, 160, 1152, 37073, 1),
        "kb_apply_si64": (1, 160, 37073, 1152, 0),
        "gram_si64": (1, 160, 160, 37073, 1),
        "overlap_si64_640": (1, 640, 160, 37073, 1),
    }
    gen = torch.Generator(device="cuda").manual_seed(7)
    out = {}
    for name, (batch, m, n, k, op) in shapes.items():
        a = torch.randn((batch, m, k), dtype=torch.complex64, device="cuda", generator=gen)
        b = torch.randn((batch, n, k) if op == 1 else (batch, k, n), dtype=torch.complex64, device="cuda",
                        generator=gen)
        c = ops.zgemm(a, b, op)  # warm-up (and any first-use setup)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            ops.zgemm(a, b, op, out=c)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        flops = 8.0 * batch * m * n * k
        tf = flops / (ms * 1e-3) / 1e12
        out[name] = {"batch": batch, "m": m, "n": n, "k": k, "op": "NC" if op else "NN", "flops": flops,
                     "avg_launch_ms": round(ms, 5), "achieved": round(tf, 2), "frac": round(tf / FP32_MFMA_PEAK_TFLOPS, 4)}
        del a, b, c
    return {"bound": "mfma", "unit": "TFLOP/s", "peak": FP32_MFMA_PEAK_TFLOPS, "dtype": "fp32 (complex64)",
            "convention": "8 real flops per complex multiply-add; HIP events over 10 launches on the launch stream",
            "products": out}


def si214_scf(log=None) -> dict:
    """The largest combined case (opt-in, about ten minutes): Si214 (3x3x3 conventional with two
    seeded vacancies, sigma 0.05 A), PBE, ecut 15 Ha, grid 108^3, ~81k plane waves, the four
    k-points of the time-reversal-reduced shifted 2x2x2 mesh batched in one Davidson, 440 bands,
    Fermi-Dirac smearing 0.01 Ry, SCF to 1e-6 Ha/atom and density residual 1e-6: plain FP64
    from the atomic start against FP32 H.psi/GEMMs (FP64 Cholesky, eigensolve, mixing) from
    the learned start. There is no warm-up (one SCF iteration here is a minute of work): each
    form pays its own first-use costs (GEMM split tuning) inside its timed run. The FP64 planes
    of 108^3 do not fit one workgroup's LDS, so the FP64 form's local H.psi and density take
    the full-grid transforms."""
    import dataclasses

    import numpy as np
    import torch

    from amaterasu_amd.forces import compute_forces
    from amaterasu_amd.learned import MODEL_PATH, DensityModel, learned_density
    from amaterasu_amd.scf import HARTREE_IN_EV, SCFSettings, run_scf
    from amaterasu_amd.workloads import si214_kgrid_system

    t0 = time.perf_counter()
    system = si214_kgrid_system()
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    base = SCFSettings(tol_energy=1e-6 * HARTREE_IN_EV, tol_density=1e-6, max_steps=60, scheme="fermi-dirac",
                       smearing=0.005)
    learned = MODEL_PATH.exists()
    model = DensityModel.load() if learned else None
    runs = {}
    for name, prec, use_learned in (("fp64", "fp64", False), ("mixed_learned", "fp32", True)):
        torch.cuda.synchronize()
        t = time.perf_counter()
        rho = learned_density(system, model) if (use_learned and learned) else system.initial_density()
        r = run_scf(system, dataclasses.replace(base, precision=prec), density=rho,
                    log=(lambda line, name=name: log(f"si214 {name}: {line}")) if log else None)
        torch.cuda.synchronize()
        runs[name] = (time.perf_counter() - t, r, compute_forces(system, r.density, r.wavefunctions, r.occupations))
        if log:
            log(f"si214 scf {name}: {runs[name][0]:.2f} s, {r.iterations} iterations")
    (ta, a, fa), (tb, b, fb) = runs["fp64"], runs["mixed_learned"]
    occupied = a.occupations > 0.5
    ev_per_ang = HARTREE_IN_EV / 0.5291772105638411
    res = {"workload": "Si214 (3x3x3 conventional, 2 seeded vacancies, sigma 0.05 A), PBE, ecut 15 Ha, 108^3, "
                       "shifted 2x2x2 mesh (4 k, batched), 440 bands, Fermi-Dirac 0.01 Ry, tol 1e-6 Ha/atom + density "
                       "residual 1e-6",
           "setup_seconds": round(setup, 1), "learned_start": learned, "speedup": round(ta / tb, 3),
           "free_energy_difference_ha_per_atom": abs(a.energies.free_energy - b.energies.free_energy) / len(system.symbols),
           "occupied_eigenvalue_max_difference_ha": float(np.abs(a.eigenvalues - b.eigenvalues)[occupied].max()),
           "force_max_difference_ev_per_ang": float(np.abs(fa - fb).max() * ev_per_ang)}
    for name, (t, r, _) in runs.items():
        res[name] = {"seconds": round(t, 2), "iterations": r.iterations, "davidson_steps": r.davidson_steps,
                     "converged": r.converged, "free_energy_ha": r.energies.free_energy}
    return res


def pmc_probe() -> None:
    """Three launches of each H.psi kernel on the k-batch workload and one elementwise kernel
    that reads and writes 256 MiB, for the FETCH_SIZE / WRITE_SIZE counter runs (the
    elementwise kernel calibrates the counters)."""
    import torch

    from amaterasu_amd import ops
    from amaterasu_amd._lib import check, lib

    _, _, _, ham, psi = build_workload("fp32")
    plan = ham.plan
    nkk, nb, _ = psi.shape
    out = torch.zeros_like(psi)
    st = plan.sticks(0, nkk * nb)
    v = ham.veff(0, 0)
    for _ in range(3):
        for i in (1, 2, 3):
            check(lib().amt_hpsi_stage(plan.handle, 0, i, psi.data_ptr(), out.data_ptr(), nb, 0, nkk, v.data_ptr(),
                                       st.data_ptr(), ops.stream_handle()))
    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.mul(src, 2.0, out=dst)  # the only float32 scale kernel of the program: reads and writes 256 MiB
    torch.cuda.synchronize()
    print("pmc probe done: 3 x (stick_inverse, plane_vloc, stick_forward) + one 256 MiB -> 256 MiB float32 scale",
          flush=True)


def _cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_hpsi(spec, v, seconds: float = 10.0, workers: int = 1) -> dict:
    """FP64 numpy oracle of Hamiltonian.apply on a bounded sample of the k-batch workload
    (k-points 0..1, all 64 bands), scipy.fft with ``workers`` threads, OpenBLAS as configured."""
    import scipy.fft

    from amaterasu_amd.workloads import seeded_coefficients
    from oracle.hamiltonian import Hamiltonian as OH
    from oracle.pseudo import read_psp8
    from oracle.system import build_system as obuild

    nk_sample = 2
    syso = obuild(spec.cell, spec.positions, spec.symbols,
                  {"Si": read_psp8(REPO / "amaterasu_amd" / "data" / "Si.pbe.psp8.gz")}, float(spec.ecuts[0]),
                  spec.ecut_rho, spec.kpoints[:nk_sample], wave_ecuts=spec.ecuts[:nk_sample], with_ewald=False)
    ho = OH(syso, v[None])
    blocks = [seeded_coefficients(s.g2, spec.nbands, seed=1000 + k) for k, s in enumerate(syso.wave_spheres)]
    with scipy.fft.set_workers(workers):
        ho.apply(blocks[0][:2], 0)  # warm-up
        done, t0 = 0, time.perf_counter()
        while True:
            for k in range(nk_sample):
                ho.apply(blocks[k], k)
                done += spec.nbands
            if time.perf_counter() - t0 > seconds:
                break
        dt = time.perf_counter() - t0
    return {"bands_per_s": done / dt, "band_applications": done, "seconds": round(dt, 2), "fft_workers": workers}


def cpu_si64_davidson_step() -> dict:
    """One Davidson pass of the FP64 numpy oracle on the Si64 Gamma workload (160 bands, npw
    37 073, 84^3, 1152 projectors): orthonormalise the start, H.psi of the block, Rayleigh-Ritz,
    then one pass (residuals, preconditioner, orthogonalisation, H.psi of the corrections,
    Rayleigh-Ritz on 320 rows). Half of it is the cost of one Davidson step."""
    from amaterasu_amd.workloads import si64_gamma, smooth_potential
    from oracle.eigen import davidson, random_bands
    from oracle.hamiltonian import Hamiltonian as OH
    from oracle.pseudo import read_psp8
    from oracle.system import build_system as obuild

    spec = si64_gamma()
    syso = obuild(spec.cell, spec.positions, spec.symbols,
                  {"Si": read_psp8(REPO / "amaterasu_amd" / "data" / "Si.pbe.psp8.gz")}, float(spec.ecuts[0]),
                  spec.ecut_rho, spec.kpoints, with_ewald=False)
    ho = OH(syso, smooth_potential(syso.grid.shape, seed=64)[None])
    sp = syso.wave_spheres[0]
    guess = random_bands(len(sp), spec.nbands, sp.kinetic_energy, seed=12345)
    t0 = time.perf_counter()
    davidson(lambda b: ho.apply(b, 0), guess, sp.kinetic_energy, tolerance=0.0, max_iterations=1)
    dt = time.perf_counter() - t0
    return {"seconds_two_steps": round(dt, 2), "seconds_per_step": round(dt / 2, 2)}


def cpu_config1_scf() -> dict:
    """The whole SCF of the small plain case on the FP64 numpy oracle: diamond Si2 (a = 5.431 A),
    ecut 15 Ha, diamond symmetry (28^3), the 2 irreducible k of the shifted 2x2x2 mesh, 8 bands,
    PBE, fixed occupations, to 1e-8 Ha per cell."""
    import numpy as np

    from oracle.pseudo import read_psp8
    from oracle.scf import SCFSettings as OS
    from oracle.scf import run_scf as orun
    from oracle.symmetry import find_symmetry, ibz_kpoints
    from oracle.system import build_system as obuild
    from oracle.units import ANG_IN_BOHR
    from oracle.xc import XCFunctional

    a = 5.431 * ANG_IN_BOHR
    cell = 0.5 * a * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
    pos = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]]) @ cell
    sym = find_symmetry(cell, pos @ np.linalg.inv(cell), ["Si", "Si"])
    k, w = ibz_kpoints(sym, [2, 2, 2], [0.5, 0.5, 0.5])
    s = obuild(cell, pos, ["Si", "Si"], {"Si": read_psp8(REPO / "amaterasu_amd" / "data" / "Si.pbe.psp8.gz")},
               15.0, 60.0, k, w, symmetry=sym, xc=XCFunctional("PBE"))
    tol = 1e-8 * 27.211386024367243 / 2
    t0 = time.perf_counter()
    r = orun(s, OS(tol_energy=tol, tol_density=1e-9, max_steps=60))
    dt = time.perf_counter() - t0
    return {"seconds": round(dt, 2), "iterations": r.iterations, "converged": r.converged,
            "energy_ha": r.energies.total}


def cpu_baseline(spec, v, fp64_davidson_steps: int, seconds: float = 10.0) -> dict:
    """The numpy FP64 oracle (the reference's algorithm with the same numpy/scipy calls) on this
    host. ``value``: the Si64 SCF wall time it would take, extrapolated from a bounded sample
    (one Davidson pass on the Si64 workload) times the Davidson steps of the plain FP64 GPU SCF
    (the grid-side work between steps is left out, so this is a lower bound). ``legs``: the
    k-batch H.psi rate with scipy.fft on one thread (the reference's default) and on every
    core, and the whole small-case SCF."""
    cores = _cores()
    dav = cpu_si64_davidson_step()
    legs = {"cpu_model": _cpu_model(), "cores": cores,
            "hpsi_si8_kbatch_fft_workers_1": cpu_hpsi(spec, v, seconds, 1),
            f"hpsi_si8_kbatch_fft_workers_{cores}": cpu_hpsi(spec, v, seconds, cores),
            "config1_si2_scf": cpu_config1_scf(),
            "si64_davidson": dav}
    return {
        "value": round(dav["seconds_per_step"] * fp64_davidson_steps, 1),
        "unit": "s",
        "cores": cores,
        "kind": "port",
        "sample": f"one Davidson pass of the FP64 numpy oracle on Si64 (160 bands, 84^3; scipy.fft workers=1, "
                  f"OpenBLAS over {cores} threads): {dav['seconds_per_step']} s per step, times the "
                  f"{fp64_davidson_steps} Davidson steps of the FP64 GPU SCF (lower bound: grid-side work not "
                  f"counted)",
        "legs": legs,
    }


def hpsi_bench(steps: int, warmup: int) -> tuple:
    """H.psi band-applications per second of the k-batch workload in FP32 and FP64, the
    per-kernel times, the plane kernel's HBM and VALU rooflines (secondary half of the metric)."""
    import math

    import numpy as np
    import torch

    from amaterasu_amd.hamiltonian import Hamiltonian

    spec, system, v, ham32, psi32 = build_workload("fp32")
    nkb = psi32.shape[0] * psi32.shape[1]
    t32 = time_steps(ham32, psi32, steps, warmup)
    kern32 = time_kernels(ham32, psi32)
    ham64 = Hamiltonian(system, v, precision="fp64")
    psi64 = psi32.to(torch.complex128)
    t64 = time_steps(ham64, psi64, max(5, steps // 2), 2)
    kern64 = time_kernels(ham64, psi64)
    h64 = ham64.plan.hpsi(psi64, ham64.veff(0, 1))
    h32 = ham32.plan.hpsi(psi32, ham32.veff(0, 0)).to(torch.complex128)
    rel32 = float((h32 - h64).abs().max().item() / h64.abs().max().item())

    # roofline of the dominant kernel of the step: the per-plane 2D FFT + V_eff kernel
    plan = ham32.plan
    n1, n2, n3 = plan.shape
    nst = plan.nsticks()
    nb = spec.nbands
    s_c, s_r = 8, 4
    alg_bytes = sum(nb * S * n1 * 2 * s_c for S in nst) + n1 * n2 * n3 * s_r
    plane_ms = kern32["plane_vloc"]
    achieved = alg_bytes / (plane_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": _traffic_from_profile(),
                "kernel": "k_plane_vloc (fp32)", "algorithmic_bytes_per_launch": alg_bytes,
                "avg_launch_ms": round(plane_ms, 5)}
    # the stick kernels against the same roofline: sticks S n1 s read or written, psi (and H psi) npw s
    npw_sum = sum(plan.npw) * nb
    st_bytes = sum(nb * S * n1 * s_c for S in nst)
    stick_rooflines = {}
    for name, nbytes in (("stick_inverse", st_bytes + npw_sum * s_c),
                         ("stick_forward", st_bytes + 2 * npw_sum * s_c + sum(plan.npw) * s_r)):
        gbs = nbytes / (kern32[name] * 1e-3) / 1e9
        stick_rooflines[name] = {"algorithmic_bytes_per_launch": nbytes, "achieved": round(gbs, 1),
                                 "frac": round(gbs / HBM_PEAK_GBS, 4), "avg_launch_ms": round(kern32[name], 5)}

    def fft_flops(nn):
        return 5.0 * nn * math.log2(nn)

    plane_flops = 0.0
    for sp in system.wave_spheres:
        rows_j = len(np.unique(np.asarray(sp.miller)[:, 1] % n2))
        plane_flops += nb * (n1 * 2 * (rows_j * fft_flops(n3) + n3 * fft_flops(n2)) + 2 * n1 * n2 * n3)
    valu_achieved = plane_flops / (plane_ms * 1e-3) / 1e12
    plane_valu = {"bound": "valu", "achieved": round(valu_achieved, 2), "peak": FP32_VALU_PEAK_TFLOPS,
                  "unit": "TFLOP/s", "frac": round(valu_achieved / FP32_VALU_PEAK_TFLOPS, 4),
                  "flops_per_launch": plane_flops,
                  "convention": "5 N log2 N per length-N complex FFT, k-rows with sticks only, V multiply 2 flops "
                                "per point; peak = FP32 vector FMA rate without packed math"}
    out = {
        "workload": "si8-kbatch: H.psi of Si8 cubic (a=5.431 A), ecut 20 Ha, grid 42^3, 16 seeded k-points "
                    "(npw +-8 %), 64 bands each, nproj 144, V_loc seeded smooth field; inputs resident in HBM",
        "bands_per_step": nkb,
        "fp32": {"bands_per_s": round(nkb / t32, 1), "ms_per_step": round(t32 * 1e3, 4),
                 "kernels_ms": {k: round(x, 5) for k, x in kern32.items()}},
        "fp64": {"bands_per_s": round(nkb / t64, 1), "ms_per_step": round(t64 * 1e3, 4),
                 "kernels_ms": {k: round(x, 5) for k, x in kern64.items()}},
        "fp32_vs_fp64_max_rel_error": rel32,
        "stick_kernel_rooflines_fp32": stick_rooflines,
        "plane_kernel_valu_roofline": plane_valu,
        "npw": plan.npw,
    }
    return spec, v, roofline, out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed Si64 SCF runs (mixed + learned)")
    ap.add_argument("--warmup", type=int, default=2, help="untimed Si64 SCF runs before them")
    ap.add_argument("--hpsi-steps", type=int, default=50, help="timed H.psi steps of the k-batch workload")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--hpsi-only", action="store_true",
                    help="only the k-batch H.psi part (for kernel profiles); prints its own JSON line")
    ap.add_argument("--pmc-probe", action="store_true", help="only the short program for counter runs")
    ap.add_argument("--si214", action="store_true", help="add the Si214 k-grid SCF comparison (about ten minutes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); sharding k-points and bands "
              "across GPUs is not part of this measurement.", file=sys.stderr)
        return 2
    if args.steps < 1:
        print("--steps must be at least 1", file=sys.stderr)
        return 2

    import torch

    if not torch.cuda.is_available():
        print("bench.py needs an MI355X (HIP device)", file=sys.stderr)
        return 1
    if args.pmc_probe:
        pmc_probe()
        return 0
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731

    spec, v, roofline, hpsi = hpsi_bench(args.hpsi_steps, 5)
    log(f"hpsi: fp32 {hpsi['fp32']['bands_per_s']:.0f} bands/s, fp64 {hpsi['fp64']['bands_per_s']:.0f} bands/s")
    if args.hpsi_only:
        print(json.dumps({"metric": "H·ψ bands/s (si8 k-batch, fp32)", "value": hpsi["fp32"]["bands_per_s"],
                          "unit": "bands/s", "roofline": roofline, "hpsi": hpsi}), flush=True)
        return 0
    gemm = mfma_roofline()
    scf = si64_headline(args.steps, args.warmup, log)
    line = {
        "metric": METRIC,
        "value": round(scf["seconds_per_scf"], 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(scf["seconds_per_scf"] * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {
            "workload": "si64-scf: Si64 2x2x2 conventional (a=5.431 A, seeded displacements sigma 0.05 A), PBE, "
                        "ecut 20 Ha, 84^3, Gamma, 160 bands, SCF to 1e-6 Ha/atom + density residual 1e-6 from the "
                        "learned start; one step = one whole SCF",
            "precision_policy": "FP32 Davidson blocks for the whole SCF: FFTs, V_loc, projector and subspace GEMMs "
                                "in FP32 (FP32 sums over at most 1024 products, those sums added in FP64), "
                                "eigensolver tolerance floored at 2e-6 while FP32 blocks are used; FP64 Cholesky, "
                                "subspace eigensolve, occupations, energies, XC, Hartree and mixing",
            "stated_errors_vs_fp64": scf["errors_vs_fp64"],
            "parallelism": "single GPU",
        },
        "si64_scf": scf,
        "hpsi_bands_per_s": {"fp32": hpsi["fp32"]["bands_per_s"], "fp64": hpsi["fp64"]["bands_per_s"]},
        "roofline": roofline,
        "mfma_roofline": gemm,
        "hpsi": hpsi,
    }
    if args.si214:
        line["si214_scf"] = si214_scf(log=log)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(spec, v, scf["fp64"]["davidson_steps"], args.cpu_seconds)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
