#!/usr/bin/env python3
"""H.psi throughput of the batched k-point workload on one MI355X.

A step is one batched application of the Kohn-Sham Hamiltonian (kinetic + local
V_eff through the pruned FFTs + Kleinman-Bylander projectors) to every band of every
k-point of the Si8 kernel-batch workload: conventional 8-atom Si cell, ecut 20 Ha on a
42^3 grid, 16 seeded k-points with plane-wave counts spread +-8 %, 64 bands each
(1024 band-applications per step), inputs resident in HBM. ``value`` is
band-applications per second in the mixed-precision (FP32) form; the FP64 form is
reported alongside. The ``roofline`` object covers the dominant kernel (the per-plane
2D FFT / V_eff kernel), timed with HIP events on the stream it runs on; the
``cpu_baseline`` object times the FP64 numpy oracle of the same operator on a bounded
sample of the same workload on this host. The ``si64_scf`` object is the other half of
the metric: wall time of the Si64 Gamma-point SCF to 1e-6 Ha/atom in plain FP64 and in
mixed precision from the learned start, with the energy and eigenvalue differences.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-scf] [--no-cpu-baseline]
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
REPO = Path(__file__).resolve().parent


def _traffic_from_profile():
    """HBM bytes per launch of the plane kernel, from the committed rocprofv3 PMC summary."""
    path = REPO / "profiles" / "plane_vloc_si8_kbatch_fp32_pmc.txt"
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


def si64_scf(log=None) -> dict:
    """Wall time of the Si64 Gamma-point SCF to 1e-6 Ha/atom (and density residual 1e-6):
    the plain FP64 form from the atomic superposition against the mixed-precision form (FP32
    FFTs, V_eff and projector / subspace GEMMs; FP64 Cholesky, subspace eigensolve,
    occupations, energies and mixing) from the learned start (the atomic start when no
    trained model is present). The learned start's own cost (features, network, clip and
    rescale) is inside its timed region; the system setup and an untimed two-iteration
    warm-up of both forms are not. Energy and eigenvalue errors are those of the timed
    runs. Forces respond to the density at first order, so at this tolerance they differ by
    the SCF stopping point (~1e-4 eV/A between any two starts); their precision error is
    stated from a second, untimed pair of runs converged to a density residual of 1e-9, with
    the mixed form finishing in FP64 blocks."""
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

    def start(use_learned):
        return learned_density(system, model) if (use_learned and learned) else system.initial_density()

    for prec in ("fp64", "fp32"):  # untimed warm-up
        run_scf(system, dataclasses.replace(base, precision=prec, max_steps=2), density=start(True))
    torch.cuda.synchronize()
    runs = {}
    for name, prec, use_learned in (("fp64", "fp64", False), ("mixed_learned" if learned else "mixed", "fp32", True)):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = run_scf(system, dataclasses.replace(base, precision=prec), density=start(use_learned))
        torch.cuda.synchronize()
        runs[name] = (time.perf_counter() - t, r)
        if log:
            log(f"si64 scf {name}: {runs[name][0]:.3f} s, {r.iterations} iterations")
    (ta, a), (tb, b) = runs["fp64"], runs["mixed_learned" if learned else "mixed"]
    nocc = int(round(system.nelectrons / 2))
    # force precision check at tight convergence (untimed)
    tight = dataclasses.replace(base, tol_energy=1e-9 * HARTREE_IN_EV, tol_density=1e-9, max_steps=80)
    fa = run_scf(system, dataclasses.replace(tight, precision="fp64"), density=start(False))
    fb = run_scf(system, dataclasses.replace(tight, precision="mixed"), density=start(True))
    ev_per_ang = HARTREE_IN_EV / 0.5291772105638411
    f64 = compute_forces(system, fa.density, fa.wavefunctions, fa.occupations)
    fmx = compute_forces(system, fb.density, fb.wavefunctions, fb.occupations)
    res = {"workload": "Si64 2x2x2 conventional (a=5.431 A, seeded displacements sigma 0.05 A), PBE, ecut 20 Ha, "
                       "84^3, Gamma, 160 bands, tol 1e-6 Ha/atom + density residual 1e-6",
           "setup_seconds": round(setup, 3), "learned_start": learned,
           "speedup": round(ta / tb, 3),
           "energy_difference_ha_per_atom": abs(a.energies.total - b.energies.total) / len(system.symbols),
           "occupied_eigenvalue_max_difference_ha": float(np.abs(a.eigenvalues - b.eigenvalues)[..., :nocc].max()),
           "force_max_difference_ev_per_ang_tight": float(np.abs(fmx - f64).max() * ev_per_ang),
           "force_max_ev_per_ang": float(np.abs(f64).max() * ev_per_ang)}
    for name, (t, r) in runs.items():
        res[name] = {"seconds": round(t, 3), "iterations": r.iterations, "davidson_steps": r.davidson_steps,
                     "converged": r.converged, "energy_ha": r.energies.total,
                     "first_density_residual": r.history[0][1]}
    return res


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
    """Three launches of the plane kernel on the k-batch workload and one elementwise kernel
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
        for i in (1, 2):
            check(lib().amt_hpsi_stage(plan.handle, 0, i, psi.data_ptr(), out.data_ptr(), nb, 0, nkk, v.data_ptr(),
                                       st.data_ptr(), ops.stream_handle()))
    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.mul(src, 2.0, out=dst)  # the only float32 scale kernel of the program: reads and writes 256 MiB
    torch.cuda.synchronize()
    print("pmc probe done: 3 x (stick_inverse, plane_vloc) + one 256 MiB -> 256 MiB float32 scale", flush=True)


def cpu_baseline(spec, v, seconds: float = 15.0):
    """FP64 numpy oracle of Hamiltonian.apply on a bounded sample of the same workload."""
    import numpy as np

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
    ho.apply(blocks[0][:2], 0)  # warm-up
    done, t0 = 0, time.perf_counter()
    while True:
        for k in range(nk_sample):
            ho.apply(blocks[k], k)
            done += spec.nbands
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        cores = os.cpu_count()
    return {
        "value": done / dt,
        "unit": "bands/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{done} band-applications of the FP64 numpy oracle (scipy.fft workers=1, OpenBLAS over "
                  f"{cores} threads) on k-points 0..{nk_sample - 1} of the same workload, {dt:.1f} s",
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-scf", action="store_true", help="skip the Si64 SCF wall-time comparison")
    ap.add_argument("--pmc-probe", action="store_true", help="only the short program for counter runs")
    ap.add_argument("--si214", action="store_true", help="add the Si214 k-grid SCF comparison (about ten minutes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); sharding k-points and bands "
              "across GPUs is not part of this measurement.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs an MI355X (HIP device)", file=sys.stderr)
        return 1
    if args.pmc_probe:
        pmc_probe()
        return 0

    spec, system, v, ham32, psi32 = build_workload("fp32")
    nkb = psi32.shape[0] * psi32.shape[1]
    t32 = time_steps(ham32, psi32, args.steps, args.warmup)
    kern32 = time_kernels(ham32, psi32)

    # FP64 form of the same step, and the FP32 error against it
    from amaterasu_amd.hamiltonian import Hamiltonian

    ham64 = Hamiltonian(system, v, precision="fp64")
    psi64 = psi32.to(torch.complex128)
    t64 = time_steps(ham64, psi64, max(5, args.steps // 2), 2)
    h64 = ham64.plan.hpsi(psi64, ham64.veff(0, 1))
    h32 = ham32.plan.hpsi(psi32, ham32.veff(0, 0)).to(torch.complex128)
    rel32 = float((h32 - h64).abs().max().item() / h64.abs().max().item())

    # roofline of the dominant kernel: the per-plane 2D FFT + V_eff kernel
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

    # the same kernel against the FP32 vector-ALU roofline (what bounds it: the pruned
    # decomposition keeps its HBM traffic small). Flops by the usual 5 N log2 N per length-N
    # complex FFT; along k only the rows j that hold sticks are transformed; V multiply 2 flops
    # per point; peak = FP32 FMA without packed math (disabled on this pool): 78.6 TFLOP/s.
    import math

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

    line = {
        "metric": METRIC,
        "value": round(nkb / t32, 1),
        "unit": "bands/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t32 * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "si8-kbatch: H.psi of Si8 cubic (a=5.431 A), ecut 20 Ha, grid 42^3, 16 seeded k-points "
                        "(npw +-8 %), 64 bands each, nproj 144, V_loc seeded smooth field",
            "bands_per_step": nkb,
            "grid": list(plan.shape),
            "npw": plan.npw,
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "plane_kernel_valu_roofline": plane_valu,
        "fp64": {"value": round(nkb / t64, 1), "unit": "bands/s", "ms_per_step": round(t64 * 1e3, 4)},
        "fp32_vs_fp64_max_rel_error": rel32,
        "kernels_ms_fp32": {k: round(x, 5) for k, x in kern32.items()},
    }
    if not args.no_scf:
        line["si64_scf"] = si64_scf(log=lambda m: print(m, file=sys.stderr, flush=True))
    if args.si214:
        line["si214_scf"] = si214_scf(log=lambda m: print(m, file=sys.stderr, flush=True))
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(spec, v, args.cpu_seconds)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
