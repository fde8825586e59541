"""H.psi on the GPU against the oracle's Hamiltonian.apply (hamiltonian/operator.py:60-69).

FP64 must match to ~1e-12 relative (summation order differs in the FFTs and GEMMs);
FP32 is reported against FP64 with a stated bound."""

from __future__ import annotations

import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _block(n, nb, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(nb, n)) + 1j * rng.normal(size=(nb, n))


def test_apply_matches_oracle_si2(si2_system):
    from amaterasu_amd.hamiltonian import Hamiltonian
    from oracle.hamiltonian import Hamiltonian as OH

    rng = np.random.default_rng(0)
    v = rng.normal(size=(1, *si2_system.grid.shape))
    h, o = Hamiltonian(si2_system, v), OH(si2_system, v)
    for ik in range(2):
        c = _block(len(si2_system.wave_spheres[ik]), 5, ik)
        got, ref = h.apply(c, ik), o.apply(c, ik)
        assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12
        one = h.apply(c[2], ik)  # a single band (1-D) keeps its shape
        assert one.shape == c[2].shape and np.abs(one - ref[2]).max() / np.abs(ref).max() < 1e-12


def test_apply_generic_grid_matches_oracle(si2_system):
    """A non-cubic grid runs the generic (runtime-radix) kernels instead of the
    compile-time-length ones; both must agree with the oracle."""
    from amaterasu_amd.hamiltonian import Hamiltonian
    from oracle.basis import FFTGrid
    from oracle.hamiltonian import Hamiltonian as OH

    s = copy.copy(si2_system)
    s.grid = FFTGrid((24, 20, 22), si2_system.cell)
    rng = np.random.default_rng(11)
    v = rng.normal(size=(1, *s.grid.shape))
    h, o = Hamiltonian(s, v), OH(s, v)
    assert not h.plan.info[0] == h.plan.info[1] == h.plan.info[2]
    for ik in range(2):
        c = _block(len(s.wave_spheres[ik]), 3, 20 + ik)
        got, ref = h.apply(c, ik), o.apply(c, ik)
        assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12
        got32 = Hamiltonian(s, v, precision="fp32").apply(c, ik)
        assert np.abs(got32 - ref).max() / np.abs(ref).max() < 5e-6


def test_apply_is_hermitian(si2_system):
    from amaterasu_amd.hamiltonian import Hamiltonian

    rng = np.random.default_rng(1)
    h = Hamiltonian(si2_system, rng.normal(size=(1, *si2_system.grid.shape)))
    b = _block(len(si2_system.wave_spheres[1]), 4, 2)
    m = b.conj() @ h.apply(b, 1).T
    assert np.abs(m - m.conj().T).max() < 1e-11


def test_wrong_potential_shape(si2_system):
    from amaterasu_amd.hamiltonian import Hamiltonian

    with pytest.raises(ValueError, match="effective potential has shape"):
        Hamiltonian(si2_system, np.zeros((2, *si2_system.grid.shape)))


def test_fp32_apply_error_is_bounded(si2_system):
    from amaterasu_amd.hamiltonian import Hamiltonian
    from oracle.hamiltonian import Hamiltonian as OH

    rng = np.random.default_rng(3)
    v = rng.normal(size=(1, *si2_system.grid.shape))
    c = _block(len(si2_system.wave_spheres[0]), 6, 4)
    ref = OH(si2_system, v).apply(c, 0)
    got = Hamiltonian(si2_system, v, precision="fp32").apply(c, 0)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 5e-6


def test_kbatch_matches_per_k_oracle(si_pseudo):
    """The ragged k-batch (padded [k][band][npw_pad] layout, one launch per kernel)
    against the oracle applied k by k, on 3 k-points of the Si8 kernel batch."""
    import torch

    from amaterasu_amd.hamiltonian import Hamiltonian
    from amaterasu_amd.system import build_system, load_dataset
    from amaterasu_amd.workloads import seeded_coefficients, si8_kbatch, smooth_potential
    from oracle.hamiltonian import Hamiltonian as OH
    from oracle.system import build_system as obuild

    spec = si8_kbatch(nk=3, nbands=6)
    sysd = build_system(spec.cell, spec.positions, spec.symbols, {"Si": load_dataset()}, spec.ecut_rho,
                        spec.kpoints, spec.ecuts)
    syso = obuild(spec.cell, spec.positions, spec.symbols, {"Si": si_pseudo}, float(spec.ecuts[0]), spec.ecut_rho,
                  spec.kpoints, wave_ecuts=spec.ecuts, with_ewald=False)
    assert sysd.grid.shape == syso.grid.shape == (42, 42, 42)
    v = smooth_potential(sysd.grid.shape, seed=spec.seed)
    hd = Hamiltonian(sysd, v)
    ho = OH(syso, v)
    plan = hd.plan
    psi = torch.zeros((3, 6, plan.npw_pad), dtype=torch.complex128, device="cuda")
    blocks = []
    for k, s in enumerate(sysd.wave_spheres):
        c = seeded_coefficients(s.g2, 6, seed=100 + k)
        blocks.append(c)
        psi[k, :, : len(s.g2)] = torch.from_numpy(c).cuda()
    out = hd.apply_batched(psi).cpu().numpy()
    for k, s in enumerate(sysd.wave_spheres):
        ref = ho.apply(blocks[k], k)
        got = out[k, :, : len(s.g2)]
        assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12
        assert np.all(out[k, :, len(s.g2):] == 0)  # padding stays zero


def test_kbatch_fp32_band_pairs_across_k_points(si_pseudo):
    """FP32 k-batch on the 42^3 grid with an odd band count (5 bands at each of 3 k-points,
    so neighbouring band rows of the launch belong to different k-points, with different
    stick maps and row lists). Against the oracle k by k, at the FP32 bound of the
    single-k test."""
    import torch

    from amaterasu_amd.hamiltonian import Hamiltonian
    from amaterasu_amd.system import build_system, load_dataset
    from amaterasu_amd.workloads import seeded_coefficients, si8_kbatch, smooth_potential
    from oracle.hamiltonian import Hamiltonian as OH
    from oracle.system import build_system as obuild

    spec = si8_kbatch(nk=3, nbands=5)
    sysd = build_system(spec.cell, spec.positions, spec.symbols, {"Si": load_dataset()}, spec.ecut_rho,
                        spec.kpoints, spec.ecuts)
    syso = obuild(spec.cell, spec.positions, spec.symbols, {"Si": si_pseudo}, float(spec.ecuts[0]), spec.ecut_rho,
                  spec.kpoints, wave_ecuts=spec.ecuts, with_ewald=False)
    v = smooth_potential(sysd.grid.shape, seed=spec.seed)
    hd = Hamiltonian(sysd, v, precision="fp32")
    ho = OH(syso, v)
    plan = hd.plan
    psi = torch.zeros((3, 5, plan.npw_pad), dtype=torch.complex64, device="cuda")
    blocks = []
    for k, s in enumerate(sysd.wave_spheres):
        c = seeded_coefficients(s.g2, 5, seed=200 + k)
        blocks.append(c)
        psi[k, :, : len(s.g2)] = torch.from_numpy(c).to("cuda", torch.complex64)
    out = hd.apply_batched(psi).cpu().numpy()
    for k, s in enumerate(sysd.wave_spheres):
        ref = ho.apply(blocks[k], k)
        got = out[k, :, : len(s.g2)]
        assert np.abs(got - ref).max() / np.abs(ref).max() < 5e-6
        assert np.all(out[k, :, len(s.g2):] == 0)


def test_empty_lattice_free_electron_bands(si2_system):
    """Gate of test_hamiltonian.py:221-244 on the GPU path: V = 0, no projectors ->
    eigenvalues |k+G|^2/2 to 1e-11 Ha."""
    from amaterasu_amd.eigen import davidson, random_bands
    from amaterasu_amd.hamiltonian import Hamiltonian
    from oracle.pseudo import KBProjectors

    s = copy.copy(si2_system)
    s.projectors = [KBProjectors(np.zeros((0, len(sp)), dtype=complex), np.zeros((0, 0)), []) for sp in
                    s.wave_spheres]
    h = Hamiltonian(s, np.zeros((1, *s.grid.shape)))
    for ik, sp in enumerate(s.wave_spheres):
        r = davidson(lambda b, ik=ik: h.apply(b, ik), random_bands(len(sp), 10, sp.kinetic_energy, seed=ik),
                     sp.kinetic_energy, tolerance=1e-12, max_iterations=200)
        assert np.abs(r.eigenvalues - np.sort(sp.kinetic_energy)[:10]).max() < 1e-11


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_full_grid_route_matches_the_plane_kernels(si2_system, precision):
    """Grids whose (n2 x n3) plane does not fit one workgroup's LDS (FP64 planes above ~100^2,
    e.g. the 108^3 grid of Si214) take the full-grid route: the library's scatter, 3D
    transforms and gather. Forced here on the Si2 grid, it must give the pruned kernels' H.psi
    and density on the same inputs."""
    import torch

    from amaterasu_amd.device import DevicePlan
    from amaterasu_amd.ops import FP32, FP64

    s = si2_system
    plan = DevicePlan(s.grid.shape, s.wave_spheres, s.projectors)
    prec, dt = (FP64, torch.complex128) if precision == "fp64" else (FP32, torch.complex64)
    rng = np.random.default_rng(7)
    psi = torch.zeros((2, 6, plan.npw_pad), dtype=dt, device="cuda")
    for k in range(2):
        psi[k, :, : plan.npw[k]] = torch.from_numpy(_block(plan.npw[k], 6, 10 + k)).to(dt)
    v = torch.from_numpy(rng.normal(size=s.grid.shape)).cuda()
    rows, wts = np.arange(12), rng.uniform(0.1, 1.0, size=12)
    rho_a = torch.zeros(s.grid.shape, dtype=torch.float64, device="cuda")
    rho_b = torch.zeros_like(rho_a)
    ref = plan.hpsi(psi, v)
    plan.accumulate_density(psi, rows, wts, rho_a)
    plan.set_route(prec, True)
    assert not plan.plane_ok[prec]
    got = plan.hpsi(psi, v)
    plan.accumulate_density(psi, rows, wts, rho_b)
    tol = 1e-12 if precision == "fp64" else 2e-6
    assert (torch.abs(got - ref).max() / torch.abs(ref).max()).item() < tol
    assert (torch.abs(rho_b - rho_a).max() / rho_a.abs().max()).item() < tol


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_full_operator_entry_point_matches_the_call_sequence(si_pseudo, precision):
    """amt_hpsi (T + V_loc + V_NL in one call, projector tables attached to the plan with
    amt_plan_set_projectors) against the explicit sequence amt_hpsi_local, amt_zgemm (NC),
    amt_kb_dij, amt_zgemm (NN) through the C ABI on 3 ragged k-points of the Si8 batch:
    bit-identical, and a plan without tables gives the local part alone."""
    import torch

    from amaterasu_amd import ops
    from amaterasu_amd._lib import check, lib
    from amaterasu_amd.device import DevicePlan
    from amaterasu_amd.system import build_system, load_dataset
    from amaterasu_amd.workloads import seeded_coefficients, si8_kbatch, smooth_potential

    spec = si8_kbatch(nk=3, nbands=7)
    s = build_system(spec.cell, spec.positions, spec.symbols, {"Si": load_dataset()}, spec.ecut_rho, spec.kpoints,
                     spec.ecuts)
    plan = DevicePlan(s.grid.shape, s.wave_spheres, s.projectors)
    prec = 1 if precision == "fp64" else 0
    cdt = torch.complex128 if prec else torch.complex64
    psi = torch.zeros((3, 7, plan.npw_pad), dtype=cdt, device="cuda")
    for k, sp in enumerate(s.wave_spheres):
        psi[k, :, : len(sp.g2)] = torch.from_numpy(seeded_coefficients(sp.g2, 7, seed=40 + k)).to("cuda", cdt)
    v = torch.from_numpy(smooth_potential(s.grid.shape, seed=4)).to("cuda", ops.real_dtype(prec))
    mat, dij, start, length = plan.kb_tables(prec)
    stream = ops.stream_handle()
    # explicit sequence
    ref = torch.zeros_like(psi)
    st = plan.sticks(prec, 21)
    check(lib().amt_hpsi_local(plan.handle, prec, psi.data_ptr(), ref.data_ptr(), 7, 0, 3, v.data_ptr(),
                               st.data_ptr(), stream), "hpsi_local")
    local = ref.clone()
    proj = ops.zgemm(psi, mat, 1)
    pd = torch.empty_like(proj)
    check(lib().amt_kb_dij(prec, proj.data_ptr(), dij.data_ptr(), start.data_ptr(), length.data_ptr(), pd.data_ptr(),
                           7, plan.nproj, 7 * plan.nproj, plan.nproj * plan.nproj, 7 * plan.nproj, 3, stream), "kb_dij")
    ops.zgemm(pd, mat, 0, beta=1.0, out=ref)
    # one call
    check(lib().amt_plan_set_projectors(plan.handle, prec, plan.nproj, mat.data_ptr(), dij.data_ptr(),
                                        start.data_ptr(), length.data_ptr()), "plan_set_projectors")
    nbytes = lib().amt_hpsi_workspace(plan.handle, prec, 7, 3)
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    got = torch.zeros_like(psi)
    check(lib().amt_hpsi(plan.handle, prec, psi.data_ptr(), got.data_ptr(), 7, 0, 3, v.data_ptr(), ws.data_ptr(),
                         nbytes, stream), "hpsi")
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
    # a single k-point from the middle of the batch
    one = torch.zeros_like(psi[1:2])
    check(lib().amt_hpsi(plan.handle, prec, psi[1:2].data_ptr(), one.data_ptr(), 7, 1, 1, v.data_ptr(),
                         ws.data_ptr(), nbytes, stream), "hpsi")
    torch.cuda.synchronize()
    assert (torch.abs(one - ref[1:2]).max() / torch.abs(ref).max()).item() < (1e-14 if prec else 1e-6)
    # detached tables: the local part alone
    check(lib().amt_plan_set_projectors(plan.handle, prec, 0, None, None, None, None), "plan_set_projectors")
    bare = torch.zeros_like(psi)
    check(lib().amt_hpsi(plan.handle, prec, psi.data_ptr(), bare.data_ptr(), 7, 0, 3, v.data_ptr(), ws.data_ptr(),
                         nbytes, stream), "hpsi")
    torch.cuda.synchronize()
    assert torch.equal(bare, local)
    with pytest.raises(Exception, match="workspace smaller"):
        check(lib().amt_hpsi(plan.handle, prec, psi.data_ptr(), bare.data_ptr(), 7, 0, 3, v.data_ptr(),
                             ws.data_ptr(), 16, stream), "hpsi")


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_launch_shapes_give_identical_results(si_pseudo, precision):
    """Every launch shape the autotuner may pick (plane kernel of 256 or 512 threads, stick
    kernels over the default or half the sticks per workgroup) gives bit-identical H.psi on
    the Si8 k-batch (42^3); the hook rejects other shapes."""
    import torch

    from amaterasu_amd import ops
    from amaterasu_amd._lib import AmtError, check, lib
    from amaterasu_amd.device import DevicePlan
    from amaterasu_amd.system import build_system, load_dataset
    from amaterasu_amd.workloads import seeded_coefficients, si8_kbatch, smooth_potential

    spec = si8_kbatch(nk=3, nbands=9)
    s = build_system(spec.cell, spec.positions, spec.symbols, {"Si": load_dataset()}, spec.ecut_rho, spec.kpoints,
                     spec.ecuts)
    plan = DevicePlan(s.grid.shape, s.wave_spheres, s.projectors)
    prec = 1 if precision == "fp64" else 0
    cdt = torch.complex128 if prec else torch.complex64
    psi = torch.zeros((3, 9, plan.npw_pad), dtype=cdt, device="cuda")
    for k, sp in enumerate(s.wave_spheres):
        psi[k, :, : len(sp.g2)] = torch.from_numpy(seeded_coefficients(sp.g2, 9, seed=60 + k)).to("cuda", cdt)
    v = torch.from_numpy(smooth_potential(s.grid.shape, seed=6)).to("cuda", ops.real_dtype(prec))
    st = plan.sticks(prec, 27)
    wdef = 32 if prec else 64
    outs = []
    for nth in (256, 512):
        for w in (wdef, wdef // 2):
            check(lib().amt_plan_set_launch(plan.handle, prec, nth, w), "plan_set_launch")
            out = torch.zeros_like(psi)
            check(lib().amt_hpsi_local(plan.handle, prec, psi.data_ptr(), out.data_ptr(), 9, 0, 3, v.data_ptr(),
                                       st.data_ptr(), ops.stream_handle()), "hpsi_local")
            outs.append(out)
    torch.cuda.synchronize()
    for o in outs[1:]:
        assert torch.equal(o, outs[0])
    with pytest.raises(AmtError, match="plane_threads must be"):
        check(lib().amt_plan_set_launch(plan.handle, prec, 384, 0), "plan_set_launch")
