"""amt_hpsi_local_noalloc (the local H.psi over the library's fixed scratch array, no device
allocation) against amt_hpsi_local with a caller workspace, through the C ABI: bit-identical
output on two calls in a row, the second with a smaller input, in FP32 and FP64; the
band-slice path (a k-point whose bands do not fit one launch) on the Si64 grid; and the
error for a plan whose single band row needs more sticks than the scratch holds."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plan(shape, spheres):
    from amaterasu_amd.device import DevicePlan

    return DevicePlan(shape, spheres, None)


def _psi(plan, nk, nb, dtype, seed):
    import torch

    rng = np.random.default_rng(seed)
    psi = torch.zeros((nk, nb, plan.npw_pad), dtype=dtype, device="cuda")
    for k in range(nk):
        n = plan.npw[k]
        c = (rng.normal(size=(nb, n)) + 1j * rng.normal(size=(nb, n))) / (1.0 + np.arange(n))[None, :] ** 0.5
        psi[k, :, :n] = torch.from_numpy(c).to(dtype)
    return psi


def _both(plan, psi, v, k0):
    """(amt_hpsi_local with a torch workspace, amt_hpsi_local_noalloc) on the same inputs."""
    import torch

    from amaterasu_amd import ops
    from amaterasu_amd._lib import check, lib

    prec = ops.precision_of(psi)
    nk, nb, _ = psi.shape
    ref = torch.zeros_like(psi)
    ws = torch.empty(lib().amt_plan_stick_bytes(plan.handle, prec, nk * nb), dtype=torch.uint8, device="cuda")
    s = ops.stream_handle()
    check(lib().amt_hpsi_local(plan.handle, prec, psi.data_ptr(), ref.data_ptr(), nb, k0, nk, v.data_ptr(),
                               ws.data_ptr(), s), "hpsi_local")
    got = torch.zeros_like(psi)  # padded columns are written by neither entry point
    check(lib().amt_hpsi_local_noalloc(plan.handle, prec, psi.data_ptr(), got.data_ptr(), nb, k0, nk, v.data_ptr(),
                                       s), "hpsi_local_noalloc")
    torch.cuda.synchronize()
    return ref, got


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_noalloc_equals_workspace_entry_point_kbatch(precision):
    """Si8 k-batch grid (42^3, ragged spheres): 6 k x 64 bands, then 2 k x 5 bands from k0 = 3
    (the second call reuses the scratch the first left behind)."""
    import torch

    from amaterasu_amd.basis import GSphere
    from amaterasu_amd.workloads import si8_kbatch, smooth_potential

    spec = si8_kbatch(nk=6, nbands=64)
    spheres = [GSphere.build(spec.cell, float(e), kpoint=k) for e, k in zip(spec.ecuts, spec.kpoints)]
    plan = _plan((42, 42, 42), spheres)
    dt, rt = (torch.complex64, torch.float32) if precision == "fp32" else (torch.complex128, torch.float64)
    v = torch.from_numpy(smooth_potential((42, 42, 42), seed=5)).to("cuda", rt)
    for nk, nb, k0, seed in ((6, 64, 0, 1), (2, 5, 3, 2)):
        psi = _psi(plan, nk, nb, dt, seed) if k0 == 0 else _psi_at(plan, k0, nk, nb, dt, seed)
        print(f"noalloc kbatch {precision}: nk {nk} nb {nb} k0 {k0}", flush=True)
        ref, got = _both(plan, psi, v, k0)
        assert torch.equal(ref, got)
        assert float(ref.abs().max()) > 0


def _psi_at(plan, k0, nk, nb, dtype, seed):
    import torch

    rng = np.random.default_rng(seed)
    psi = torch.zeros((nk, nb, plan.npw_pad), dtype=dtype, device="cuda")
    for kk in range(nk):
        n = plan.npw[k0 + kk]
        c = rng.normal(size=(nb, n)) + 1j * rng.normal(size=(nb, n))
        psi[kk, :, :n] = torch.from_numpy(c).to(dtype)
    return psi


def test_noalloc_band_slices_on_the_si64_grid():
    """Si64 Gamma (84^3, npw 37 073) in FP64: more bands than one launch of the scratch holds,
    so one k-point goes in band slices; then a call of 3 bands."""
    import torch

    from amaterasu_amd._lib import FP64, lib
    from amaterasu_amd.basis import GSphere
    from amaterasu_amd.workloads import si64_gamma, smooth_potential

    spec = si64_gamma()
    plan = _plan((84, 84, 84), [GSphere.build(spec.cell, 20.0)])
    row = lib().amt_plan_stick_bytes(plan.handle, FP64, 1)
    per_launch = (64 << 20) // row
    nb = per_launch + 7
    assert 1 <= per_launch < nb
    v = torch.from_numpy(smooth_potential((84, 84, 84), seed=6)).to("cuda", torch.float64)
    for n, seed in ((nb, 3), (3, 4)):
        print(f"noalloc si64 fp64: nb {n} (rows per launch {per_launch})", flush=True)
        ref, got = _both(plan, _psi(plan, 1, n, torch.complex128, seed), v, 0)
        assert torch.equal(ref, got)


def test_noalloc_rejects_a_row_beyond_the_scratch():
    """A long grid (960 x 96 x 96) whose 'sphere' has a stick at each of 95 x 95 (j, k) columns
    of the plane: one FP32 band row needs 960 * 9025 * 8 bytes of sticks, more than the 64 MiB
    scratch. The call fails with the library's error before any launch."""
    import types

    import torch

    from amaterasu_amd._lib import AmtError, FP32, check, lib

    j, k = np.meshgrid(np.arange(-47, 48), np.arange(-47, 48), indexing="ij")
    miller = np.stack([np.zeros(j.size, dtype=int), j.ravel(), k.ravel()], axis=1)
    sphere = types.SimpleNamespace(miller=miller, g2=np.arange(len(miller), dtype=float))
    plan = _plan((960, 96, 96), [sphere])
    assert lib().amt_plan_stick_bytes(plan.handle, FP32, 1) > (64 << 20)
    psi = torch.zeros((1, 1, plan.npw_pad), dtype=torch.complex64, device="cuda")
    out = torch.zeros_like(psi)
    v = torch.zeros((960, 96, 96), dtype=torch.float32, device="cuda")
    with pytest.raises(AmtError, match="more than the fixed scratch"):
        check(lib().amt_hpsi_local_noalloc(plan.handle, FP32, psi.data_ptr(), out.data_ptr(), 1, 0, 1,
                                           v.data_ptr(), ctypes.c_void_p(0)), "hpsi_local_noalloc")
