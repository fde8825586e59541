"""The MFMA complex GEMM (NN and NC, FP64 and FP32, split-K, batched) against numpy."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _c(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-13), ("complex64", 2e-5)])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 33, 9), (64, 64, 16), (100, 70, 300), (8, 36, 754),
                                   (160, 1152, 3000), (37, 5, 40000), (300, 270, 1000), (257, 513, 700),
                                   (256, 384, 3000), (150, 90, 700), (470, 33, 5000), (150, 20000, 1100)])
def test_zgemm_nn_and_nc(dtype, tol, m, n, k):
    """Every tile shape: 64- and 128-row tiles, 80-row tiles with ragged edges (m = 150, 470),
    and an 80-row candidate whose single K chunk would need the two-level kernel (m = 150,
    k = 1100 with 626 output tiles: it falls back to 64-row tiles)."""
    import torch

    from amaterasu_amd import ops

    a, b_nn, b_nc = _c((m, k), 1), _c((k, n), 2), _c((n, k), 3)
    ta = torch.from_numpy(a).to("cuda", getattr(torch, dtype))
    nn = ops.zgemm(ta, torch.from_numpy(b_nn).to("cuda", ta.dtype), 0).cpu().numpy()
    nc = ops.zgemm(ta, torch.from_numpy(b_nc).to("cuda", ta.dtype), 1).cpu().numpy()
    ref_nn, ref_nc = a @ b_nn, a @ b_nc.conj().T
    scale = np.sqrt(k) * 2.0
    assert np.abs(nn - ref_nn).max() / scale < tol
    assert np.abs(nc - ref_nc).max() / scale < tol


def test_zgemm_alpha_beta_and_batch():
    import torch

    from amaterasu_amd import ops

    a, b, c = _c((3, 20, 50), 4), _c((3, 50, 30), 5), _c((3, 20, 30), 6)
    tc = torch.from_numpy(c).cuda()
    ops.zgemm(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), 0, alpha=-1.0, beta=1.0, out=tc)
    assert np.abs(tc.cpu().numpy() - (c - a @ b)).max() < 1e-12
    bc = _c((3, 30, 50), 7)
    out = ops.zgemm(torch.from_numpy(a).cuda(), torch.from_numpy(bc).cuda(), 1).cpu().numpy()
    assert np.abs(out - a @ bc.conj().transpose(0, 2, 1)).max() < 1e-12


def test_mfma_lane_layout_single_tile():
    """A 16x16x4 product with distinct small integers: exact in both precisions, so any
    lane-layout mistake (transposed fragment, wrong row of the accumulator) shows up as
    a wrong integer."""
    import torch

    from amaterasu_amd import ops

    a = (np.arange(16 * 4).reshape(16, 4) % 7 + 1j * (np.arange(64).reshape(16, 4) % 5)).astype(np.complex128)
    b = (np.arange(4 * 16).reshape(4, 16) % 3 - 1j * (np.arange(64).reshape(4, 16) % 4)).astype(np.complex128)
    for dt in (torch.complex128, torch.complex64):
        got = ops.zgemm(torch.from_numpy(a).to("cuda", dt), torch.from_numpy(b).to("cuda", dt), 0).cpu().numpy()
        assert np.array_equal(got.astype(np.complex128), a @ b)


def test_zgemm_strided_views():
    """Row-strided and batch-strided views (a prefix of a larger subspace block) go to the
    kernel without copies and give the same products."""
    import torch

    from amaterasu_amd import ops

    g = torch.Generator().manual_seed(5)
    big = torch.randn((3, 40, 130), dtype=torch.complex128, generator=g).cuda()
    a = torch.randn((3, 7, 130), dtype=torch.complex128, generator=g).cuda()
    view = big[:, :25]  # batch stride 40*130, rows of 130
    c = ops.zgemm(a, view, 1)
    want = a.cpu() @ view.cpu().conj().transpose(-1, -2)
    assert torch.allclose(c.cpu(), want, rtol=1e-12, atol=1e-11)
    rot = torch.randn((3, 5, 25), dtype=torch.complex128, generator=g).cuda()
    out = torch.zeros((3, 9, 130), dtype=torch.complex128).cuda()
    ops.zgemm(rot, view, 0, out=out[:, 2:7])
    assert torch.allclose(out[:, 2:7].cpu(), rot.cpu() @ view.cpu(), rtol=1e-12, atol=1e-11)
    assert float(out[:, :2].abs().max()) == 0.0 and float(out[:, 7:].abs().max()) == 0.0


@pytest.mark.parametrize("split", [1, 0])
def test_zgemm_fp32_long_contraction_accumulates_in_fp64(split):
    """Gram matrix of unit rows over K = 131072 plane waves in FP32 against the FP64 product
    of the same (FP32-rounded) inputs. With one split (the whole contraction in one
    workgroup) the FP32 matrix-core sums are carried into FP64 every 64 of K; with the
    heuristic split (256 chunks of 512) each chunk keeps one FP32 accumulator and the
    partials are summed in FP64. Either way the error stays near the final rounding
    (~1e-7), where one FP32 accumulator over all of K loses ~1e-4 on the unit diagonal."""
    import torch

    from amaterasu_amd import ops
    from amaterasu_amd._lib import FP32, check, lib

    m, k = 128, 131072
    g = torch.Generator().manual_seed(9)
    a = torch.randn((m, k), dtype=torch.complex128, generator=g)
    a = (a / torch.linalg.vector_norm(a, dim=1, keepdim=True)).to(torch.complex64).cuda()
    a64 = a.to(torch.complex128)
    check(lib().amt_zgemm_set_split(FP32, 1, m, m, k, 1, split), "zgemm_set_split")
    try:
        got = ops.zgemm(a, a, 1).to(torch.complex128)
    finally:
        check(lib().amt_zgemm_set_split(FP32, 1, m, m, k, 1, 0), "zgemm_set_split")
    want = ops.zgemm(a64, a64, 1)
    assert float((torch.diagonal(want) - 1.0).abs().max()) < 1e-6
    assert float((got - want).abs().max()) < 3e-7


def test_zgemm_fp32_large_tiles_long_contraction():
    """FP32 NC products with both output sides >= 256 take 128 x 128 tiles with K chunks of
    512 and the FP64 split reduction: the same K = 131 072 unit-row Gram check (batched, so
    the batch offsets of the partial products are exercised too), and an NN rotation of the
    same rows."""
    import torch

    from amaterasu_amd import ops

    nb, m, k = 2, 256, 131072
    g = torch.Generator().manual_seed(10)
    a = torch.randn((nb, m, k), dtype=torch.complex128, generator=g)
    a = (a / torch.linalg.vector_norm(a, dim=-1, keepdim=True)).to(torch.complex64).cuda()
    a64 = a.to(torch.complex128)
    got = ops.zgemm(a, a, 1).to(torch.complex128)
    want = ops.zgemm(a64, a64, 1)
    assert float((got - want).abs().max()) < 3e-7
    rot = torch.randn((nb, 300, m), dtype=torch.complex128, generator=g).cuda() / m**0.5
    got = ops.zgemm(rot.to(torch.complex64), a, 0).to(torch.complex128)
    want = ops.zgemm(rot.to(torch.complex64).to(torch.complex128), a64, 0)
    # single-level FP32 over K = 256 in Gauss's three-product form: the imaginary part's
    # rounding is bounded by eps * sum |a_re + a_im| |b_re + b_im|, up to twice the bound of
    # four real products (measured 1.1e-6 here; the four-product form passed 1e-6)
    assert float((got - want).abs().max() / want.abs().max()) < 2e-6
