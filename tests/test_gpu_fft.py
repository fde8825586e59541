"""The batched mixed-radix 3D FFT and the sphere<->grid maps against numpy / the oracle.

Convention (basis/fft.py:1-14): to_real = ifftn with norm="forward" (no 1/N),
to_reciprocal = fftn with norm="forward" (1/N). FP64 must agree to ~1e-13 relative,
FP32 to ~1e-6 relative."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

def _smooth(n):
    for p in (2, 3, 5, 7, 11):
        while n % p == 0:
            n //= p
    return n == 1


LENGTHS = [n for n in range(2, 129) if _smooth(n)]


def _rand(shape, seed, dtype=np.complex128):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(dtype)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_every_smooth_length_along_every_axis(axis):
    import torch

    from amaterasu_amd import ops

    for n in LENGTHS:
        shape = [3, 4, 5]
        shape[axis] = n
        x = _rand((2, *shape), seed=n)
        y = ops.fft3d(torch.from_numpy(x).cuda(), +1).cpu().numpy()
        ref = np.fft.ifftn(x, axes=(-3, -2, -1), norm="forward")
        err = np.abs(y - ref).max() / np.abs(ref).max()
        assert err < 1e-13, (n, axis, err)


@pytest.mark.parametrize("shape", [(28, 28, 28), (42, 42, 42), (35, 35, 35), (44, 22, 33), (84, 84, 84),
                                   (108, 108, 108), (11, 121, 7)])
def test_grids_of_the_workloads_fp64_and_fp32(shape):
    import torch

    from amaterasu_amd import ops

    x = _rand(shape, seed=sum(shape))
    ref_r = np.fft.ifftn(x, norm="forward")
    ref_g = np.fft.fftn(x, norm="forward")
    t = torch.from_numpy(x).cuda()
    n = np.prod(shape)
    r64 = ops.fft3d(t, +1).cpu().numpy()
    g64 = ops.fft3d(t, -1, 1.0 / n).cpu().numpy()
    assert np.abs(r64 - ref_r).max() / np.abs(ref_r).max() < 1e-13
    assert np.abs(g64 - ref_g).max() / np.abs(ref_g).max() < 1e-13
    r32 = ops.fft3d(t.to(torch.complex64), +1).cpu().numpy()
    assert np.abs(r32 - ref_r).max() / np.abs(ref_r).max() < 2e-6


def test_fftgrid_drop_in_matches_the_oracle(si2_system):
    """FFTGrid.to_real/to_reciprocal/scatter/gather (basis/fft.py:192-243) on the GPU."""
    from amaterasu_amd.basis import FFTGrid

    o = si2_system.grid
    g = FFTGrid(o.shape, o.cell)
    sphere = si2_system.wave_spheres[1]
    rng = np.random.default_rng(5)
    c = rng.normal(size=(3, len(sphere))) + 1j * rng.normal(size=(3, len(sphere)))
    grid = g.scatter(c, sphere)
    assert np.array_equal(grid, o.scatter(c, sphere))  # bit-exact placement
    assert np.array_equal(g.gather(grid, sphere), c)
    u = g.wave_to_real(c, sphere)
    ref = o.wave_to_real(c, sphere)
    assert np.abs(u - ref).max() / np.abs(ref).max() < 1e-13
    back = g.real_to_wave(u, sphere)
    assert np.abs(back - c).max() < 1e-13  # test_fft.py:49-55 round trip
    with pytest.raises(ValueError, match="alias"):
        FFTGrid((8, 8, 8), o.cell).scatter(np.ones(len(si2_system.density_sphere)), si2_system.density_sphere)


def test_fftgrid_gradient_divergence_and_constructors_match_the_oracle(si2_system):
    """FFTGrid.gradient/divergence/gvectors_cartesian/for_sphere (basis/fft.py:164-304) on
    the GPU against the oracle's, numpy in and out and device in and out."""
    import torch

    from amaterasu_amd.basis import FFTGrid
    from oracle.basis import FFTGrid as OFFTGrid

    o = si2_system.grid
    g = FFTGrid(o.shape, o.cell)
    assert str(g).startswith(f"{o.shape[0]}x{o.shape[1]}x{o.shape[2]} = {g.size} points")
    sphere = si2_system.density_sphere
    assert FFTGrid.for_sphere(sphere).shape == OFFTGrid.for_sphere(sphere).shape
    assert np.array_equal(g.gvectors_cartesian(zero_nyquist=True), o.gvectors_cartesian(zero_nyquist=True))
    rng = np.random.default_rng(8)
    u = rng.normal(size=o.shape)
    ref = o.gradient(u)
    got = g.gradient(u)
    assert got.shape == ref.shape and np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    f = rng.normal(size=(*o.shape, 3))
    ref = o.divergence(f)
    got = g.divergence(f)
    assert got.shape == ref.shape and np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    dev = g.gradient(torch.from_numpy(u).cuda())
    assert isinstance(dev, torch.Tensor) and dev.is_cuda
    assert np.abs(dev.cpu().numpy() - o.gradient(u)).max() < 1e-12 * np.abs(o.gradient(u)).max()
