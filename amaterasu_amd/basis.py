"""Plane-wave spheres, the grid rule and an FFT grid whose transforms run on the GPU.

``GSphere.build`` follows ``basis/gvectors.py:119-163``. The coefficient layout of every
band is the sphere order, so the order must be bit-identical to the reference's: the
same float64 operations produce ``g2`` (a matmul of the Miller indices with the
reciprocal cell plus k, then a row-wise dot) and the same key sorts it (g2, then
h1, h2, h3). The grid rule is ``basis/fft.py:57-81`` (2,3,5,7,11-smooth lengths).

``FFTGrid`` keeps the reference's method surface (``basis/fft.py:118-308``). Its
transforms, scatter and gather run in libamaterasu_amd; numpy arrays in give numpy
arrays out, HIP tensors stay on the device.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.fft
import torch

from . import ops


def reciprocal_cell(cell) -> np.ndarray:
    """Rows b_i with b_i . a_j = 2 pi delta_ij (``gvectors.py:36-47``)."""
    cell = np.asarray(cell, dtype=float)
    if cell.shape != (3, 3):
        raise ValueError(f"cell must be 3x3, got {cell.shape}")
    vol = float(np.linalg.det(cell))
    if abs(vol) < 1e-12:
        raise ValueError(f"cell is singular (volume {vol:.3e} Bohr^3)")
    return 2.0 * np.pi * np.linalg.inv(cell).T


@dataclass
class GSphere:
    """Plane waves with |k+G|^2/2 <= ecut in the reference order."""

    miller: np.ndarray
    cartesian: np.ndarray
    g2: np.ndarray
    kpoint: np.ndarray
    ecut: float
    cell: np.ndarray

    def __len__(self) -> int:
        return len(self.miller)

    @property
    def kinetic_energy(self) -> np.ndarray:
        return 0.5 * self.g2

    @classmethod
    def build(cls, cell, ecut: float, kpoint=(0.0, 0.0, 0.0), *, sort: bool = True) -> "GSphere":
        if ecut <= 0:
            raise ValueError(f"ecut must be positive, got {ecut}")
        cell = np.asarray(cell, dtype=float)
        recip = reciprocal_cell(cell)
        k = np.asarray(kpoint, dtype=float)
        kc = k @ recip
        radius = np.sqrt(2.0 * ecut) + float(np.linalg.norm(kc))
        box = np.ceil(radius * np.linalg.norm(cell, axis=1) / (2.0 * np.pi)).astype(int)
        h = np.stack(np.meshgrid(*[np.arange(-b, b + 1) for b in box], indexing="ij"), axis=-1).reshape(-1, 3)
        cart = h @ recip + kc
        g2 = np.einsum("ij,ij->i", cart, cart)
        inside = g2 <= 2.0 * ecut + 1e-12
        h, cart, g2 = h[inside], cart[inside], g2[inside]
        if sort:
            order = np.lexsort((h[:, 2], h[:, 1], h[:, 0], g2))
            h, cart, g2 = h[order], cart[order], g2[order]
        return cls(miller=h, cartesian=cart, g2=g2, kpoint=k, ecut=float(ecut), cell=cell)


def fast_length(minimum: int, multiple: int = 1) -> int:
    """Smallest 2,3,5,7,11-smooth length >= minimum, divisible by ``multiple``."""
    if multiple <= 1:
        return int(scipy.fft.next_fast_len(int(minimum)))
    n = -(-int(minimum) // multiple) * multiple
    for _ in range(1000):
        if scipy.fft.next_fast_len(n) == n:
            return n
        n += multiple
    return int(scipy.fft.next_fast_len(int(minimum)))


def grid_shape_for_cutoff(cell, ecut_rho: float, multiple: int = 1) -> tuple:
    """``FFTGrid.for_cutoff`` (``fft.py:151-172``): n >= 2 h_max + 1 of the density sphere."""
    s = GSphere.build(cell, ecut_rho, sort=False)
    hmax = np.abs(s.miller).max(axis=0)
    return tuple(fast_length(int(2 * h + 1), multiple) for h in hmax)


def _host(x) -> bool:
    return not isinstance(x, torch.Tensor)


@dataclass
class FFTGrid:
    """A real-space grid with GPU transforms (``fft.py:118-308``)."""

    shape: tuple
    cell: np.ndarray
    backend: str = "hip"
    _plans: dict = field(default_factory=dict, repr=False, compare=False)

    @property
    def size(self) -> int:
        return int(np.prod(self.shape))

    @property
    def volume(self) -> float:
        return abs(float(np.linalg.det(self.cell)))

    @property
    def cell_volume_per_point(self) -> float:
        return self.volume / self.size

    def __str__(self) -> str:
        n1, n2, n3 = self.shape
        return f"{n1}x{n2}x{n3} = {self.size} points ({self.backend} backend)"

    @classmethod
    def for_cutoff(cls, cell, ecut: float, *, backend: str = "hip", multiple: int = 1) -> "FFTGrid":
        """``fft.py:151-162``: smallest efficient grid holding the ecut sphere."""
        return cls(shape=grid_shape_for_cutoff(cell, ecut, multiple), cell=np.asarray(cell, dtype=float),
                   backend=backend)

    @classmethod
    def for_sphere(cls, sphere, *, backend: str = "hip", multiple: int = 1) -> "FFTGrid":
        """``fft.py:164-172``: smallest efficient grid holding ``sphere`` without aliasing."""
        hmax = np.abs(sphere.miller).max(axis=0)
        return cls(shape=tuple(fast_length(int(2 * h + 1), multiple) for h in hmax),
                   cell=np.asarray(sphere.cell, dtype=float), backend=backend)

    def indices(self, sphere):
        """``fft.py:174-190``."""
        hmax = np.abs(sphere.miller).max(axis=0)
        if np.any(2 * hmax + 1 > np.asarray(self.shape)):
            raise ValueError(
                f"grid {self.shape} is too small for Miller indices up to {tuple(int(h) for h in hmax)}; "
                "high-G components would alias onto low-G ones"
            )
        return tuple(sphere.miller[:, a] % self.shape[a] for a in range(3))

    def _plan(self, sphere):
        from .device import DevicePlan

        key = id(sphere)
        plan = self._plans.get(key)
        if plan is None or plan[0] is not sphere:
            self.indices(sphere)  # the reference's aliasing error, raised on the host
            plan = (sphere, DevicePlan(self.shape, [sphere]))
            self._plans[key] = plan
        return plan[1]

    def _transform(self, array, sign: int, scale: float):
        host = _host(array)
        x = torch.from_numpy(np.ascontiguousarray(np.asarray(array), dtype=np.complex128)) if host else array
        x = x.to("cuda")
        if not x.is_complex():
            x = x.to(torch.complex128)
        y = ops.fft3d(x.contiguous(), sign, scale)
        return y.cpu().numpy() if host else y

    def to_real(self, grid_reciprocal):
        """``fft.py:219-226``: u(r) = sum_G c_G exp(i G.r), no 1/N."""
        return self._transform(grid_reciprocal, +1, 1.0)

    def to_reciprocal(self, grid_real):
        """``fft.py:228-235``: c_G = (1/N) sum_r u(r) exp(-i G.r)."""
        return self._transform(grid_real, -1, 1.0 / self.size)

    def scatter(self, coefficients, sphere):
        """``fft.py:192-206``."""
        plan = self._plan(sphere)
        host = _host(coefficients)
        c = torch.from_numpy(np.ascontiguousarray(np.asarray(coefficients), dtype=np.complex128)) if host \
            else coefficients
        c = c.to("cuda")
        one = c.dim() == 1
        c2 = c[None, :] if one else c
        out = plan.scatter(c2.contiguous(), 0)
        out = out[0] if one else out
        return out.cpu().numpy() if host else out

    def gather(self, grid, sphere):
        """``fft.py:208-212``."""
        plan = self._plan(sphere)
        host = _host(grid)
        g = torch.from_numpy(np.ascontiguousarray(np.asarray(grid), dtype=np.complex128)) if host else grid
        g = g.to("cuda")
        if not g.is_complex():
            g = g.to(torch.complex128)
        one = g.dim() == 3
        out = plan.gather((g[None] if one else g).contiguous(), 0)
        out = out[0] if one else out
        return out.cpu().numpy() if host else out

    def wave_to_real(self, coefficients, sphere):
        return self.to_real(self.scatter(coefficients, sphere))

    def real_to_wave(self, values, sphere):
        return self.gather(self.to_reciprocal(values), sphere)

    def points_fractional(self) -> np.ndarray:
        axes = [np.arange(n) / n for n in self.shape]
        return np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1)

    def points_cartesian(self) -> np.ndarray:
        return self.points_fractional() @ self.cell

    def gvectors_cartesian(self, *, zero_nyquist: bool = False) -> np.ndarray:
        """``fft.py:254-286``: Cartesian G of every grid point (FFT frequency order), 1/Bohr,
        shape ``(*shape, 3)``; ``zero_nyquist`` blanks the n/2 component of even axes."""
        axes = []
        for n in self.shape:
            f = np.fft.fftfreq(n, d=1.0 / n)
            if zero_nyquist and n % 2 == 0:
                f[n // 2] = 0.0
            axes.append(f)
        miller = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1)
        return miller @ reciprocal_cell(self.cell)

    def _ig(self) -> torch.Tensor:
        """i G (Nyquist blanked) on the device, shape (3, *shape), complex128 (cached)."""
        ig = self._plans.get("_ig")
        if ig is None:
            g = torch.from_numpy(np.ascontiguousarray(np.moveaxis(self.gvectors_cartesian(zero_nyquist=True), -1, 0)))
            ig = (1j * g).to("cuda", torch.complex128)
            self._plans["_ig"] = ig
        return ig

    def gradient(self, values):
        """``fft.py:288-295``: gradient of a real periodic field via i G, shape ``(*shape, 3)``."""
        host = _host(values)
        v = torch.as_tensor(np.asarray(values) if host else values).to("cuda", torch.complex128)
        c = ops.fft3d(v.reshape(self.shape).contiguous(), -1, 1.0 / self.size)
        d = ops.fft3d((self._ig() * c[None]).contiguous(), +1, 1.0, inplace=True).real
        out = d.permute(1, 2, 3, 0).contiguous()
        return out.cpu().numpy() if host else out

    def divergence(self, field):
        """``fft.py:297-304``: divergence of a real vector field of shape ``(*shape, 3)``."""
        host = _host(field)
        f = torch.as_tensor(np.asarray(field) if host else field).to("cuda", torch.complex128)
        c = ops.fft3d(f.permute(3, 0, 1, 2).contiguous(), -1, 1.0 / self.size, inplace=True)
        acc = (self._ig() * c).sum(0).contiguous()
        out = ops.fft3d(acc, +1, 1.0, inplace=True).real.contiguous()
        return out.cpu().numpy() if host else out

    def integrate(self, values):
        """``fft.py:306-308``."""
        if isinstance(values, torch.Tensor):
            return torch.sum(values, dim=(-3, -2, -1)) * self.cell_volume_per_point
        return np.sum(values, axis=(-3, -2, -1)) * self.cell_volume_per_point


__all__ = ["FFTGrid", "GSphere", "fast_length", "grid_shape_for_cutoff", "reciprocal_cell"]
