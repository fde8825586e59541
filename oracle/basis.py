"""Plane-wave spheres, the FFT grid rule, sphere<->grid maps and the transform pair.

Restates ``amaterasu/basis/gvectors.py`` and ``amaterasu/basis/fft.py``.
The G ordering has to be bit-identical to the reference (it defines the layout of
every coefficient array), so the arithmetic that produces ``g2`` -- a float64
matmul of the Miller indices with the reciprocal cell, plus ``k``, then a row-wise
dot -- is the same sequence of numpy operations, and the sort key is the same
(g2 primary, then h1, h2, h3).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.fft


def reciprocal_cell(cell) -> np.ndarray:
    """Rows b_i with b_i.a_j = 2 pi delta_ij (``gvectors.py:36-47``)."""
    cell = np.asarray(cell, dtype=float)
    if cell.shape != (3, 3):
        raise ValueError(f"cell must be 3x3, got {cell.shape}")
    det = float(np.linalg.det(cell))
    if abs(det) < 1e-12:
        raise ValueError(f"cell is singular (volume {det:.3e} Bohr^3)")
    return 2.0 * np.pi * np.linalg.inv(cell).T


def miller_bounds(cell, gmax: float) -> np.ndarray:
    """|h_i| <= gmax |a_i| / 2pi, rounded up (``gvectors.py:50-58``)."""
    return np.ceil(gmax * np.linalg.norm(np.asarray(cell, dtype=float), axis=1) / (2.0 * np.pi)).astype(int)


def estimate_count(volume: float, ecut: float) -> float:
    """Asymptotic plane-wave count Omega (2 ecut)^(3/2) / (6 pi^2) (``gvectors.py:61-72``)."""
    return volume * (2.0 * ecut) ** 1.5 / (6.0 * np.pi**2)


@dataclass
class GSphere:
    """All |k+G|^2/2 <= ecut, sorted by (g2, h1, h2, h3) (``gvectors.py:75-163``)."""

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

    @property
    def gmax(self) -> float:
        return float(np.sqrt(2.0 * self.ecut))

    @classmethod
    def build(cls, cell, ecut: float, kpoint=(0.0, 0.0, 0.0), *, sort: bool = True) -> "GSphere":
        # gvectors.py:133-163
        if ecut <= 0:
            raise ValueError(f"ecut must be positive, got {ecut}")
        cell = np.asarray(cell, dtype=float)
        recip = reciprocal_cell(cell)
        kpoint = np.asarray(kpoint, dtype=float)
        kcart = kpoint @ recip
        gmax = np.sqrt(2.0 * ecut)
        box = miller_bounds(cell, gmax + float(np.linalg.norm(kcart)))
        axes = [np.arange(-b, b + 1) for b in box]
        miller = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)
        cart = miller @ recip + kcart
        g2 = np.einsum("ij,ij->i", cart, cart)
        keep = g2 <= 2.0 * ecut + 1e-12
        miller, cart, g2 = miller[keep], cart[keep], g2[keep]
        if sort:
            order = np.lexsort((miller[:, 2], miller[:, 1], miller[:, 0], g2))
            miller, cart, g2 = miller[order], cart[order], g2[order]
        return cls(miller=miller, cartesian=cart, g2=g2, kpoint=kpoint, ecut=float(ecut), cell=cell)


def fast_length(minimum: int, multiple: int = 1) -> int:
    """Smallest 2,3,5,7,11-smooth length >= minimum divisible by multiple (``fft.py:72-81``).

    Uses scipy's ``next_fast_len`` exactly as the reference does.
    """
    if multiple <= 1:
        return int(scipy.fft.next_fast_len(int(minimum)))
    candidate = -(-int(minimum) // multiple) * multiple
    for _ in range(1000):
        if scipy.fft.next_fast_len(candidate) == candidate:
            return candidate
        candidate += multiple
    return int(scipy.fft.next_fast_len(int(minimum)))


def fast_grid_shape(min_shape, multiple: int = 1) -> tuple[int, int, int]:
    """``fft.py:57-69``."""
    return tuple(fast_length(int(n), multiple) for n in min_shape)


def symmetry_compatible_multiple(translations, maximum: int = 12) -> int:
    """Smallest m with m*t integral for every fractional translation (``fft.py:84-115``)."""
    t = np.asarray(translations, dtype=float).ravel()
    t = t - np.floor(t)
    if t.size == 0 or np.all(np.abs(t) < 1e-8):
        return 1
    for m in range(1, maximum + 1):
        s = t * m
        if np.all(np.abs(s - np.rint(s)) < 1e-6):
            return m
    return 1


@dataclass
class FFTGrid:
    """Uniform real-space grid and the sphere<->grid maps (``fft.py:118-308``)."""

    shape: tuple
    cell: np.ndarray

    @property
    def size(self) -> int:
        return int(np.prod(self.shape))

    @property
    def volume(self) -> float:
        return abs(float(np.linalg.det(self.cell)))

    @property
    def cell_volume_per_point(self) -> float:
        return self.volume / self.size

    @classmethod
    def for_sphere(cls, sphere: GSphere, *, multiple: int = 1) -> "FFTGrid":
        """``fft.py:164-172``: n >= 2 h_max + 1, rounded to a fast length."""
        hmax = np.abs(sphere.miller).max(axis=0)
        return cls(shape=fast_grid_shape(2 * hmax + 1, multiple), cell=sphere.cell)

    @classmethod
    def for_cutoff(cls, cell, ecut: float, *, multiple: int = 1) -> "FFTGrid":
        """``fft.py:151-162``."""
        return cls.for_sphere(GSphere.build(np.asarray(cell, dtype=float), ecut, sort=False), multiple=multiple)

    def indices(self, sphere: GSphere):
        """``fft.py:174-190``: miller % shape per axis; raises on aliasing."""
        hmax = np.abs(sphere.miller).max(axis=0)
        if np.any(2 * hmax + 1 > np.asarray(self.shape)):
            raise ValueError(
                f"grid {self.shape} is too small for Miller indices up to {tuple(int(h) for h in hmax)}; "
                "high-G components would alias onto low-G ones"
            )
        return tuple(sphere.miller[:, a] % self.shape[a] for a in range(3))

    def scatter(self, coefficients, sphere: GSphere) -> np.ndarray:
        """``fft.py:192-206``: zero grid, then place the sphere coefficients."""
        i, j, k = self.indices(sphere)
        c = np.asarray(coefficients)
        out = np.zeros((*c.shape[:-1], *self.shape), dtype=complex)
        out[..., i, j, k] = c
        return out

    def gather(self, grid, sphere: GSphere) -> np.ndarray:
        """``fft.py:208-212``."""
        i, j, k = self.indices(sphere)
        return np.asarray(grid)[..., i, j, k]

    def to_real(self, g) -> np.ndarray:
        """``fft.py:219-226``: ifftn over the last 3 axes, norm="forward" (no 1/N)."""
        g = np.asarray(g)
        return scipy.fft.ifftn(g, axes=(-3, -2, -1), norm="forward")

    def to_reciprocal(self, r) -> np.ndarray:
        """``fft.py:228-235``: fftn over the last 3 axes, norm="forward" (1/N)."""
        r = np.asarray(r)
        return scipy.fft.fftn(r, axes=(-3, -2, -1), norm="forward")

    def wave_to_real(self, coefficients, sphere: GSphere) -> np.ndarray:
        return self.to_real(self.scatter(coefficients, sphere))

    def real_to_wave(self, values, sphere: GSphere) -> np.ndarray:
        return self.gather(self.to_reciprocal(values), sphere)

    def points_fractional(self) -> np.ndarray:
        axes = [np.arange(n) / n for n in self.shape]
        return np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1)

    def points_cartesian(self) -> np.ndarray:
        return self.points_fractional() @ self.cell

    def gvectors_cartesian(self, *, zero_nyquist: bool = False) -> np.ndarray:
        """``fft.py:254-286``: G of every grid point in fftfreq order."""
        freqs = []
        for n in self.shape:
            f = np.fft.fftfreq(n, d=1.0 / n)
            if zero_nyquist and n % 2 == 0:
                f[n // 2] = 0.0
            freqs.append(f)
        m = np.stack(np.meshgrid(*freqs, indexing="ij"), axis=-1)
        return m @ reciprocal_cell(self.cell)

    def gradient(self, values) -> np.ndarray:
        """``fft.py:288-295``."""
        t = self.to_reciprocal(np.asarray(values, dtype=complex))
        g = self.gvectors_cartesian(zero_nyquist=True)
        return np.stack([self.to_real(1j * g[..., a] * t).real for a in range(3)], axis=-1)

    def divergence(self, field) -> np.ndarray:
        """``fft.py:297-304``."""
        g = self.gvectors_cartesian(zero_nyquist=True)
        acc = np.zeros(self.shape, dtype=complex)
        for a in range(3):
            acc += 1j * g[..., a] * self.to_reciprocal(np.asarray(field[..., a], dtype=complex))
        return self.to_real(acc).real

    def integrate(self, values):
        """``fft.py:306-308``: (Omega/N) sum_r f(r)."""
        return np.sum(values, axis=(-3, -2, -1)) * self.cell_volume_per_point
