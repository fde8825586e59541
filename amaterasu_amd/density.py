"""Electron density on the GPU with the call surface of ``amaterasu.scf.density``.

Mirrors ``scf/density.py:30-84``. The k-weighted sum over occupied bands of |u(r)|^2 is
one library launch per (spin, k): sticks are transformed along the first grid axis,
then each i-plane workgroup transforms every listed band in LDS and accumulates
(2/nspin) w_k f_nk |u|^2 in FP64 registers, writing each density plane once. The
symmetrisation is the caller's ``system.symmetrize_density`` (the identity for a cell
with one symmetry operation, ``hamiltonian/system.py:198-199``); the floor at zero and
the charge rescaling run on the device in FP64.
"""

from __future__ import annotations

import numpy as np
import torch

from .hamiltonian import device_plan


def electron_count(system, density) -> float:
    """``density.py:30-32``."""
    if isinstance(density, torch.Tensor):
        return float(density.sum().item()) * system.grid.cell_volume_per_point
    return float(np.sum(density)) * system.grid.cell_volume_per_point


def _trivial_symmetry(system) -> bool:
    sym = getattr(system, "symmetry", None)
    return sym is None or getattr(sym, "n_operations", 1) <= 1


def compute_density(system, wavefunctions, occupations, *, symmetrize: bool = True, device: bool = False):
    """n_s(r) = (1/Omega) sum_k w_k sum_n (2/nspin) f |u(r)|^2, symmetrised, floored at zero
    and rescaled to the pre-floor charge (``density.py:35-84``).

    ``wavefunctions[spin][ik]`` are (nbands, npw_k) numpy arrays or device tensors.
    Returns a numpy (nspin, *grid) array, or the device tensor when ``device=True``.
    """
    plan = device_plan(system)
    shape = tuple(system.grid.shape)
    nspin = system.nspin
    g = 2.0 / nspin
    occ = np.asarray(occupations.cpu() if isinstance(occupations, torch.Tensor) else occupations, dtype=float)
    weights = np.asarray(system.kpoints.weights, dtype=float)
    rho = torch.zeros((nspin, *shape), dtype=torch.float64, device="cuda")
    for s in range(nspin):
        for ik in range(len(system.wave_spheres)):
            f = g * weights[ik] * occ[s, ik]
            act = np.flatnonzero(f > 1e-12)
            if act.size == 0:
                continue
            wf = wavefunctions[s][ik]
            psi = wf if isinstance(wf, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(wf))
            psi = psi.to("cuda")
            if psi.dtype not in (torch.complex64, torch.complex128):
                psi = psi.to(torch.complex128)
            npw = plan.npw[ik]
            if psi.shape[-1] != plan.npw_pad:
                pad = torch.zeros((psi.shape[0], plan.npw_pad), dtype=psi.dtype, device=psi.device)
                pad[:, :npw] = psi
                psi = pad
            plan.accumulate_density(psi.contiguous(), act.astype(np.int32), f[act], rho[s], k0=ik, beta=1.0,
                                    scale=1.0)
    rho /= system.grid.volume
    if symmetrize:
        if not _trivial_symmetry(system):
            rho = torch.from_numpy(np.asarray(system.symmetrize_density(rho.cpu().numpy()))).to("cuda")
        before = rho.sum(dim=(-3, -2, -1))
        clipped = torch.clamp(rho, min=0.0)
        after = clipped.sum(dim=(-3, -2, -1))
        scale = torch.where(after > 0, before / torch.where(after > 0, after, torch.ones_like(after)),
                            torch.ones_like(after))
        rho = clipped * scale[:, None, None, None]
    return rho if device else rho.cpu().numpy()


__all__ = ["compute_density", "electron_count"]
