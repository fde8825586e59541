"""The Kohn-Sham operator on the GPU, with the call surface of amaterasu's ``Hamiltonian``.

Mirrors ``amaterasu/hamiltonian/operator.py:37-95``:

* ``Hamiltonian(system, effective_potential, coupling=None)`` reshapes a 3-D potential to
  ``(nspin, *grid)`` and raises ``ValueError`` on any other shape (``operator.py:49-58``);
* ``apply(coefficients, ik, spin=0)`` is ``T c + gather(FFT(V . IFFT(scatter c))) + V_NL c``
  (``operator.py:60-69``) -- here the pruned stick/plane FFT kernels, the fused V_eff and
  kinetic epilogues and the MFMA projector GEMMs of libamaterasu_amd;
* ``apply_overlap`` / ``overlap_operator`` / ``local_energy`` as in ``operator.py:71-95``.

``system`` is any object with the reference ``PlaneWaveSystem`` attributes the operator
reads: ``grid.shape``, ``grid.cell_volume_per_point``, ``nspin``, ``wave_spheres`` (each
with ``miller`` and ``g2``) and ``projectors`` (each with ``matrix``, ``dij``, ``labels``).
numpy coefficients in give numpy out; HIP tensors in stay on the device.
"""

from __future__ import annotations

import weakref

import numpy as np
import torch

from . import ops
from ._lib import FP32, FP64, AmtError
from .device import DevicePlan

_PLANS: "weakref.WeakKeyDictionary" = weakref.WeakKeyDictionary()


def device_plan(system) -> DevicePlan:
    """The (cached) device tables of a system: one plan covering all its k-points."""
    plan = None
    try:
        plan = _PLANS.get(system)
    except TypeError:  # not weak-referenceable
        plan = getattr(system, "_amt_plan", None)
    if plan is None or plan.shape != tuple(system.grid.shape) or plan.nk != len(system.wave_spheres):
        plan = DevicePlan(system.grid.shape, system.wave_spheres, system.projectors)
        try:
            _PLANS[system] = plan
        except TypeError:
            system._amt_plan = plan
    elif plan._proj_src is not system.projectors:
        plan.set_projectors(system.projectors)
    return plan


def _precision(name) -> int:
    if name in (FP64, "fp64", "float64", None):
        return FP64
    if name in (FP32, "fp32", "float32"):
        return FP32
    raise ValueError(f"precision must be 'fp64' or 'fp32', got {name!r}")


class Hamiltonian:
    """A Kohn-Sham Hamiltonian at one fixed effective potential, applied on the GPU."""

    def __init__(self, system, effective_potential, coupling=None, *, precision="fp64",
                 shared_potential: bool = False):
        """``shared_potential=True`` copies the potential into the plan's own device buffers
        (:meth:`DevicePlan.shared_potential`) instead of fresh tensors: every Hamiltonian of
        the plan built that way then reads the potential given last, at fixed addresses, and
        ``graph_key`` names that state for Davidson graph replay (None otherwise)."""
        if isinstance(effective_potential, torch.Tensor):
            pot = effective_potential.to(torch.float64)
        else:
            pot = torch.from_numpy(np.asarray(effective_potential, dtype=float))
        if pot.dim() == 3:
            pot = pot[None, ...]
        expected = (system.nspin, *system.grid.shape)
        if tuple(pot.shape) != expected:
            raise ValueError(f"effective potential has shape {tuple(pot.shape)}, expected {expected}")
        self.system = system
        self.coupling = coupling
        self.precision = _precision(precision)
        self.plan = device_plan(system)
        self.graph_key = None
        if shared_potential:
            self._veff = self.plan.shared_potential(pot)
            if coupling is None:
                self.graph_key = ("shared-potential", self.plan.serial)
        else:
            self._veff = {FP64: pot.to("cuda").contiguous()}
            self._veff[FP32] = self._veff[FP64].to(torch.float32)

    @property
    def effective_potential(self) -> np.ndarray:
        """Host copy, shape (nspin, *grid), as the reference stores it."""
        return self._veff[FP64].cpu().numpy()

    def veff(self, spin: int = 0, precision: int | None = None) -> torch.Tensor:
        return self._veff[self.precision if precision is None else precision][spin]

    def _coupling(self, spin):
        return None if self.coupling is None else np.asarray(self.coupling)[spin]

    def apply(self, coefficients, ik: int, spin: int = 0):
        """``H |psi>`` for one band or a block of bands at k-point ``ik``."""
        host = not isinstance(coefficients, torch.Tensor)
        c = torch.from_numpy(np.asarray(coefficients, dtype=complex)) if host else coefficients
        one = c.dim() == 1
        c2 = c[None, :] if one else c
        npw = self.plan.npw[ik]
        if c2.shape[-1] != npw:
            raise ValueError(f"coefficients have {c2.shape[-1]} plane waves, k-point {ik} has {npw}")
        prec = self.precision if host else ops.precision_of(c2)
        c2 = c2.to(device="cuda", dtype=ops.complex_dtype(prec))
        if npw != self.plan.npw_pad:
            pad = torch.zeros((c2.shape[0], self.plan.npw_pad), dtype=c2.dtype, device=c2.device)
            pad[:, :npw] = c2
            c2 = pad
        h = self.plan.hpsi(c2.contiguous(), self.veff(spin, prec), k0=ik, coupling=self._coupling(spin))
        h = h[:, :npw]
        if one:
            h = h[0]
        if host:
            return h.to(torch.complex128).cpu().numpy()
        return h

    def apply_batched(self, psi: torch.Tensor, spin: int = 0, k0: int = 0) -> torch.Tensor:
        """H.psi for the padded k-batch ``[nk][nb][npw_pad]`` in one launch per kernel."""
        return self.plan.hpsi(psi, self.veff(spin, ops.precision_of(psi)), k0=k0, coupling=self._coupling(spin))

    def apply_overlap(self, coefficients, ik: int):
        """``S |psi>``: the identity for norm-conserving datasets (``operator.py:71-81``)."""
        if getattr(self.system, "paw", None) is not None:
            raise NotImplementedError("PAW overlap is not part of the HIP path")
        return coefficients

    @property
    def overlap_operator(self):
        return None if getattr(self.system, "paw", None) is None else self.apply_overlap

    def local_energy(self, density) -> float:
        """``integral V_eff n`` in Hartree (``operator.py:90-95``), reduced on the device in FP64."""
        d = density if isinstance(density, torch.Tensor) else torch.from_numpy(np.asarray(density, dtype=float))
        d = d.to("cuda", torch.float64)
        return float(torch.sum(self._veff[FP64] * d).item()) * self.system.grid.cell_volume_per_point


__all__ = ["Hamiltonian", "device_plan", "AmtError"]
