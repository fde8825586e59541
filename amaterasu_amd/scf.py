"""The self-consistent field loop on the GPU.

Follows the norm-conserving path of ``amaterasu/scf/driver.py:267-464`` step for step --
effective potential (``:225-252``), Davidson per (spin, k) with the tolerance schedule
``clip(1e-3 min(residual, 1), 1e-11, 1e-6)`` and 40/20 steps (``:334-355``), occupations
(``scf/occupations.py:157-244``), the output density (``scf/density.py``), the energy at
the output density (``:375-395``), the eigenvalue-sum identity (``:397-411``), the
integrated density residual (``:413-418``), the convergence test on energy per atom and
residual (``:428``) and Pulay mixing with the Kerker preconditioner
(``scf/mixing.py:49-201``) -- with every array on the device: wavefunctions, density,
potential and mixer history stay in HBM, the grid-side work is FP64 (torch + the
library's FFT), and H.psi / Davidson / density run in libamaterasu_amd.

``system`` is the reference's ``PlaneWaveSystem`` surface (grid, wave_spheres,
projectors, kpoints, nspin, nbands, nelectrons, v_local, core_density, ewald, xc,
density_sphere or ecut_rho, symmetry, initial_density()); amaterasu's own system and
this package's :class:`~amaterasu_amd.system.DeviceSystem` both provide it.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import ops
from .batched import solve_batched
from .density import _trivial_symmetry
from .eigen import davidson, graphs_enabled, orthonormalize
from .hamiltonian import Hamiltonian, device_plan
from .xc import GridOps, XCFunctional

HARTREE_IN_EV = 27.211386024367243
EV_IN_HARTREE = 1.0 / HARTREE_IN_EV


@dataclass
class SCFSettings:
    """The parts of amaterasu's ``Config`` the loop reads (defaults of ``config/schema.py:153-215``)."""

    max_steps: int = 40
    tol_energy: float = 1e-8  # eV/atom, as scf.tol_energy
    tol_density: float = 1e-7
    scheme: str = "fixed"
    smearing: float = 0.0  # Hartree
    smearing_order: int = 1  # Methfessel-Paxton order
    mixing: str = "pulay"
    beta: float = 0.4
    history: int = 8
    kerker_q0: float = 1.5
    eigensolver: str = "davidson"
    #: H.psi / Davidson block precision: "fp64", "fp32", or "mixed" = FP32 while the density
    #: residual is above ``fp64_below``, FP64 after (grid-side work is always FP64).
    precision: str = "fp64"
    fp64_below: float = 1e-4
    #: FP32 blocks cannot resolve residual norms below a few 1e-7 ||H||; the eigensolver
    #: tolerance is floored here while they are in use.
    fp32_tol_floor: float = 2e-6
    #: Run the Davidson of all k-points of a spin channel together (one launch per kernel
    #: over every k); each k-point still follows its own iteration.
    batched_k: bool = True
    #: Replay the expansion half of each per-k Davidson pass from HIP graphs captured per
    #: (subspace size, new directions) (eigen.StepGraphs); the potential then lives in one
    #: device buffer per plan, rewritten every iteration. Default: ``AMT_DAVIDSON_GRAPHS=1``.
    davidson_graphs: bool = field(default_factory=graphs_enabled)


@dataclass
class Energies:
    kinetic: float = 0.0
    local: float = 0.0
    hartree: float = 0.0
    xc: float = 0.0
    nonlocal_: float = 0.0
    ewald: float = 0.0
    entropy: float = 0.0

    @property
    def total(self) -> float:
        return self.kinetic + self.local + self.hartree + self.xc + self.nonlocal_ + self.ewald

    @property
    def free_energy(self) -> float:
        return self.total - self.entropy

    @property
    def extrapolated(self) -> float:
        """E0 = (E + F)/2, the sigma -> 0 estimate (``driver.py:83-86``)."""
        return self.total - 0.5 * self.entropy


@dataclass
class SCFResult:
    converged: bool
    iterations: int
    energies: Energies
    eigenvalues: np.ndarray
    occupations: np.ndarray
    fermi_level: float
    gap: float | None
    density: np.ndarray
    wavefunctions: list
    history: list = field(default_factory=list)
    electron_count: float = 0.0
    seconds: float = 0.0
    davidson_steps: int = 0
    #: total moment in Bohr magnetons, sum_k w_k sum_n (f_up - f_down) (``occupations.py:236-237``)
    magnetization: float = 0.0


class GridSide:
    """Hartree, Kerker and symmetrisation on the full grid (FP64, device)."""

    def __init__(self, system):
        self.grid = system.grid
        self.g = GridOps(system.grid.shape, system.grid.cell)
        ecut_rho = getattr(system, "ecut_rho", None)
        if ecut_rho is None:
            ecut_rho = float(system.density_sphere.ecut)
        self.sphere = (self.g.g2 <= 2.0 * ecut_rho + 1e-12)
        self.nonzero = self.sphere & (self.g.g2 > 1e-12)
        self.coulomb = torch.where(self.nonzero, 4.0 * np.pi / torch.where(self.nonzero, self.g.g2,
                                                                            torch.ones_like(self.g.g2)),
                                   torch.zeros_like(self.g.g2))
        self._sym = None
        if not _trivial_symmetry(system):
            self._sym = self._symmetry_maps(system)

    def hartree(self, total: torch.Tensor):
        """``hartree.py:28-50``: (E_H, V_H(r))."""
        c = self.g.to_reciprocal(total)
        e = 0.5 * self.g.volume * float((self.coulomb * (c.real**2 + c.imag**2)).sum().item())
        return e, self.g.to_real(c * self.coulomb).real.contiguous()

    def kerker(self, q0: float):
        """``mixing.py:49-106``: G^2/(G^2+q0^2) on the charge residual over every grid
        frequency (G = 0 removed); with two spin channels the moment passes unfiltered."""
        f = torch.where(self.g.g2 > 1e-12, self.g.g2 / (self.g.g2 + q0 * q0), torch.zeros_like(self.g.g2))

        def filt(x):
            return self.g.to_real(self.g.to_reciprocal(x) * f).real

        def apply(r):
            if q0 <= 0:
                return r
            if r.shape[0] == 1:
                return filt(r[0])[None]
            charge, moment = filt(r[0] + r[1]), r[0] - r[1]
            return torch.stack([0.5 * (charge + moment), 0.5 * (charge - moment)])

        return apply

    def _symmetry_maps(self, system):
        """Per operation: the grid index of R^T G for each density-sphere G, and exp(-2 pi i G.f)."""
        shape = np.array(self.grid.shape)
        mill = np.stack(np.meshgrid(*[np.fft.fftfreq(n, d=1.0 / n) for n in shape], indexing="ij"), -1)
        mask = self.sphere.cpu().numpy()
        h = np.rint(mill[mask]).astype(int)
        flat = np.ravel_multi_index(tuple((h % shape).T), tuple(shape))
        lookup = np.full(int(np.prod(shape)), -1, dtype=np.int64)
        lookup[flat] = np.arange(len(h))
        imgs, phases = [], []
        for rot, tr in zip(system.symmetry.rotations, system.symmetry.translations):
            r = h @ np.asarray(rot, dtype=int)
            idx = lookup[np.ravel_multi_index(tuple((r % shape).T), tuple(shape))]
            if np.any(idx < 0):
                raise ValueError("a symmetry operation maps a G-vector outside the cutoff sphere")
            imgs.append(idx)
            phases.append(np.exp(-2j * np.pi * (h @ np.asarray(tr, dtype=float))))
        dev = self.g.g.device
        return (torch.from_numpy(flat).to(dev), torch.from_numpy(np.stack(imgs)).to(dev),
                torch.from_numpy(np.stack(phases)).to(dev))

    def symmetrize(self, rho: torch.Tensor) -> torch.Tensor:
        """``system.py:192-212`` with ``symmetry.py:327-370`` on the device."""
        if self._sym is None:
            return rho
        flat, imgs, phases = self._sym
        out = torch.empty_like(rho)
        for s in range(rho.shape[0]):
            c = self.g.to_reciprocal(rho[s]).reshape(-1)[flat]
            acc = (c[imgs] * phases).sum(0) / imgs.shape[0]
            grid = torch.zeros(self.g.size, dtype=torch.complex128, device=rho.device)
            grid[flat] = acc
            out[s] = self.g.to_real(grid.reshape(self.grid.shape)).real
        return out


def _hermite_gauss(x: torch.Tensor, mmax: int) -> list:
    """H_m(x) e^{-x^2} for m = 0..mmax (H_{m+1} = 2x H_m - 2m H_{m-1})."""
    e = torch.exp(-torch.clamp(x * x, max=200.0))
    h = [e, 2.0 * x * e]
    for m in range(1, mmax):
        h.append(2.0 * x * h[m] - 2.0 * m * h[m - 1])
    return h


def _mp_coefficient(n: int) -> float:
    """Methfessel-Paxton A_n = (-1)^n / (n! 4^n sqrt(pi)) (PRB 40, 3616)."""
    return (-1.0) ** n / (math.factorial(n) * 4.0**n * np.sqrt(np.pi))


def occupations(eig: torch.Tensor, weights: torch.Tensor, nelectrons: float, scheme: str, smearing: float,
                nspin: int, tolerance: float = 1e-12, order: int = 1):
    """``occupations.py:134-244`` (fixed, Fermi-Dirac, Gaussian, Methfessel-Paxton of ``order``,
    Marzari-Vanderbilt) in FP64 on the device.
    Returns (occupations, fermi level, entropy term, gap or None)."""
    if eig.dim() != 3:
        raise ValueError(f"eigenvalues must be (nspin, nk, nbands), got {tuple(eig.shape)}")
    if scheme == "tetrahedron":  # occupations.py:192-196
        raise NotImplementedError("the linear tetrahedron method is not implemented; use occupations.scheme: "
                                  "gaussian, methfessel-paxton or marzari-vanderbilt")
    g = 2.0 / nspin
    w = weights.to(eig)
    if scheme == "fixed" or smearing <= 0.0:
        flat = eig.reshape(-1)
        order = torch.sort(flat, stable=True).indices
        sw = (w[None, :, None] * g).expand(eig.shape).reshape(-1)[order]
        cum = torch.cumsum(sw, 0)
        occ = torch.zeros_like(flat)
        full = cum <= nelectrons + 1e-12
        occ[order[full]] = 1.0
        b = int(full.sum().item())
        if b < len(order):
            rem = nelectrons - (float(cum[b - 1].item()) if b else 0.0)
            occ[order[b]] = max(0.0, rem / float(sw[b].item()))
        homo = float(flat[order[max(b - 1, 0)]].item())
        lumo = float(flat[order[b]].item()) if b < len(order) else homo
        gap = lumo - homo
        return occ.reshape(eig.shape), 0.5 * (homo + lumo), 0.0, (None if gap <= 0 else gap)

    def occfun(x):
        if scheme == "fermi-dirac":
            return 1.0 / (1.0 + torch.exp(torch.clamp(x, -200.0, 200.0)))
        if scheme == "marzari-vanderbilt":
            sh = x + 1.0 / np.sqrt(2.0)
            return 0.5 * torch.erfc(sh) + torch.exp(-torch.clamp(sh * sh, max=200.0)) / np.sqrt(2.0 * np.pi)
        if scheme == "gaussian":
            return 0.5 * torch.erfc(x)
        if scheme == "methfessel-paxton":  # occupations.py:47-68,84-89
            h = _hermite_gauss(x, 2 * order)
            return 0.5 * torch.erfc(x) + sum(_mp_coefficient(n) * h[2 * n - 1] for n in range(1, order + 1))
        raise ValueError(f"unknown smearing scheme {scheme!r}")

    def entfun(x):
        if scheme == "fermi-dirac":
            f = torch.clamp(occfun(x), 1e-300, 1.0 - 1e-16)
            return -(f * torch.log(f) + (1.0 - f) * torch.log1p(-f))
        if scheme == "marzari-vanderbilt":
            sh = x + 1.0 / np.sqrt(2.0)
            return sh * torch.exp(-torch.clamp(sh * sh, max=200.0)) / np.sqrt(2.0 * np.pi)
        if scheme == "methfessel-paxton":  # S' = x f' (occupations.py:106-111)
            h = _hermite_gauss(x, 2 * order)
            return 0.5 * h[0] / np.sqrt(np.pi) + sum(_mp_coefficient(n) * (0.5 * h[2 * n] + 2 * n * h[2 * n - 2])
                                                     for n in range(1, order + 1))
        return 0.5 * torch.exp(-torch.clamp(x * x, max=200.0)) / np.sqrt(np.pi)

    def electrons(mu):
        return g * float((w[None, :, None] * occfun((eig - mu) / smearing)).sum().item())

    lo = float(eig.min().item()) - 20.0 * smearing - 1.0
    hi = float(eig.max().item()) + 20.0 * smearing + 1.0
    if electrons(hi) < nelectrons - 1e-9:
        raise ValueError(f"only {electrons(hi):.4f} of {nelectrons:g} electrons fit in {eig.shape[2]} bands")
    for _ in range(200):
        mu = 0.5 * (lo + hi)
        if electrons(mu) < nelectrons:
            lo = mu
        else:
            hi = mu
        if hi - lo < tolerance:
            break
    mu = 0.5 * (lo + hi)
    x = (eig - mu) / smearing
    occ = occfun(x)
    ent = smearing * g * float((w[None, :, None] * entfun(x)).sum().item())
    gap = None
    if float(torch.minimum(occ, 1.0 - occ).max().item()) < 1e-6:
        filled, empty = eig[occ > 0.5], eig[occ <= 0.5]
        if filled.numel() and empty.numel():
            gap = float(empty.min().item() - filled.max().item())
    return occ, mu, ent, (None if gap is None or gap <= 0 else gap)


class PulayMixer:
    """``mixing.py:131-201`` with the history in device memory.

    The residual overlap matrix is kept incrementally: each new residual is dotted with the
    stored ones in one fused multiply-and-reduce (a history x grid matrix product here is a
    skinny GEMM the BLAS library handles poorly), and the oldest row and column are dropped
    with the oldest residual."""

    def __init__(self, beta=0.4, history=8, precondition=None, linear=False):
        self.beta, self.history, self.linear = beta, history, linear
        self.precondition = precondition or (lambda r: r)
        self.inputs, self.residuals = [], []
        self._ov = np.zeros((0, 0))

    def _drop_oldest(self) -> None:
        del self.inputs[0]
        del self.residuals[0]
        self._ov = self._ov[1:, 1:]

    def __call__(self, din: torch.Tensor, dout: torch.Tensor) -> torch.Tensor:
        r = dout - din
        self.inputs.append(din.clone())
        self.residuals.append(r)
        if not self.linear:
            n = len(self.residuals)
            flat = torch.stack([x.reshape(-1) for x in self.residuals])
            dots = (flat * r.reshape(1, -1)).sum(1).cpu().numpy()
            ov = np.zeros((n, n))
            ov[: n - 1, : n - 1] = self._ov
            ov[n - 1, :] = ov[:, n - 1] = dots
            self._ov = ov
        if len(self.inputs) > self.history:
            self._drop_oldest()
        if self.linear:
            return din + self.beta * self.precondition(r)
        while True:
            n = len(self.residuals)
            if n == 1:
                c = np.ones(1)
                break
            ov = self._ov
            scale = float(np.trace(ov)) / n
            m = np.ones((n + 1, n + 1))
            m[:n, :n] = ov / scale
            m[n, n] = 0.0
            t = np.zeros(n + 1)
            t[n] = 1.0
            try:
                sol = np.linalg.solve(m, t)
            except np.linalg.LinAlgError:
                sol = None
            if sol is not None and np.all(np.isfinite(sol)) and np.max(np.abs(sol[:n])) < 1e4:
                c = sol[:n]
                break
            self._drop_oldest()
        ct = torch.from_numpy(c).to(din)
        xin = torch.einsum("i,i...->...", ct, torch.stack(self.inputs))
        xr = torch.einsum("i,i...->...", ct, torch.stack(self.residuals))
        return xin + self.beta * self.precondition(xr)


class BroydenMixer:
    """Johnson's modified Broyden (PRB 38, 12807; ``mixing.py:217-268``) with the history in
    device memory: secant pairs of successive residual and input differences, normalised by
    the residual change; ``w0`` regularises the normal equations."""

    def __init__(self, beta=0.4, history=8, precondition=None, w0=0.01, weight=1.0):
        self.beta, self.history, self.w0, self.weight = beta, history, w0, weight
        self.precondition = precondition or (lambda r: r)
        self.inputs, self.residuals = [], []

    def __call__(self, din: torch.Tensor, dout: torch.Tensor) -> torch.Tensor:
        r = dout - din
        self.inputs.append(din.clone())
        self.residuals.append(r)
        if len(self.inputs) > self.history:
            del self.inputs[0]
            del self.residuals[0]
        step = self.beta * self.precondition(r)
        if len(self.residuals) < 2:
            return din + step
        df, dn = [], []
        for i in range(len(self.residuals) - 1):
            diff = self.residuals[i + 1] - self.residuals[i]
            nrm = float(torch.linalg.vector_norm(diff).item())
            if nrm < 1e-30:
                continue
            df.append(diff.reshape(-1) / nrm)
            dn.append((self.inputs[i + 1] - self.inputs[i] + self.beta * self.precondition(diff)).reshape(-1) / nrm)
        if not df:
            return din + step
        dfm, dnm = torch.stack(df), torch.stack(dn)
        m = self.weight**2 * (dfm @ dfm.T) + self.w0**2 * torch.eye(len(df), dtype=dfm.dtype, device=dfm.device)
        gamma = torch.linalg.solve(m, self.weight * (dfm @ r.reshape(-1))) * self.weight
        return din + step - (gamma @ dnm).reshape(din.shape)


def make_mixer(name: str, beta: float = 0.4, history: int = 8, precondition=None):
    """``mixing.py:263-286``: linear, anderson (Pulay over at most two iterations), pulay,
    broyden; an unknown name is a ValueError."""
    if name == "linear":
        return PulayMixer(beta, history, precondition, linear=True)
    if name == "anderson":
        return PulayMixer(beta, min(history, 2), precondition)
    if name == "pulay":
        return PulayMixer(beta, history, precondition)
    if name == "broyden":
        return BroydenMixer(beta, history, precondition)
    raise ValueError(f"unknown mixing algorithm {name!r}; available: linear, anderson, pulay, broyden")


def _as_device(x, dtype=torch.float64):
    if x is None:
        return None
    if isinstance(x, torch.Tensor):
        return x.to("cuda", dtype)
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=float))).to("cuda", dtype)


def _xc_of(system):
    x = getattr(system, "xc", None)
    if isinstance(x, XCFunctional):
        return x
    name = getattr(x, "name", None) or (x if isinstance(x, str) else "PBE")
    return XCFunctional(name)


def effective_potential(system, grid_side: GridSide, xc: XCFunctional, density: torch.Tensor, vloc, core):
    """``driver.py:225-252``: v_loc + V_H[n] + v_xc[n + n_core/nspin]."""
    eh, vh = grid_side.hartree(density.sum(0))
    xd = density if core is None else density + core[None] / system.nspin
    exc, vxc = xc.evaluate_on_grid(xd, grid_side.g)
    return (vloc[None] + vh[None] + vxc).contiguous(), eh, exc


def initial_wavefunctions(system, plan, dtype, seed=12345):
    """``driver.py:255-264``: the reference's default_rng draws (seed + 991 spin + ik), damped by
    1/(1+T), orthonormalised on the device, in the padded layout [spin][k][band][npw_pad]."""
    extra = getattr(system, "extra", None)
    key = ("_initial_wavefunctions", seed, system.nspin, plan.nk, system.nbands, plan.npw_pad)
    if isinstance(extra, dict) and key in extra:  # the same seeded draws, already orthonormalised
        return extra[key].to(dtype, copy=True)
    wf = torch.zeros((system.nspin, plan.nk, system.nbands, plan.npw_pad), dtype=torch.complex128, device="cuda")
    for sp in range(system.nspin):
        for ik, s in enumerate(system.wave_spheres):
            n = len(s.g2)
            rng = np.random.default_rng(seed + 991 * sp + ik)
            b = rng.normal(size=(system.nbands, n)) + 1j * rng.normal(size=(system.nbands, n))
            b = b / (1.0 + 0.5 * np.asarray(s.g2))
            wf[sp, ik, :, :n] = orthonormalize(torch.from_numpy(b).to("cuda"))
    if isinstance(extra, dict):
        extra[key] = wf
    return wf.to(dtype, copy=True)


def run_scf(system, settings: SCFSettings, *, density=None, wavefunctions=None, log=None) -> SCFResult:
    """The SCF loop of ``driver.py:267-464`` with all state on the device."""
    t_start = time.perf_counter()
    plan = device_plan(system)
    if settings.precision not in ("fp64", "fp32", "mixed"):
        raise ValueError(f"precision must be fp64, fp32 or mixed, got {settings.precision!r}")
    cdt = torch.complex128 if settings.precision == "fp64" else torch.complex64
    nspin, nk, nb = system.nspin, plan.nk, system.nbands
    natoms = len(system.symbols)
    side = GridSide(system)
    xc = _xc_of(system)
    vloc = _as_device(system.v_local)
    core = _as_device(system.core_density)
    if density is None:
        density = system.initial_density()
    rho = _as_device(density).reshape(nspin, *system.grid.shape).clone()
    wf = wavefunctions if wavefunctions is not None else initial_wavefunctions(system, plan, cdt)
    wf = wf.to(cdt)
    weights = torch.as_tensor(np.asarray(system.kpoints.weights, dtype=float), device="cuda")
    mixer = make_mixer(settings.mixing, settings.beta, settings.history,
                       side.kerker(settings.kerker_q0) if settings.kerker_q0 > 0 else None)
    ethr = settings.tol_energy * natoms * EV_IN_HARTREE
    dv = system.grid.cell_volume_per_point
    hist, prev, resid, conv = [], np.inf, np.inf, False
    energies = Energies(ewald=float(system.ewald))
    eig = torch.zeros((nspin, nk, nb), dtype=torch.float64, device="cuda")
    occ = torch.zeros_like(eig)
    fermi, gap, dout = 0.0, None, rho
    it = dav_steps = 0
    kin = plan.kinetic[1]
    for it in range(1, settings.max_steps + 1):
        pot, _, _ = effective_potential(system, side, xc, rho, vloc, core)
        if settings.precision == "mixed" and cdt == torch.complex64 and resid < settings.fp64_below:
            cdt = torch.complex128
            wf = wf.to(cdt)
        prec = "fp64" if cdt == torch.complex128 else "fp32"
        ham = Hamiltonian(system, pot, precision=prec, shared_potential=settings.davidson_graphs)
        tol = float(np.clip(1e-3 * min(resid, 1.0), 1e-11, 1e-6))
        if prec == "fp32":
            tol = max(tol, settings.fp32_tol_floor)
        steps = 40 if it == 1 else 20
        steps_before = dav_steps
        for sp in range(nspin):
            if settings.batched_k and nk > 1:
                kin_b = plan.kinetic[1 if cdt == torch.complex128 else 0]
                r = solve_batched(lambda b, sp=sp: ham.apply_batched(b, sp), wf[sp].contiguous(), kin_b,
                                  plan.npw, tolerance=tol, max_iterations=steps)
                eig[sp] = r.eigenvalues
                wf[sp] = r.coefficients.to(cdt)
                dav_steps += int(r.iterations.sum())
                continue
            for ik in range(nk):
                n = plan.npw[ik]
                gkey = (ham.graph_key, ik, sp) if ham.graph_key is not None else None
                r = davidson(lambda b, ik=ik, sp=sp: ham.apply(b, ik, sp), wf[sp, ik, :, :n].contiguous(),
                             kin[ik, :n], tolerance=tol, max_iterations=steps, precision=prec, graph_key=gkey)
                eig[sp, ik] = r.eigenvalues
                wf[sp, ik, :, :n] = r.coefficients.to(cdt)
                dav_steps += r.iterations
        occ, fermi, ent, gap = occupations(eig, weights, system.nelectrons, settings.scheme, settings.smearing,
                                           nspin, order=settings.smearing_order)
        # output density: one batched launch over all k per spin
        dout = torch.zeros_like(rho)
        f_all = (2.0 / nspin) * weights[None, :, None] * occ
        for sp in range(nspin):
            f = f_all[sp].reshape(-1).cpu().numpy()
            act = np.flatnonzero(f > 1e-12)
            plan.accumulate_density(wf[sp], act.astype(np.int32), f[act], dout[sp], k0=0, beta=0.0,
                                    scale=1.0 / system.grid.volume)
        dout = side.symmetrize(dout)
        before = dout.sum(dim=(-3, -2, -1))
        clipped = torch.clamp(dout, min=0.0)
        after = clipped.sum(dim=(-3, -2, -1))
        dout = clipped * torch.where(after > 0, before / torch.where(after > 0, after, torch.ones_like(after)),
                                     torch.ones_like(after))[:, None, None, None]
        # energies at the output density
        eh, _ = side.hartree(dout.sum(0))
        xd = dout if core is None else dout + core[None] / nspin
        exc, _ = xc.evaluate_on_grid(xd, side.g)
        ekin = enl = 0.0
        kb = plan.kb_tables(1)
        for sp in range(nspin):
            psi = wf[sp].to(torch.complex128)
            dens2 = psi.real**2 + psi.imag**2
            # sum over plane waves as a reduction: as a matrix-vector product (K = npw, one output
            # column) it ran on a handful of workgroups
            ekin += float((f_all[sp] * (dens2 * kin[:, None, :]).sum(-1)).sum().item())
            if kb is not None:
                mat, dij = kb[0], kb[1]
                p = ops.zgemm(psi.contiguous(), mat, 1)
                pd = torch.einsum("knj,kji->kni", p, dij.to(torch.complex128))
                quad = (p.conj() * pd).real.sum(-1)
                enl += float((f_all[sp] * quad).sum().item())
        energies = Energies(kinetic=ekin, local=float((vloc * dout.sum(0)).sum().item()) * dv, hartree=eh, xc=exc,
                            nonlocal_=enl, ewald=float(system.ewald), entropy=ent)
        band_sum = (2.0 / nspin) * float((weights[None, :, None] * occ * eig).sum().item())
        local_e = float((pot * dout).sum().item()) * dv
        identity = band_sum - (energies.kinetic + local_e + energies.nonlocal_)
        resid = float(torch.abs(dout - rho).sum().item()) * dv / system.nelectrons
        e = energies.free_energy
        delta = e - prev
        hist.append((e, resid, identity))
        if log:
            log(f"{it:4d} {e * HARTREE_IN_EV:18.9f} {delta * HARTREE_IN_EV / natoms:14.2e} {resid:13.3e} "
                f"{identity * HARTREE_IN_EV:12.2e}  {prec} tol {tol:.1e} davidson {dav_steps - steps_before}")
        if it > 1 and abs(delta) < ethr and resid < settings.tol_density:
            conv = True
            break
        prev = e
        rho = mixer(rho, dout)
    wfs = [[wf[sp, ik, :, : plan.npw[ik]].to(torch.complex128).cpu().numpy() for ik in range(nk)]
           for sp in range(nspin)]
    count = float(dout.sum().item()) * dv
    occ_h = occ.cpu().numpy()
    magnetization = 0.0
    if occ_h.shape[0] == 2:
        per_spin = np.einsum("k,skn->s", np.asarray(system.kpoints.weights, dtype=float), occ_h)
        magnetization = float(per_spin[0] - per_spin[1])
    return SCFResult(conv, it, energies, eig.cpu().numpy(), occ_h, fermi, gap, dout.cpu().numpy(), wfs,
                     hist, count, time.perf_counter() - t_start, dav_steps, magnetization)


__all__ = ["BroydenMixer", "Energies", "GridSide", "PulayMixer", "SCFResult", "SCFSettings", "make_mixer",
           "occupations", "run_scf"]
