"""The SCF loop and its grid-side pieces (FP64, CPU).

Restates ``amaterasu/hamiltonian/hartree.py:28-50`` (Hartree), the fixed and smeared
branches of ``amaterasu/scf/occupations.py:134-244`` (Fermi-Dirac, Gaussian,
Marzari-Vanderbilt), ``amaterasu/scf/mixing.py:49-201`` (linear and Pulay mixing with
the Kerker preconditioner) and the norm-conserving path of ``run_scf``
(``amaterasu/scf/driver.py:225-464``): effective potential, Davidson per (spin, k) with
the tolerance schedule of ``driver.py:334-335``, occupations, density, energies, the
eigenvalue-sum identity, the density residual and the convergence test of
``driver.py:428``.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import math

import numpy as np
from scipy.special import erfc

from .density import compute_density, electron_count
from .eigen import random_bands, solve
from .hamiltonian import Hamiltonian
from .units import EV_IN_HARTREE, HARTREE_IN_EV


# ------------------------------------------------------------------------- Hartree
def hartree(grid, sphere, density):
    """``hartree.py:28-50``: V_H(G) = 4 pi n(G)/G^2 on the density sphere, G=0 dropped."""
    density = np.asarray(density, dtype=float)
    if density.shape != tuple(grid.shape):
        raise ValueError(f"density has shape {density.shape}, grid is {grid.shape}")
    c = grid.gather(grid.to_reciprocal(density.astype(complex)), sphere)
    nz = sphere.g2 > 1e-12
    vg = np.zeros_like(c)
    vg[nz] = 4.0 * np.pi * c[nz] / sphere.g2[nz]
    energy = 0.5 * grid.volume * float(np.sum(4.0 * np.pi * np.abs(c[nz]) ** 2 / sphere.g2[nz]))
    return energy, np.ascontiguousarray(grid.to_real(grid.scatter(vg, sphere)).real)


# --------------------------------------------------------------------- occupations
_MAXARG = 200.0


def _hermite_gauss(x, mmax):
    """H_m(x) e^{-x^2} for m = 0..mmax, by H_{m+1} = 2x H_m - 2m H_{m-1}."""
    e = np.exp(-np.minimum(_MAXARG, x**2))
    h = [e, 2.0 * x * e]
    for m in range(1, mmax):
        h.append(2.0 * x * h[m] - 2.0 * m * h[m - 1])
    return h


def _mp_coefficient(n):
    """Methfessel-Paxton (PRB 40, 3616) A_n = (-1)^n / (n! 4^n sqrt(pi))."""
    return (-1.0) ** n / (math.factorial(n) * 4.0**n * np.sqrt(np.pi))


def occupation_function(x, scheme, order=1):
    """``occupations.py:71-89``."""
    x = np.asarray(x, dtype=float)
    if scheme == "methfessel-paxton":  # occupations.py:47-68,84-89
        h = _hermite_gauss(x, 2 * order)
        return 0.5 * erfc(x) + sum(_mp_coefficient(n) * h[2 * n - 1] for n in range(1, order + 1))
    if scheme == "fermi-dirac":
        return 1.0 / (1.0 + np.exp(np.clip(x, -_MAXARG, _MAXARG)))
    if scheme == "marzari-vanderbilt":
        sh = x + 1.0 / np.sqrt(2.0)
        return 0.5 * erfc(sh) + np.exp(-np.minimum(_MAXARG, sh**2)) / np.sqrt(2.0 * np.pi)
    if scheme in ("gaussian", "fixed"):
        return 0.5 * erfc(x)
    raise ValueError(f"unknown smearing scheme {scheme!r}")


def entropy_function(x, scheme, order=1):
    """``occupations.py:92-111``; for Methfessel-Paxton S = e^{-x^2}/(2 sqrt pi) +
    sum_n A_n (H_{2n}/2 + 2n H_{2n-2}) e^{-x^2}, the function with S' = x f'."""
    x = np.asarray(x, dtype=float)
    if scheme == "methfessel-paxton":  # occupations.py:106-111
        h = _hermite_gauss(x, 2 * order)
        return 0.5 * h[0] / np.sqrt(np.pi) + sum(_mp_coefficient(n) * (0.5 * h[2 * n] + 2 * n * h[2 * n - 2])
                                                 for n in range(1, order + 1))
    if scheme == "fermi-dirac":
        f = np.clip(occupation_function(x, scheme), 1e-300, 1.0 - 1e-16)
        return -(f * np.log(f) + (1.0 - f) * np.log1p(-f))
    if scheme == "marzari-vanderbilt":
        sh = x + 1.0 / np.sqrt(2.0)
        return sh * np.exp(-np.minimum(_MAXARG, sh**2)) / np.sqrt(2.0 * np.pi)
    if scheme in ("gaussian", "fixed"):
        return 0.5 * np.exp(-np.minimum(_MAXARG, x**2)) / np.sqrt(np.pi)
    raise ValueError(f"unknown smearing scheme {scheme!r}")


@dataclass
class OccupationResult:
    occupations: np.ndarray
    fermi_level: float
    entropy_term: float
    gap: float | None


def compute_occupations(eig, weights, nelectrons, *, scheme="gaussian", smearing=0.0, order=1, nspin=1,
                        tolerance=1e-12):
    """``occupations.py:157-244``: zero-temperature filling or a 200-step bisection on mu."""
    eig = np.asarray(eig, dtype=float)
    w = np.asarray(weights, dtype=float)
    g = 2.0 / nspin
    if scheme == "fixed" or smearing <= 0.0:
        flat = eig.ravel()
        order = np.argsort(flat, kind="stable")
        sw = np.broadcast_to(w[None, :, None] * g, eig.shape).ravel()[order]
        cum = np.cumsum(sw)
        occ = np.zeros_like(flat)
        full = cum <= nelectrons + 1e-12
        occ[order[full]] = 1.0
        b = int(np.count_nonzero(full))
        if b < len(order):
            rem = nelectrons - (cum[b - 1] if b else 0.0)
            occ[order[b]] = max(0.0, rem / sw[b])
        homo = flat[order[max(b - 1, 0)]]
        lumo = flat[order[b]] if b < len(order) else homo
        gap = lumo - homo
        return OccupationResult(occ.reshape(eig.shape), float(0.5 * (homo + lumo)), 0.0,
                                None if gap <= 0 else float(gap))
    lo = float(eig.min()) - 20.0 * smearing - 1.0
    hi = float(eig.max()) + 20.0 * smearing + 1.0

    def electrons(mu):
        return g * float(np.einsum("k,skn->", w, occupation_function((eig - mu) / smearing, scheme, order)))

    if electrons(hi) < nelectrons - 1e-9:
        raise ValueError("not enough bands for the electrons")
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
    occ = occupation_function(x, scheme, order)
    ent = smearing * g * float(np.einsum("k,skn->", w, entropy_function(x, scheme, order)))
    gap = None
    if float(np.max(np.minimum(occ, 1.0 - occ))) < 1e-6:
        filled, empty = eig[occ > 0.5], eig[occ <= 0.5]
        if filled.size and empty.size:
            gap = float(empty.min() - filled.max())
    return OccupationResult(occ, float(mu), float(ent), None if gap is None or gap <= 0 else gap)


# -------------------------------------------------------------------------- mixing
class Kerker:
    """G^2/(G^2 + q0^2) over every grid frequency, G = 0 -> 0 (``mixing.py:49-106``)."""

    def __init__(self, grid, q0):
        g = grid.gvectors_cartesian()
        g2 = np.einsum("...i,...i->...", g, g)
        with np.errstate(invalid="ignore"):
            self.factor = np.where(g2 > 1e-12, g2 / (g2 + q0**2), 0.0)
        self.grid, self.q0 = grid, q0

    def _filter(self, f):
        return self.grid.to_real(self.grid.to_reciprocal(f.astype(complex)) * self.factor).real

    def __call__(self, r):
        """``mixing.py:97-106``: the charge residual is filtered, the moment passes through."""
        if self.q0 <= 0:
            return r
        if r.shape[0] == 1:
            return self._filter(r[0])[None]
        charge, moment = self._filter(r[0] + r[1]), r[0] - r[1]
        return np.stack([0.5 * (charge + moment), 0.5 * (charge - moment)])


class PulayMixer:
    """DIIS with a bordered, trace-scaled system; drop-oldest on failure (``mixing.py:131-201``)."""

    def __init__(self, beta=0.4, history=8, precondition=None, linear=False):
        self.beta, self.history, self.linear = beta, history, linear
        self.precondition = precondition or (lambda r: r)
        self.inputs, self.residuals = [], []

    def _record(self, din, dout):
        r = dout - din
        self.inputs.append(din.copy())
        self.residuals.append(r)
        if len(self.inputs) > self.history:
            del self.inputs[0]
            del self.residuals[0]
        return r

    def _coefficients(self):
        while True:
            n = len(self.residuals)
            if n == 1:
                return np.ones(1)
            flat = np.array([r.ravel() for r in self.residuals])
            ov = flat @ flat.T
            scale = float(np.trace(ov)) / n
            m = np.ones((n + 1, n + 1))
            m[:n, :n] = ov / scale
            m[n, n] = 0.0
            t = np.zeros(n + 1)
            t[n] = 1.0
            try:
                sol = np.linalg.solve(m, t)
            except np.linalg.LinAlgError:
                pass
            else:
                if np.all(np.isfinite(sol)) and np.max(np.abs(sol[:n])) < 1e4:
                    return sol[:n]
            del self.inputs[0]
            del self.residuals[0]

    def __call__(self, din, dout):
        r = self._record(din, dout)
        if self.linear:
            return din + self.beta * self.precondition(r)
        c = self._coefficients()
        xin = np.einsum("i,i...->...", c, np.array(self.inputs))
        xr = np.einsum("i,i...->...", c, np.array(self.residuals))
        return xin + self.beta * self.precondition(xr)


class BroydenMixer(PulayMixer):
    """Johnson's modified Broyden (PRB 38, 12807; ``mixing.py:217-268``): secant pairs of
    successive residual and input differences, normalised by the residual change, weighted
    by ``weight`` against the regulariser ``w0``."""

    def __init__(self, beta=0.4, history=8, precondition=None, w0=0.01, weight=1.0):
        super().__init__(beta, history, precondition)
        self.w0, self.weight = w0, weight

    def __call__(self, din, dout):
        r = self._record(din, dout)
        step = self.beta * self.precondition(r)
        if len(self.residuals) < 2:
            return din + step
        df, dn = [], []
        for i in range(len(self.residuals) - 1):
            diff = self.residuals[i + 1] - self.residuals[i]
            nrm = float(np.linalg.norm(diff))
            if nrm < 1e-30:
                continue
            df.append(diff.ravel() / nrm)
            dn.append((self.inputs[i + 1] - self.inputs[i] + self.beta * self.precondition(diff)).ravel() / nrm)
        if not df:
            return din + step
        df, dn = np.array(df), np.array(dn)
        m = self.weight**2 * (df @ df.T) + self.w0**2 * np.eye(len(df))
        gamma = np.linalg.solve(m, self.weight * (df @ r.ravel())) * self.weight
        return din + step - (gamma @ dn).reshape(din.shape)


def make_mixer(name, beta=0.4, history=8, precondition=None):
    """``mixing.py:263-286``: linear, anderson (Pulay over at most two iterations), pulay,
    broyden."""
    if name == "linear":
        return PulayMixer(beta, history, precondition, linear=True)
    if name == "anderson":
        return PulayMixer(beta, min(history, 2), precondition)
    if name == "pulay":
        return PulayMixer(beta, history, precondition)
    if name == "broyden":
        return BroydenMixer(beta, history, precondition)
    raise ValueError(f"unknown mixing algorithm {name!r}")


# ------------------------------------------------------------------------------ SCF
@dataclass
class SCFSettings:
    """The parts of amaterasu's Config the loop reads (defaults of ``config/schema.py:153-215``)."""

    max_steps: int = 40
    tol_energy: float = 1e-8  # eV/atom
    tol_density: float = 1e-7
    scheme: str = "fixed"
    smearing: float = 0.0  # Ha
    smearing_order: int = 1  # Methfessel-Paxton order
    mixing: str = "pulay"
    beta: float = 0.4
    history: int = 8
    kerker_q0: float = 1.5
    eigensolver: str = "davidson"


@dataclass
class Energies:
    """``driver.py:61-86``."""

    kinetic: float = 0.0
    local: float = 0.0
    hartree: float = 0.0
    xc: float = 0.0
    nonlocal_: float = 0.0
    ewald: float = 0.0
    entropy: float = 0.0

    @property
    def total(self):
        return self.kinetic + self.local + self.hartree + self.xc + self.nonlocal_ + self.ewald

    @property
    def free_energy(self):
        return self.total - self.entropy


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


def effective_potential(system, density):
    """``driver.py:225-252`` (norm-conserving): v_loc + V_H + v_xc[n + n_core]."""
    eh, vh = hartree(system.grid, system.density_sphere, density.sum(axis=0))
    xd = density if system.core_density is None else density + system.core_density[None] / system.nspin
    exc, vxc = system.xc.evaluate_on_grid(xd, system.grid)
    return np.ascontiguousarray(system.v_local[None] + vh[None] + vxc), eh, exc


def initial_wavefunctions(system, seed=12345):
    """``driver.py:255-264``: seed + 991 spin + ik."""
    return [[random_bands(len(s), system.nbands, s.kinetic_energy, seed=seed + 991 * sp + ik)
             for ik, s in enumerate(system.wave_spheres)] for sp in range(system.nspin)]


def run_scf(system, cfg: SCFSettings, *, wavefunctions=None, density=None, log=None):
    """``driver.py:267-464`` for norm-conserving datasets and mixing."""
    grid = system.grid
    natoms = len(system.symbols)
    density = system.initial_density() if density is None else np.array(density, dtype=float)
    wf = wavefunctions or initial_wavefunctions(system)
    mixer = make_mixer(cfg.mixing, cfg.beta, cfg.history, Kerker(grid, cfg.kerker_q0) if cfg.kerker_q0 > 0 else None)
    ethr = cfg.tol_energy * natoms * EV_IN_HARTREE
    hist, prev, resid, conv = [], np.inf, np.inf, False
    energies, occ = Energies(ewald=system.ewald), None
    eig = np.zeros((system.nspin, system.nkpoints, system.nbands))
    it = 0
    for it in range(1, cfg.max_steps + 1):
        pot, _, _ = effective_potential(system, density)
        ham = Hamiltonian(system, pot)
        tol = float(np.clip(1e-3 * min(resid, 1.0), 1e-11, 1e-6))
        steps = 40 if it == 1 else 20
        eig = np.empty((system.nspin, system.nkpoints, system.nbands))
        for sp in range(system.nspin):
            for ik, sphere in enumerate(system.wave_spheres):
                r = solve(cfg.eigensolver, lambda b, ik=ik, sp=sp, h=ham: h.apply(b, ik, sp), wf[sp][ik],
                          sphere.kinetic_energy, tolerance=tol, max_iterations=steps)
                eig[sp, ik] = r.eigenvalues
                wf[sp][ik] = r.coefficients
        occ = compute_occupations(eig, system.kpoints.weights, system.nelectrons, scheme=cfg.scheme,
                                  smearing=cfg.smearing, order=cfg.smearing_order, nspin=system.nspin)
        dout = compute_density(system, wf, occ.occupations)
        eh, _ = hartree(grid, system.density_sphere, dout.sum(axis=0))
        xd = dout if system.core_density is None else dout + system.core_density[None] / system.nspin
        exc, _ = system.xc.evaluate_on_grid(xd, grid)
        g = 2.0 / system.nspin
        kin = nl = 0.0
        for sp in range(system.nspin):
            for ik, sphere in enumerate(system.wave_spheres):
                f = g * system.kpoints.weights[ik] * occ.occupations[sp, ik]
                kin += float(f @ (np.abs(wf[sp][ik]) ** 2 @ sphere.kinetic_energy))
                nl += system.projectors[ik].energy(wf[sp][ik], f)
        energies = Energies(kinetic=kin, local=float(np.sum(system.v_local * dout.sum(axis=0)))
                            * grid.cell_volume_per_point, hartree=eh, xc=exc, nonlocal_=nl, ewald=system.ewald,
                            entropy=occ.entropy_term)
        band_sum = g * float(np.einsum("k,skn,skn->", system.kpoints.weights, occ.occupations, eig))
        identity = band_sum - (energies.kinetic + ham.local_energy(dout) + energies.nonlocal_)
        resid = float(np.sum(np.abs(dout - density))) * grid.cell_volume_per_point / system.nelectrons
        e = energies.free_energy
        delta = e - prev
        hist.append((e, resid, identity))
        if log:
            log(f"{it:4d} {e * HARTREE_IN_EV:18.9f} {delta * HARTREE_IN_EV / natoms:14.2e} {resid:13.3e} "
                f"{identity * HARTREE_IN_EV:12.2e}")
        if it > 1 and abs(delta) < ethr and resid < cfg.tol_density:
            conv = True
            break
        prev = e
        density = mixer(density, dout)
    return SCFResult(conv, it, energies, eig, occ.occupations, occ.fermi_level, occ.gap, dout, wf, hist,
                     electron_count(system, dout))
