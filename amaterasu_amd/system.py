"""A device-side counterpart of amaterasu's ``PlaneWaveSystem`` for the H.psi path.

Holds the attributes the operator, the eigensolver and the density read
(``hamiltonian/system.py:62-94``): the FFT grid, one sphere per k-point, the projector
matrices (device tensors), the k-point weights, ``nspin`` and ``nbands``. Built from
explicit k-points (the reference's spglib reduction is outside this path); a P1 cell
needs no density symmetrisation (``system.py:198-199``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .basis import FFTGrid, GSphere
from .pseudo import build_projectors, read_psp8

DATA = Path(__file__).resolve().parent / "data"


def default_band_count(nelectrons: float, nspin: int = 1, moment: float = 0.0) -> int:
    """``hamiltonian/system.py:48-59``: 20 % extra, never fewer than 4 extra."""
    occ = 0.5 * nelectrons + (0.5 * abs(moment) if nspin == 2 else 0.0)
    return max(math.ceil(occ) + 4, math.ceil(1.2 * occ))


@dataclass
class KPoints:
    points: np.ndarray
    weights: np.ndarray

    def __len__(self) -> int:
        return len(self.points)


@dataclass
class DeviceSystem:
    cell: np.ndarray
    positions: np.ndarray
    symbols: list
    grid: FFTGrid
    wave_spheres: list
    projectors: list
    kpoints: KPoints
    nspin: int = 1
    nbands: int = 0
    nelectrons: float = 0.0
    symmetry: object = None
    paw: object = None
    extra: dict = field(default_factory=dict)
    ecut_rho: float = 0.0
    xc: object = None
    v_local: object = None
    core_density: object = None
    ewald: float = 0.0
    datasets: dict = field(default_factory=dict)

    def initial_density(self):
        """Superposed atomic valence densities (``hamiltonian/system.py:147-161``), on the device."""
        if "initial_density" not in self.extra:
            raise ValueError("attach_ionic_fields() has not been called for this system")
        return self.extra["initial_density"]

    @property
    def volume(self) -> float:
        return self.grid.volume

    @property
    def nkpoints(self) -> int:
        return len(self.kpoints)


def load_dataset(name: str = "Si.pbe.psp8.gz"):
    return read_psp8(DATA / name)


def build_system(cell, positions, symbols, datasets: dict, ecut_rho: float, kpoints, ecuts, *, weights=None,
                 nbands=None, nspin: int = 1, with_projectors: bool = True, grid_shape=None) -> DeviceSystem:
    """Grid from the density cutoff (``system.py:262-271``), one sphere per k at its own
    cutoff, projector matrices on the device."""
    cell = np.asarray(cell, dtype=float)
    kpts = np.atleast_2d(np.asarray(kpoints, dtype=float))
    ecuts = np.broadcast_to(np.asarray(ecuts, dtype=float), (len(kpts),))
    grid = FFTGrid.for_cutoff(cell, ecut_rho) if grid_shape is None else FFTGrid(tuple(grid_shape), cell)
    spheres = [GSphere.build(cell, float(e), kpoint=k) for e, k in zip(ecuts, kpts)]
    vol = grid.volume
    projs = [build_projectors(s, positions, symbols, datasets, vol) for s in spheres] if with_projectors else None
    w = np.full(len(kpts), 1.0 / len(kpts)) if weights is None else np.asarray(weights, dtype=float)
    nel = float(sum(datasets[s].valence_charge for s in symbols))
    nb = nbands or default_band_count(nel, nspin)
    return DeviceSystem(cell=cell, positions=np.asarray(positions, dtype=float), symbols=list(symbols), grid=grid,
                        wave_spheres=spheres, projectors=projs, kpoints=KPoints(kpts, w), nspin=nspin, nbands=nb,
                        nelectrons=nel, ecut_rho=float(ecut_rho))




# --------------------------------------------------------------------------- ionic fields
def _radial_on_shells(fn, g2: np.ndarray):
    """Evaluate a radial form factor once per distinct |G| and expand to every G."""
    q = np.sqrt(np.asarray(g2, dtype=float))
    uq, inv = np.unique(q, return_inverse=True)
    return fn(uq)[inv]


def local_form_factor(ds, q):
    """v_loc(q) with the erf-screened Coulomb split and the alpha term at q = 0
    (``pseudo/formfactors.py:74-103``)."""
    from scipy.special import erf

    q = np.atleast_1d(np.asarray(q, dtype=float))
    r, z = ds.r, float(ds.valence_charge)
    out = np.empty_like(q)
    zero = q < 1e-12
    if np.any(zero):
        out[zero] = 4.0 * np.pi * ds.integrate(r * (r * ds.local_potential + z))
    if np.any(~zero):
        qq = q[~zero]
        scr = r * ds.local_potential + z * erf(r)
        out[~zero] = 4.0 * np.pi / qq * ds.integrate(scr[None, :] * np.sin(qq[:, None] * r[None, :])) \
            - 4.0 * np.pi * z * np.exp(-0.25 * qq**2) / qq**2
    return out


def density_form_factor(ds, values, q):
    """int rho(r) j_0(q r) dr for a density stored as 4 pi r^2 rho (``formfactors.py:117-123``)."""
    from scipy.special import spherical_jn

    q = np.atleast_1d(np.asarray(q, dtype=float))
    return ds.integrate(np.asarray(values)[None, :] * spherical_jn(0, q[:, None] * ds.r[None, :]))


def local_form_factor_derivative(ds, q):
    """d v_loc/dq of both pieces of the erf split, zero at q = 0 where v_loc depends on q^2
    only (``pseudo/formfactors.py:160-185``). With s(r) = r V_loc + Z erf(r):
    d/dq [(4 pi/q) int s sin(qr) dr] = -(4 pi/q^2) int s sin + (4 pi/q) int s r cos(qr) dr,
    d/dq [-4 pi Z e^{-q^2/4}/q^2] = 4 pi Z e^{-q^2/4} (1/(2q) + 2/q^3)."""
    from scipy.special import erf

    q = np.atleast_1d(np.asarray(q, dtype=float))
    r, z = ds.r, float(ds.valence_charge)
    out = np.zeros_like(q)
    nz = q >= 1e-12
    if np.any(nz):
        qq = q[nz]
        scr = r * ds.local_potential + z * erf(r)
        x = qq[:, None] * r[None, :]
        out[nz] = (-4.0 * np.pi / qq**2 * ds.integrate(scr[None, :] * np.sin(x))
                   + 4.0 * np.pi / qq * ds.integrate((scr * r)[None, :] * np.cos(x))
                   + 4.0 * np.pi * z * np.exp(-0.25 * qq**2) * (0.5 / qq + 2.0 / qq**3))
    return out


def density_form_factor_derivative(ds, values, q):
    """d/dq int rho(r) j_0(q r) dr = int rho r j_0'(q r) dr (``formfactors.py:133-150``)."""
    from scipy.special import spherical_jn

    q = np.atleast_1d(np.asarray(q, dtype=float))
    return ds.integrate((np.asarray(values) * ds.r)[None, :] * spherical_jn(0, q[:, None] * ds.r[None, :],
                                                                             derivative=True))


def superpose(grid, ecut_rho, positions, symbols, radial_of_symbol):
    """F(r) = sum_G [(1/Omega) sum_a f_a(|G|) e^{-i G.tau_a}] e^{iG.r} over the density sphere,
    assembled on the device (``hamiltonian/local.py:42-55``); returns the real part (FP64)."""
    import torch

    from .basis import GSphere
    from .xc import GridOps

    sphere = GSphere.build(grid.cell, ecut_rho)
    gops = GridOps(grid.shape, grid.cell)
    coef = torch.zeros(len(sphere), dtype=torch.complex128, device="cuda")
    cart = torch.from_numpy(sphere.cartesian).to("cuda")
    pos = torch.from_numpy(np.asarray(positions, dtype=float)).to("cuda")
    radial = {s: torch.from_numpy(_radial_on_shells(f, sphere.g2)).to("cuda") for s, f in radial_of_symbol.items()}
    for atom, sym in enumerate(symbols):
        coef += radial[sym] * torch.exp(-1j * (cart * pos[atom]).sum(1))
    coef /= grid.volume
    shape = np.array(grid.shape)
    flat = torch.from_numpy(np.ravel_multi_index(tuple((sphere.miller % shape).T), tuple(shape))).to("cuda")
    g = torch.zeros(int(np.prod(shape)), dtype=torch.complex128, device="cuda")
    g[flat] = coef
    field = gops.to_real(g.reshape(tuple(shape)))
    scale = float(field.real.abs().max().item())
    if scale > 0 and float(field.imag.abs().max().item()) > 1e-8 * scale:
        raise ValueError("a superposed atomic field came out complex; the sphere map or a phase is wrong")
    return field.real.contiguous()


def ewald_energy(cell, positions, charges, eta=None, tolerance=1e-14) -> float:
    """Ion-ion energy: erfc real-space sum over folded separations, Gaussian reciprocal sum,
    self energy and neutralising background (``hamiltonian/ewald.py:94-225``)."""
    from scipy.special import erfc

    cell = np.asarray(cell, dtype=float)
    pos = np.atleast_2d(np.asarray(positions, dtype=float))
    z = np.asarray(charges, dtype=float)
    vol = abs(float(np.linalg.det(cell)))
    eta = 1.8 / vol ** (1.0 / 3.0) if eta is None else eta
    ztot = float(z.sum())
    rc = 1.0 / eta
    while erfc(eta * rc) / rc >= tolerance / max(1.0, ztot**2):
        rc *= 1.2
    sep = pos[:, None, :] - pos[None, :, :]
    fr = sep @ np.linalg.inv(cell)
    sep = (fr - np.rint(fr)) @ cell
    recip = 2.0 * np.pi * np.linalg.inv(cell).T
    reps = np.ceil(rc * np.linalg.norm(recip, axis=1) / (2.0 * np.pi)).astype(int) + 1
    offs = np.stack(np.meshgrid(*[np.arange(-n, n + 1) for n in reps], indexing="ij"), -1).reshape(-1, 3) @ cell
    zz = z[:, None] * z[None, :]
    real = 0.0
    for o in offs:
        d = np.linalg.norm(sep + o, axis=-1)
        small = d < 1e-10
        safe = np.where(small, 1.0, d)
        pair = np.where(small, 0.0, zz * erfc(eta * safe) / safe)
        if np.max(np.abs(pair)) < tolerance and np.min(safe) > rc:
            continue
        real += float(pair.sum())
    gc = 2.0 * eta
    while 2.0 * np.pi / vol * max(ztot**2, 1.0) * np.exp(-(gc**2) / (4.0 * eta**2)) / gc**2 >= tolerance:
        gc *= 1.2
    b = np.ceil(gc * np.linalg.norm(cell, axis=1) / (2.0 * np.pi)).astype(int)
    m = np.stack(np.meshgrid(*[np.arange(-n, n + 1) for n in b], indexing="ij"), -1).reshape(-1, 3)
    g = m @ recip
    g2 = np.einsum("ij,ij->i", g, g)
    keep = (g2 > 1e-12) & (g2 <= gc**2)
    g, g2 = g[keep], g2[keep]
    s = np.exp(1j * (g @ pos.T)) @ z
    rsum = float(2.0 * np.pi / vol * np.sum(np.abs(s) ** 2 * np.exp(-g2 / (4.0 * eta**2)) / g2))
    return 0.5 * real + rsum - eta / np.sqrt(np.pi) * float(np.sum(z**2)) - np.pi / (2.0 * vol * eta**2) * ztot**2


def attach_ionic_fields(system: DeviceSystem, datasets: dict, ecut_rho: float, xc: str = "PBE",
                        magmoms=None) -> DeviceSystem:
    """Local potential, partial core, superposed valence start and Ewald energy for an SCF
    (``hamiltonian/system.py:337-378`` norm-conserving branch, ``local.py:58-143``). With
    ``nspin == 2`` the start is split as (n +- m)/2, m the per-atom valence densities scaled by
    magmom / (their charge) and clipped to |m| <= n (``system.py:147-190``)."""
    import torch

    syms = system.symbols
    system.datasets = dict(datasets)
    system.ecut_rho = float(ecut_rho)
    system.xc = xc
    system.v_local = superpose(system.grid, ecut_rho, system.positions, syms,
                               {s: (lambda q, d=d: local_form_factor(d, q)) for s, d in datasets.items()})
    if any(d.core_density is not None for d in datasets.values()):
        system.core_density = superpose(
            system.grid, ecut_rho, system.positions, syms,
            {s: (lambda q, d=d: density_form_factor(d, d.core_density, q) if d.core_density is not None
                 else np.zeros_like(q)) for s, d in datasets.items()})
    else:
        system.core_density = None
    vradial = {s: (lambda q, d=d: density_form_factor(d, d.valence_density, q)) for s, d in datasets.items()}
    valence = superpose(system.grid, ecut_rho, system.positions, syms, vradial).clamp(min=1e-12)
    dv = system.grid.cell_volume_per_point
    valence *= system.nelectrons / (float(valence.sum().item()) * dv)
    system.extra["initial_magmoms"] = np.zeros(len(syms)) if magmoms is None else np.asarray(magmoms, dtype=float)
    if system.nspin == 1:
        system.extra["initial_density"] = valence[None].contiguous()
    else:
        mom = np.zeros(len(syms)) if magmoms is None else np.asarray(magmoms, dtype=float)
        m = torch.zeros_like(valence)
        for atom, sym in enumerate(syms):
            if abs(mom[atom]) < 1e-12:
                continue
            single = superpose(system.grid, ecut_rho, system.positions[atom:atom + 1], [sym],
                               {sym: vradial[sym]}).clamp(min=1e-12)
            charge = float(single.sum().item()) * dv
            if charge > 0:
                m += (mom[atom] / charge) * single
        m = torch.maximum(torch.minimum(m, valence), -valence)
        system.extra["initial_density"] = torch.stack([0.5 * (valence + m), 0.5 * (valence - m)]).contiguous()
    system.ewald = ewald_energy(system.cell, system.positions, [datasets[s].valence_charge for s in syms])
    return system


__all__ = ["DeviceSystem", "KPoints", "attach_ionic_fields", "build_system", "default_band_count", "ewald_energy",
           "load_dataset", "superpose"]
