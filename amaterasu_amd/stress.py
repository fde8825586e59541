"""Stress tensor (norm-conserving), on the GPU: ``compute_stress`` of
``amaterasu/properties/stress.py:63-290``.

sigma_uv = (1/Omega) dE/d eps_uv for the strain r -> (1 + eps) r, in Ha/Bohr^3, with the
plane-wave coefficients held on their Miller indices (exact at self-consistency):

* kinetic: -(1/Omega) sum w f |c|^2 (k+G)_u (k+G)_v (``stress.py:114-123``);
* Hartree: -E_H delta/Omega + sum_G 4 pi |n(G)|^2 G_u G_v / G^4 (``:126-137``);
* local: -E_loc delta/Omega - (1/Omega) sum_G Re[v'(|G|)/|G| S^*(G) n(G)] G_u G_v, with
  S the species structure factor (``:139-155``);
* exchange-correlation: delta (E_xc - int v_xc n)/Omega minus, for a GGA, the
  vsigma-weighted outer products of the spin-channel gradients (``:158-186``), from the
  same device XC evaluation the SCF uses;
* core: the rigid partial core, -(1/Omega) sum_G Re[V^*(G) rho_c'(|G|)/|G| S(G)] G_u G_v
  with V the spin-averaged v_xc (``:189-214``);
* nonlocal: -(R + R^T + E_nl delta)/Omega with
  R_{mu nu} = sum_n w_n Re sum_i conj((P D)_ni) sum_G c_nG q_nu conj(dP_i/dq_mu)
  (``:217-255``): per k-point one MFMA GEMM of the three q_nu-scaled copies of the band
  block (3 nb x npw) against the three projector q-gradients (3 nproj x npw), all nine
  (mu, nu) blocks at once, plus the projection GEMM; the q-gradients
  (``_projector_q_gradient``, ``:293-330``) are assembled on the device from the radial
  transforms and their q-derivatives (host quadrature once per |k+G| shell), the real
  harmonics and the solid-harmonic gradients;
* Ewald: the pair, reciprocal and background terms of ``hamiltonian/ewald.py:128-228``
  (host numpy);
* with symmetry, each term averaged as (1/N) sum_S M_S^T sigma M_S
  (``system/symmetry.py:310-324``).

Arguments are those of :func:`amaterasu_amd.forces.compute_forces`.
"""

from __future__ import annotations

import numpy as np
import torch
from scipy.special import erfc

from . import ops
from .forces import _datasets, _side, padded_block
from .hamiltonian import device_plan
from .pseudo import _solid_gradient, _ylm, radial_form_factor, radial_form_factor_derivative
from .system import density_form_factor_derivative, local_form_factor, local_form_factor_derivative
from .xc import XCFunctional


def ewald_stress(cell, positions, charges, eta=None, tolerance: float = 1e-14) -> np.ndarray:
    """Ion-ion stress (``hamiltonian/ewald.py:128-228``): pair separations strain as
    dr/d eps_uv = r_u r_v / r, reciprocal vectors as dG^2/d eps_uv = -2 G_u G_v with the
    structure factor invariant, and the reciprocal sum and background scale as 1/Omega."""
    cell = np.asarray(cell, dtype=float)
    pos = np.atleast_2d(np.asarray(positions, dtype=float))
    z = np.atleast_1d(np.asarray(charges, dtype=float))
    vol = abs(float(np.linalg.det(cell)))
    eta = 1.8 / vol ** (1.0 / 3.0) if eta is None else eta
    ztot = float(z.sum())
    out = np.zeros((3, 3))
    sep = pos[:, None, :] - pos[None, :, :]
    fr = sep @ np.linalg.inv(cell)
    sep = (fr - np.rint(fr)) @ cell
    recip = 2.0 * np.pi * np.linalg.inv(cell).T
    rc = 1.0 / eta
    while erfc(eta * rc) / rc >= tolerance / max(1.0, ztot**2):
        rc *= 1.2
    reps = np.ceil(rc * np.linalg.norm(recip, axis=1) / (2.0 * np.pi)).astype(int) + 1
    grid = np.stack(np.meshgrid(*[np.arange(-n, n + 1) for n in reps], indexing="ij"), -1).reshape(-1, 3)
    zz = z[:, None] * z[None, :]
    c = 2.0 * eta / np.sqrt(np.pi)
    for o in grid @ cell:
        v = sep + o
        d = np.linalg.norm(v, axis=-1)
        small = d < 1e-10
        r = np.where(small, 1.0, d)
        if np.min(r) > rc and np.max(np.abs(np.where(small, 0.0, zz * erfc(eta * r) / r))) < tolerance:
            continue
        dd = np.where(small, 0.0, -zz * (erfc(eta * r) / r**2 + c * np.exp(-((eta * r) ** 2)) / r))
        out += 0.5 * np.einsum("ab,abu,abv->uv", dd / r, v, v)
    gc = 2.0 * eta
    while 2.0 * np.pi / vol * max(ztot**2, 1.0) * np.exp(-(gc**2) / (4.0 * eta**2)) / gc**2 >= tolerance:
        gc *= 1.2
    b = np.ceil(gc * np.linalg.norm(cell, axis=1) / (2.0 * np.pi)).astype(int)
    m = np.stack(np.meshgrid(*[np.arange(-n, n + 1) for n in b], indexing="ij"), -1).reshape(-1, 3)
    g = m @ recip
    g2 = np.einsum("ij,ij->i", g, g)
    keep = (g2 > 1e-12) & (g2 <= gc**2)
    g, g2 = g[keep], g2[keep]
    s2 = np.abs(np.exp(1j * (g @ pos.T)) @ z) ** 2
    w = s2 * np.exp(-g2 / (4.0 * eta**2)) / g2
    out += 2.0 * np.pi / vol * ((w * (0.5 / eta**2 + 2.0 / g2))[:, None] * g).T @ g
    out -= np.eye(3) * (2.0 * np.pi / vol * float(w.sum()) - np.pi / (2.0 * vol * eta**2) * ztot**2)
    return out / vol


def _outer(weight: torch.Tensor, a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """sum_p weight_p a_pu b_pv for (n, 3) vector fields a, b (any leading shape)."""
    return (a.reshape(-1, 3) * weight.reshape(-1, 1)).T @ b.reshape(-1, 3)


def _dev(x) -> torch.Tensor:
    return torch.as_tensor(np.asarray(x) if not isinstance(x, torch.Tensor) else x, dtype=torch.float64).to("cuda")


def kinetic_stress(system, wavefunctions, occupations) -> np.ndarray:
    plan = device_plan(system)
    occ = np.asarray(occupations, dtype=float)
    w_k = np.asarray(system.kpoints.weights, dtype=float)
    out = torch.zeros((3, 3), dtype=torch.float64, device="cuda")
    for spin in range(system.nspin):
        for ik, sphere in enumerate(system.wave_spheres):
            n = plan.npw[ik]
            w = (2.0 / system.nspin) * w_k[ik] * occ[spin, ik]
            act = np.flatnonzero(w != 0.0)
            if act.size == 0:
                continue
            psi = padded_block(wavefunctions, spin, ik, act, n, plan.npw_pad)[:, :n]
            amp = torch.from_numpy(w[act]).to("cuda") @ (psi.real**2 + psi.imag**2)
            kg = torch.from_numpy(np.asarray(sphere.cartesian, dtype=float)).to("cuda")
            out -= _outer(amp, kg, kg)
    return out.cpu().numpy() / system.grid.volume


def _inverse(q: torch.Tensor) -> torch.Tensor:
    return torch.where(q > 1e-12, 1.0 / torch.where(q > 1e-12, q, 1.0), 0.0)


def hartree_stress(system, ng: torch.Tensor) -> np.ndarray:
    side = _side(system)
    g2 = (side.g**2).sum(-1)
    nz = g2 > 1e-12
    a2 = (ng.real**2 + ng.imag**2)[nz]
    e_h = 0.5 * system.grid.volume * float((4.0 * np.pi * a2 / g2[nz]).sum().item())
    t = _outer(4.0 * np.pi * a2 / g2[nz] ** 2, side.g[nz], side.g[nz])
    return -np.eye(3) * e_h / system.grid.volume + t.cpu().numpy()


def _species_phase(system, side) -> dict:
    """sum over the atoms of each species of e^{-iG.tau_a} on the density sphere."""
    out = {}
    for a, sym in enumerate(system.symbols):
        out[sym] = side.phase[a] if sym not in out else out[sym] + side.phase[a]
    return out


def local_stress(system, ng: torch.Tensor) -> np.ndarray:
    side = _side(system)
    ds = _datasets(system)
    qinv = _inverse(torch.sqrt((side.g**2).sum(-1)))
    energy = torch.zeros((), dtype=torch.float64, device="cuda")
    weight = torch.zeros_like(qinv)
    for sym, phase in _species_phase(system, side).items():
        sn = phase.conj() * ng
        energy = energy + (side.shells(local_form_factor(ds[sym], side.uq)) * sn).real.sum()
        weight = weight + (side.shells(local_form_factor_derivative(ds[sym], side.uq)) * qinv * sn).real
    t = -_outer(weight, side.g, side.g)
    return (t.cpu().numpy() - np.eye(3) * float(energy.item())) / system.grid.volume


def xc_core_stress(system, density) -> tuple[np.ndarray, np.ndarray | None]:
    """The exchange-correlation term and, with a partial core, the core term."""
    side = _side(system)
    d = _dev(density)
    core = getattr(system, "core_density", None)
    xd = d if core is None else d + _dev(core)[None] / d.shape[0]
    xc = system.xc if isinstance(system.xc, XCFunctional) else XCFunctional(getattr(system.xc, "name", system.xc))
    exc, vxc, grads, vsig = xc.evaluate_on_grid_details(xd, side.gops)
    dv = system.grid.volume / side.gops.size
    t = np.eye(3) * (exc - float((vxc * xd).sum().item()) * dv)
    if vsig is not None:
        gu, gd = grads
        suu, sud, sdd = vsig
        outer = 2.0 * _outer(suu, gu, gu) + _outer(sud, gu, gd) + _outer(sud, gd, gu) + 2.0 * _outer(sdd, gd, gd)
        t -= outer.cpu().numpy() * dv
    xc_t = t / system.grid.volume
    if core is None:
        return xc_t, None
    ds = _datasets(system)
    vg = side.coefficients(vxc.mean(0))
    qinv = _inverse(torch.sqrt((side.g**2).sum(-1)))
    weight = torch.zeros_like(qinv)
    for sym, phase in _species_phase(system, side).items():
        if ds[sym].core_density is None:
            continue
        slope = side.shells(density_form_factor_derivative(ds[sym], ds[sym].core_density, side.uq))
        weight = weight + (vg.conj() * slope * qinv * phase).real
    return xc_t, -_outer(weight, side.g, side.g).cpu().numpy() / system.grid.volume


def projector_q_gradient(system, ik: int, npw_pad: int | None = None) -> torch.Tensor:
    """dP_i/dq_alpha for every projector row at k-point ``ik``, (3, nproj, npw_pad) complex128
    on the device, rows in the order of ``build_projectors`` (``stress.py:293-330``):
    P = (4 pi/sqrt Omega) (-i)^l C_lm(q) f_l(q) q^-l e^{-iq.tau}, so
    dP/dq = pref (-i)^l [dC/dq f q^-l - l Y f q/q^2 + Y f' q/q] e^{-iq.tau}."""
    ds = _datasets(system)
    sphere = system.wave_spheres[ik]
    kg_h = np.asarray(sphere.cartesian, dtype=float)
    n = len(kg_h)
    npw_pad = n if npw_pad is None else npw_pad
    q_h = np.sqrt(np.asarray(sphere.g2, dtype=float))
    uq, inv = np.unique(q_h, return_inverse=True)
    inv_t = torch.from_numpy(inv.astype(np.int64)).to("cuda")
    kg = torch.from_numpy(kg_h).to("cuda")
    q = torch.from_numpy(q_h).to("cuda")
    qinv = _inverse(q)
    pos = torch.from_numpy(np.asarray(system.positions, dtype=float)).to("cuda")
    phase = torch.exp(-1j * (pos @ kg.T))  # [natoms][n]
    pref = 4.0 * np.pi / np.sqrt(system.grid.volume)
    # row offsets: atom-major, then channel, then m
    per_atom = [sum(2 * ch.l + 1 for ch in ds[s].projectors) for s in system.symbols]
    first = np.concatenate([[0], np.cumsum(per_atom)[:-1]]).astype(np.int64)
    out = torch.zeros((3, int(sum(per_atom)), npw_pad), dtype=torch.complex128, device="cuda")
    ylm, dcl = {}, {}
    for sym in dict.fromkeys(system.symbols):
        atoms = [a for a, s in enumerate(system.symbols) if s == sym]
        ph = phase[atoms]  # [na][n]
        off = 0
        for ch in ds[sym].projectors:
            l = ch.l
            if l not in ylm:
                ylm[l], dcl[l] = _ylm(l, kg), _solid_gradient(l, kg)
            f = torch.from_numpy(radial_form_factor(ds[sym], ch, uq)).to("cuda")[inv_t]
            fp = torch.from_numpy(radial_form_factor_derivative(ds[sym], ch, uq)).to("cuda")[inv_t]
            power = qinv**l if l else torch.ones_like(q)
            coef = pref * (-1j) ** l
            for mi in range(2 * l + 1):
                y = ylm[l][mi]
                ang = (dcl[l][mi] * (f * power)[:, None] - l * (y * f * qinv**2)[:, None] * kg
                       + (y * fp * qinv)[:, None] * kg)  # [n][3]
                rows = torch.from_numpy(first[atoms] + off + mi).to("cuda")
                out[:, rows, :n] = coef * ang.T.to(torch.complex128)[:, None, :] * ph[None, :, :]
            off += 2 * l + 1
    return out


def nonlocal_stress(system, wavefunctions, occupations) -> np.ndarray:
    plan = device_plan(system)
    kb = plan.kb_tables(1)
    if kb is None:
        return np.zeros((3, 3))
    mat, dij = kb[0], kb[1]
    occ = np.asarray(occupations, dtype=float)
    w_k = np.asarray(system.kpoints.weights, dtype=float)
    r = torch.zeros((3, 3), dtype=torch.float64, device="cuda")
    energy = torch.zeros((), dtype=torch.float64, device="cuda")
    for ik, sphere in enumerate(system.wave_spheres):
        n = plan.npw[ik]
        weights = [(2.0 / system.nspin) * w_k[ik] * occ[spin, ik] for spin in range(system.nspin)]
        if not any(np.any(w > 1e-12) for w in weights):
            continue
        grad = projector_q_gradient(system, ik, plan.npw_pad)  # [3][nproj][npw_pad]
        nproj = grad.shape[1]
        gflat = grad.reshape(3 * nproj, plan.npw_pad)
        kg = torch.zeros((plan.npw_pad, 3), dtype=torch.float64, device="cuda")
        kg[:n] = torch.from_numpy(np.asarray(sphere.cartesian, dtype=float)).to("cuda")
        for spin, w in enumerate(weights):
            act = np.flatnonzero(w > 1e-12)
            if act.size == 0:
                continue
            nb = len(act)
            psi = padded_block(wavefunctions, spin, ik, act, n, plan.npw_pad)
            p = ops.zgemm(psi, mat[ik], 1)  # [nb][nproj] = psi beta^H
            coupled = p @ dij[ik].to(torch.complex128)
            wt = torch.from_numpy(w[act]).to("cuda")
            energy = energy + (wt * (coupled * p.conj()).sum(1).real).sum()
            scaled = torch.cat([psi * kg[:, nu][None, :] for nu in range(3)])  # [3 nb][npw_pad]
            x = ops.zgemm(scaled, gflat, 1).view(3, nb, 3, nproj)  # [nu][n][mu][i]
            r -= torch.einsum("n,ni,vnmi->mv", wt.to(torch.complex128), coupled.conj(), x).real
        del grad, gflat
    r = r + r.T
    return (r.cpu().numpy() - np.eye(3) * float(energy.item())) / system.grid.volume


def symmetrize_tensor(tensor, system) -> np.ndarray:
    """(1/N) sum_S M_S^T sigma M_S with M_S = cell^-1 R^T cell (``symmetry.py:238-253,310-324``)."""
    tensor = np.asarray(tensor, dtype=float)
    sym = getattr(system, "symmetry", None)
    if sym is None or len(sym.rotations) <= 1:
        return tensor.copy()
    cell = np.asarray(system.cell, dtype=float)
    inv = np.linalg.inv(cell)
    acc = np.zeros((3, 3))
    for rot in sym.rotations:
        m = inv @ np.asarray(rot, dtype=float).T @ cell
        acc += m.T @ tensor @ m
    return acc / len(sym.rotations)


def compute_stress(system, density, wavefunctions, occupations, *, symmetrize: bool = True,
                   decompose: bool = False):
    """Stress in Ha/Bohr^3, shape (3, 3), or the per-term dict with ``decompose=True``
    (keys ewald, kinetic, hartree, local, nonlocal, xc and, with a partial core, core)."""
    ds = _datasets(system)
    side = _side(system)
    charges = np.array([ds[s].valence_charge for s in system.symbols])
    ng = side.coefficients(_dev(density).sum(0))
    parts = {"ewald": ewald_stress(system.cell, system.positions, charges),
             "kinetic": kinetic_stress(system, wavefunctions, occupations),
             "hartree": hartree_stress(system, ng),
             "local": local_stress(system, ng),
             "nonlocal": nonlocal_stress(system, wavefunctions, occupations)}
    parts["xc"], core = xc_core_stress(system, density)
    if core is not None:
        parts["core"] = core
    if symmetrize and getattr(system, "symmetry", None) is not None:
        parts = {k: symmetrize_tensor(v, system) for k, v in parts.items()}
    return parts if decompose else sum(parts.values())


__all__ = ["compute_stress", "ewald_stress", "hartree_stress", "kinetic_stress", "local_stress", "nonlocal_stress",
           "projector_q_gradient", "symmetrize_tensor", "xc_core_stress"]
