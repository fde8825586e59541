"""Forces on the ions (norm-conserving), on the GPU: ``compute_forces`` of
``amaterasu/properties/forces.py:48-176``.

F = F_ewald + F_local + F_nonlocal (+ F_core with a partial core), in Ha/Bohr:

* local: ``F_a = -sum_G i G v_a(|G|) e^{iG.tau_a} n(G)`` over the density sphere
  (``forces.py:111-126``), n(G) from the library FFT, the sum on the device;
* core: the partial core moving with its atom against the spin-averaged v_xc of
  n + n_core/nspin (``forces.py:96-105,129-143``);
* nonlocal: ``-sum_nk w f 2 Re[(P D)^* dP/dtau]`` with ``P = psi beta^H`` and
  ``dP/dtau_x = (psi . i(k+G)_x) beta^H`` (``forces.py:146-175``): four MFMA GEMMs per
  k-point (project, and one per Cartesian axis) on the plan's projector matrices;
* Ewald: the pair and structure-factor derivatives of ``hamiltonian/ewald.py:161-206``
  (host numpy, O(natoms^2) per lattice vector);
* with symmetry, each term averaged over the space group (``system/symmetry.py:288-307``).

``density`` is (nspin, n1, n2, n3); ``wavefunctions`` either the reference's nested list
``[spin][k]`` of (nbands, npw_k) blocks (numpy or device) or the padded device block
``[nspin][nk][nbands][npw_pad]``; ``occupations`` (nspin, nk, nbands).
"""

from __future__ import annotations

import numpy as np
import torch
from scipy.special import erfc

from . import ops
from .basis import GSphere
from .hamiltonian import device_plan
from .system import density_form_factor, local_form_factor
from .xc import GridOps, XCFunctional


def ewald_forces(cell, positions, charges, eta=None, tolerance: float = 1e-14) -> np.ndarray:
    """Ion-ion forces: erfc pair derivatives over folded separations plus the Gaussian
    reciprocal sum (``hamiltonian/ewald.py:161-206``)."""
    cell = np.asarray(cell, dtype=float)
    pos = np.atleast_2d(np.asarray(positions, dtype=float))
    z = np.atleast_1d(np.asarray(charges, dtype=float))
    vol = abs(float(np.linalg.det(cell)))
    eta = 1.8 / vol ** (1.0 / 3.0) if eta is None else eta
    ztot = float(z.sum())
    out = np.zeros((len(z), 3))
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
        out -= np.einsum("ab,abi->ai", dd / r, v)
    gc = 2.0 * eta
    while 2.0 * np.pi / vol * max(ztot**2, 1.0) * np.exp(-(gc**2) / (4.0 * eta**2)) / gc**2 >= tolerance:
        gc *= 1.2
    b = np.ceil(gc * np.linalg.norm(cell, axis=1) / (2.0 * np.pi)).astype(int)
    m = np.stack(np.meshgrid(*[np.arange(-n, n + 1) for n in b], indexing="ij"), -1).reshape(-1, 3)
    g = m @ recip
    g2 = np.einsum("ij,ij->i", g, g)
    keep = (g2 > 1e-12) & (g2 <= gc**2)
    g, g2 = g[keep], g2[keep]
    ph = np.exp(1j * (g @ pos.T))
    s = ph @ z
    w = np.exp(-g2 / (4.0 * eta**2)) / g2
    out += 4.0 * np.pi / vol * z[:, None] * ((w[:, None] * np.imag(s.conj()[:, None] * ph)).T @ g)
    return out


class _DensitySide:
    """Density sphere, its grid indices and structure factors on the device (cached per system)."""

    def __init__(self, system):
        ecut_rho = getattr(system, "ecut_rho", 0.0)
        sphere = getattr(system, "density_sphere", None) or GSphere.build(system.cell, ecut_rho)
        shape = np.array(system.grid.shape)
        self.sphere = sphere
        self.flat = torch.from_numpy(np.ravel_multi_index(tuple((sphere.miller % shape).T), tuple(shape))).to("cuda")
        self.g = torch.from_numpy(np.asarray(sphere.cartesian, dtype=float)).to("cuda")
        pos = torch.from_numpy(np.asarray(system.positions, dtype=float)).to("cuda")
        self.phase = torch.exp(-1j * (pos @ self.g.T))  # [natoms][ng] = e^{-iG.tau}
        q = np.sqrt(np.asarray(sphere.g2, dtype=float))
        self.uq, inv = np.unique(q, return_inverse=True)
        self.inv = torch.from_numpy(inv.astype(np.int64)).to("cuda")
        self.gops = GridOps(system.grid.shape, system.cell)

    def shells(self, values: np.ndarray) -> torch.Tensor:
        return torch.from_numpy(np.asarray(values, dtype=float)).to("cuda")[self.inv]

    def coefficients(self, field: torch.Tensor) -> torch.Tensor:
        """field(G) on the density sphere (the reference's gather of to_reciprocal)."""
        return self.gops.to_reciprocal(field).reshape(-1)[self.flat]


def _side(system) -> _DensitySide:
    extra = getattr(system, "extra", None)
    if isinstance(extra, dict) and "_force_side" in extra:
        return extra["_force_side"]
    side = _DensitySide(system)
    if isinstance(extra, dict):
        extra["_force_side"] = side
    return side


def _datasets(system):
    ds = getattr(system, "datasets", None) or getattr(system, "pseudopotentials", None)
    if ds is None:
        raise ValueError("forces need the system's pseudopotential datasets (system.datasets)")
    return ds


def _contract_g(w: torch.Tensor, side: _DensitySide) -> torch.Tensor:
    """sum_G w[a, G] G for all atoms at once: one split-K MFMA product [natoms][ng] x [ng][3]
    (a matrix-vector product per atom sends the long contraction to a single workgroup)."""
    gt = side.g.T.to(torch.complex128).contiguous()  # [3][ng], real, so conj(gt)^T = G
    return ops.zgemm(w.contiguous(), gt, 1)


def local_forces(system, density) -> np.ndarray:
    side = _side(system)
    ds = _datasets(system)
    total = torch.as_tensor(np.asarray(density) if not isinstance(density, torch.Tensor) else density,
                            dtype=torch.float64).to("cuda").sum(0)
    ng = side.coefficients(total)
    ff = {s: side.shells(local_form_factor(d, side.uq)) for s, d in ds.items()}
    w = torch.stack([ff[sym] for sym in system.symbols]) * side.phase.conj() * (1j * ng)[None, :]
    return _real(-_contract_g(w, side), "local force")


def core_forces(system, density) -> np.ndarray:
    side = _side(system)
    ds = _datasets(system)
    core = getattr(system, "core_density", None)
    if core is None:
        return np.zeros((len(system.symbols), 3))
    d = torch.as_tensor(np.asarray(density) if not isinstance(density, torch.Tensor) else density,
                        dtype=torch.float64).to("cuda")
    core_t = torch.as_tensor(np.asarray(core) if not isinstance(core, torch.Tensor) else core,
                             dtype=torch.float64).to("cuda")
    xc = system.xc if isinstance(system.xc, XCFunctional) else XCFunctional(getattr(system.xc, "name", system.xc))
    _, vxc = xc.evaluate_on_grid(d + core_t[None] / d.shape[0], side.gops)
    vg = side.coefficients(vxc.mean(0))
    zero = torch.zeros_like(side.g[:, 0])
    cg = {s: zero if d.core_density is None else side.shells(density_form_factor(d, d.core_density, side.uq))
          for s, d in ds.items()}
    w = torch.stack([cg[sym] for sym in system.symbols]) * side.phase * (1j * vg.conj())[None, :]
    return _real(_contract_g(w, side), "core force")


def padded_block(wavefunctions, spin: int, ik: int, rows, n: int, npw_pad: int) -> torch.Tensor:
    """Bands ``rows`` of (spin, k) as a zero-padded FP64 device block (len(rows), npw_pad), from
    the nested ``[spin][k]`` list (numpy or device) or the padded [nspin][nk][nb][npw_pad] block."""
    if isinstance(wavefunctions, torch.Tensor):
        blk = wavefunctions[spin, ik][torch.as_tensor(rows, device=wavefunctions.device)]
    else:
        b = wavefunctions[spin][ik]
        b = b if isinstance(b, torch.Tensor) else torch.from_numpy(np.asarray(b))
        blk = b[torch.as_tensor(rows, device=b.device)]
    psi = torch.zeros((len(rows), npw_pad), dtype=torch.complex128, device="cuda")
    psi[:, :n] = blk[:, :n].to("cuda", torch.complex128)
    return psi


def nonlocal_forces(system, wavefunctions, occupations) -> np.ndarray:
    plan = device_plan(system)
    kb = plan.kb_tables(1)
    natoms = len(system.symbols)
    if kb is None:
        return np.zeros((natoms, 3))
    mat, dij = kb[0], kb[1]
    labels = system.projectors[0].labels
    atom_of_row = torch.as_tensor([lab[0] for lab in labels], dtype=torch.long, device="cuda")
    nspin = system.nspin
    occ = np.asarray(occupations, dtype=float)
    w_k = np.asarray(system.kpoints.weights, dtype=float)
    out = torch.zeros((natoms, 3), dtype=torch.float64, device="cuda")
    for spin in range(nspin):
        for ik, sphere in enumerate(system.wave_spheres):
            n = plan.npw[ik]
            w = (2.0 / nspin) * w_k[ik] * occ[spin, ik]
            act = np.flatnonzero(w > 1e-12)
            if act.size == 0:
                continue
            psi = padded_block(wavefunctions, spin, ik, act, n, plan.npw_pad)
            beta = mat[ik]
            p = ops.zgemm(psi, beta, 1)  # [nb][nproj] = psi beta^H
            coupled = p @ dij[ik].to(torch.complex128)
            kg = torch.zeros((plan.npw_pad, 3), dtype=torch.float64, device="cuda")  # k+G, zero padding
            kg[:n] = torch.from_numpy(np.asarray(sphere.cartesian, dtype=float)).to("cuda")
            wt = torch.from_numpy(w[act]).to("cuda")
            for axis in range(3):
                der = ops.zgemm((psi * (1j * kg[:, axis])[None, :]).contiguous(), beta, 1)
                terms = 2.0 * (coupled.conj() * der).real  # [nb][nproj]
                per = wt @ terms
                out[:, axis].index_add_(0, atom_of_row, -per)
    return out.cpu().numpy()


def symmetrize_vectors(vectors, system) -> np.ndarray:
    """``symmetry.py:288-307``: (1/N) sum_S F_{S(a)} M_S^-1 with M_S = cell^-1 R^T cell."""
    vectors = np.asarray(vectors, dtype=float)
    sym = getattr(system, "symmetry", None)
    if sym is None or len(sym.rotations) <= 1:
        return vectors.copy()
    cell = np.asarray(system.cell, dtype=float)
    frac = np.asarray(system.positions, dtype=float) @ np.linalg.inv(cell)
    species = list(system.symbols)
    same = np.array([[a == b for b in species] for a in species])
    out = np.zeros_like(vectors)
    inv = np.linalg.inv(cell)
    for r, t in zip(sym.rotations, sym.translations):
        r = np.asarray(r, dtype=float)
        mat = inv @ r.T @ cell
        images = frac @ r.T + t
        diff = images[:, None, :] - frac[None, :, :]
        diff -= np.rint(diff)
        dist = np.where(same, np.linalg.norm(diff, axis=-1), np.inf)
        out += vectors[np.argmin(dist, axis=1)] @ np.linalg.inv(mat)
    return out / len(sym.rotations)


def _real(v: torch.Tensor, what: str) -> np.ndarray:
    """``forces.py:178-186``: the imaginary part checks the sphere bookkeeping."""
    v = v.cpu().numpy()
    resid = float(np.max(np.abs(v.imag))) if v.size else 0.0
    scale = max(float(np.max(np.abs(v.real))) if v.size else 0.0, 1e-30)
    if resid > 1e-6 * scale and resid > 1e-12:
        raise ValueError(f"the {what} came out complex (imaginary part {resid:.3e} against {scale:.3e})")
    return v.real.copy()


def compute_forces(system, density, wavefunctions, occupations, *, symmetrize: bool = True,
                   decompose: bool = False):
    """Forces in Ha/Bohr, shape (natoms, 3), or the per-term dict with ``decompose=True``."""
    ds = _datasets(system)
    charges = np.array([ds[s].valence_charge for s in system.symbols])
    parts = {"ewald": ewald_forces(system.cell, system.positions, charges),
             "local": local_forces(system, density),
             "nonlocal": nonlocal_forces(system, wavefunctions, occupations)}
    if getattr(system, "core_density", None) is not None:
        parts["core"] = core_forces(system, density)
    if symmetrize and getattr(system, "symmetry", None) is not None:
        parts = {k: symmetrize_vectors(v, system) for k, v in parts.items()}
    return parts if decompose else sum(parts.values())


__all__ = ["compute_forces", "core_forces", "ewald_forces", "local_forces", "nonlocal_forces", "symmetrize_vectors"]
