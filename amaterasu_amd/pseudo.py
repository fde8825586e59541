"""Norm-conserving datasets and Kleinman-Bylander projectors, assembled on the GPU.

``read_psp8`` parses the ABINIT psp8 format the reference reads in
``pseudo/readers/psp8.py:90-220`` (projectors stored as r*beta, densities as 4 pi rho,
multiplied by r^2 on load). ``build_projectors`` produces the matrix of
``pseudo/kb.py:260-343``:

    <k+G|beta_{a,l,n,m}> = (4 pi / sqrt(Omega)) (-i)^l Y_lm(k+G) f_ln(|k+G|) exp(-i (k+G).tau_a)

with rows atom-major, then radial channel, then m = -l..l (real Y_lm, l = 1 ordered
y, z, x) and D block-diagonal per atom. The radial transform f_l(q) = int r^2 beta j_l(qr) dr
is evaluated once per distinct |k+G| on the host with the same Simpson quadrature as
the reference (``pseudo/base.py:130-144``); the (nproj x npw) matrix -- 0.7 GB in FP64
for 64 silicon atoms -- is assembled directly in device memory.
"""

from __future__ import annotations

import gzip
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import torch
from scipy.integrate import simpson
from scipy.special import spherical_jn

from . import ops

MAX_L = 3
_XC_NAMES = {2: "LDA-PZ", 7: "LDA-PW", 11: "PBE", -116133: "PBEsol", -101130: "PBE", -1012: "LDA-PW"}


@dataclass(frozen=True)
class Channel:
    l: int
    index: int
    energy: float
    values: np.ndarray  # r beta_l(r)


@dataclass
class Dataset:
    """What the operator needs from a norm-conserving pseudopotential file."""

    symbol: str
    atomic_number: float
    valence_charge: float
    r: np.ndarray
    local_potential: np.ndarray
    projectors: list
    xc_functional: str | None = None
    core_density: np.ndarray | None = None  # 4 pi r^2 rho_core
    valence_density: np.ndarray | None = None  # 4 pi r^2 rho_valence
    dij: np.ndarray | None = None
    extra: dict = field(default_factory=dict)

    def integrate(self, values) -> np.ndarray:
        return simpson(np.asarray(values, dtype=float), x=self.r, axis=-1)


def _num(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def read_psp8(path) -> Dataset:
    path = Path(path)
    opener = gzip.open if path.suffix == ".gz" else open
    with opener(path, "rt", encoding="utf-8", errors="replace") as fh:
        lines = fh.read().splitlines()
    if len(lines) < 7:
        raise ValueError(f"{path.name}: too short to be a psp8 file")
    z, zion = _num(lines[1].split()[0]), _num(lines[1].split()[1])
    code, ixc, lmax, lloc, mmax = (int(t) for t in lines[2].split()[:5])
    if code != 8:
        raise ValueError(f"{path.name}: pspcod {code}, not 8")
    fchrg = _num(lines[3].split()[1])
    counts = [int(t) for t in lines[4].split()[:5]]
    switch = int(lines[5].split()[0])
    if switch in (2, 3):
        raise NotImplementedError(f"{path.name}: spin-orbit psp8 files are not supported")

    def table(start):
        rows = [[_num(t) for t in lines[start + i].split()[1:]] for i in range(mmax)]
        return np.asarray(rows)

    row = 6
    chans = []
    for l in range(lmax + 1):
        if l == lloc or l >= len(counts) or counts[l] == 0:
            continue
        head = lines[row].split()
        if int(head[0]) != l:
            raise ValueError(f"{path.name}: expected l={l} at line {row + 1}")
        ekb = [_num(t) for t in head[1 : 1 + counts[l]]]
        tab = table(row + 1)
        chans += [Channel(l, j, ekb[j], np.ascontiguousarray(tab[:, 1 + j])) for j in range(counts[l])]
        row += 1 + mmax
    if int(lines[row].split()[0]) != lloc:
        raise ValueError(f"{path.name}: expected the local block at line {row + 1}")
    tab = table(row + 1)
    r, vloc = np.ascontiguousarray(tab[:, 0]), np.ascontiguousarray(tab[:, 1])
    row += 1 + mmax
    core = val = None
    if fchrg > 0:
        core = table(row)[:, 1] * r**2
        row += mmax
    if switch in (1, 3):
        val = table(row)[:, 1] * r**2
        row += mmax
    return Dataset(symbol=lines[0].split()[0], atomic_number=z, valence_charge=zion, r=r, local_potential=vloc,
                   projectors=chans, xc_functional=_XC_NAMES.get(ixc), core_density=core, valence_density=val,
                   extra={"pspxc": ixc, "fchrg": fchrg})


def _ylm(l: int, v: torch.Tensor) -> torch.Tensor:
    """Real Y_lm, m = -l..l, of directions v (n, 3); the zero vector maps to +z."""
    n = torch.linalg.vector_norm(v, dim=1)
    u = torch.zeros_like(v)
    u[:, 2] = 1.0
    nz = n > 1e-14
    u[nz] = v[nz] / n[nz, None]
    x, y, z = u[:, 0], u[:, 1], u[:, 2]
    pi = np.pi
    if l == 0:
        return torch.full((1, v.shape[0]), 0.5 / np.sqrt(pi), dtype=v.dtype, device=v.device)
    if l == 1:
        return np.sqrt(0.75 / pi) * torch.stack([y, z, x])
    if l == 2:
        c = 0.5 * np.sqrt(15.0 / pi)
        return torch.stack([c * x * y, c * y * z, 0.25 * np.sqrt(5.0 / pi) * (3.0 * z * z - 1.0), c * x * z,
                            0.25 * np.sqrt(15.0 / pi) * (x * x - y * y)])
    if l == 3:
        return torch.stack([
            0.25 * np.sqrt(17.5 / pi) * y * (3.0 * x * x - y * y),
            0.5 * np.sqrt(105.0 / pi) * x * y * z,
            0.25 * np.sqrt(10.5 / pi) * y * (5.0 * z * z - 1.0),
            0.25 * np.sqrt(7.0 / pi) * z * (5.0 * z * z - 3.0),
            0.25 * np.sqrt(10.5 / pi) * x * (5.0 * z * z - 1.0),
            0.25 * np.sqrt(105.0 / pi) * z * (x * x - y * y),
            0.25 * np.sqrt(17.5 / pi) * x * (x * x - 3.0 * y * y),
        ])
    raise NotImplementedError(f"real spherical harmonics are implemented up to l={MAX_L}")


def radial_form_factor(ds: Dataset, channel: Channel, q: np.ndarray) -> np.ndarray:
    """f_l(q) = int r^2 beta_l(r) j_l(q r) dr (Simpson on the dataset mesh)."""
    q = np.atleast_1d(np.asarray(q, dtype=float))
    return ds.integrate((ds.r * channel.values)[None, :] * spherical_jn(channel.l, q[:, None] * ds.r[None, :]))


def radial_form_factor_derivative(ds: Dataset, channel: Channel, q: np.ndarray) -> np.ndarray:
    """f_l'(q) = int r^3 beta_l(r) j_l'(q r) dr, the Bessel function differentiated under the
    integral (``pseudo/formfactors.py:133-157``)."""
    q = np.atleast_1d(np.asarray(q, dtype=float))
    x = q[:, None] * ds.r[None, :]
    return ds.integrate((ds.r**2 * channel.values)[None, :] * spherical_jn(channel.l, x, derivative=True))


#: Real solid harmonics C_lm(q) = |q|^l Y_lm(q_hat) as sums of monomials
#: coefficient * x^a y^b z^c, listed as (coefficient, a, b, c), m = -l..l in the order of
#: ``_ylm`` (``pseudo/kb.py:98-139``).
def _solid_monomials():
    pi = np.pi
    k1 = np.sqrt(0.75 / pi)
    k2, k20, k22 = 0.5 * np.sqrt(15.0 / pi), 0.25 * np.sqrt(5.0 / pi), 0.25 * np.sqrt(15.0 / pi)
    k33, k32, k31 = 0.25 * np.sqrt(17.5 / pi), 0.5 * np.sqrt(105.0 / pi), 0.25 * np.sqrt(10.5 / pi)
    k30, k32c = 0.25 * np.sqrt(7.0 / pi), 0.25 * np.sqrt(105.0 / pi)
    return {
        0: [[(0.5 / np.sqrt(pi), 0, 0, 0)]],
        1: [[(k1, 0, 1, 0)], [(k1, 0, 0, 1)], [(k1, 1, 0, 0)]],
        2: [[(k2, 1, 1, 0)], [(k2, 0, 1, 1)], [(2 * k20, 0, 0, 2), (-k20, 2, 0, 0), (-k20, 0, 2, 0)],
            [(k2, 1, 0, 1)], [(k22, 2, 0, 0), (-k22, 0, 2, 0)]],
        3: [[(3 * k33, 2, 1, 0), (-k33, 0, 3, 0)], [(k32, 1, 1, 1)],
            [(4 * k31, 0, 1, 2), (-k31, 2, 1, 0), (-k31, 0, 3, 0)],
            [(2 * k30, 0, 0, 3), (-3 * k30, 2, 0, 1), (-3 * k30, 0, 2, 1)],
            [(4 * k31, 1, 0, 2), (-k31, 3, 0, 0), (-k31, 1, 2, 0)],
            [(k32c, 2, 0, 1), (-k32c, 0, 2, 1)], [(k33, 3, 0, 0), (-3 * k33, 1, 2, 0)]],
    }


_SOLID = _solid_monomials()


def _solid_gradient(l: int, v: torch.Tensor) -> torch.Tensor:
    """dC_lm/dq_alpha of the vectors v (n, 3), shape (2l+1, n, 3): each monomial
    differentiated term by term (``pseudo/kb.py:142-197``)."""
    if not 0 <= l <= MAX_L:
        raise NotImplementedError(f"solid harmonics are implemented up to l={MAX_L}")
    out = torch.zeros((2 * l + 1, v.shape[0], 3), dtype=v.dtype, device=v.device)
    powers = [[torch.ones_like(v[:, a])] for a in range(3)]
    for a in range(3):
        for _ in range(l):
            powers[a].append(powers[a][-1] * v[:, a])
    for m, terms in enumerate(_SOLID[l]):
        for coef, *e in terms:
            for axis in range(3):
                if e[axis] == 0:
                    continue
                f = [e[0], e[1], e[2]]
                f[axis] -= 1
                out[m, :, axis] += (coef * e[axis]) * powers[0][f[0]] * powers[1][f[1]] * powers[2][f[2]]
    return out


@dataclass
class Projectors:
    """``KBProjectors``-compatible: ``matrix`` (nproj, npw) complex, ``dij``, ``labels``.

    ``matrix`` is a device tensor; ``host_matrix()`` copies it to numpy for host consumers
    (forces and stress read ``.matrix`` directly in the reference, ``properties/forces.py:152-171``).
    """

    matrix: torch.Tensor
    dij: np.ndarray
    labels: list

    def __len__(self) -> int:
        return len(self.labels)

    def host_matrix(self) -> np.ndarray:
        return self.matrix.cpu().numpy()

    # The operator methods of ``KBProjectors`` (``pseudo/kb.py:216-260``) on the device
    # GEMMs. numpy in gives numpy out; device tensors stay on the device (FP64, or FP32
    # for complex64 input in ``project`` and ``apply``).
    def _block(self, coefficients):
        host = not isinstance(coefficients, torch.Tensor)
        c = torch.from_numpy(np.asarray(coefficients, dtype=complex)) if host else coefficients
        if c.shape[-1] != self.matrix.shape[-1]:
            raise ValueError(f"coefficients have {c.shape[-1]} plane waves, the projectors {self.matrix.shape[-1]}")
        one = c.dim() == 1
        c2 = (c[None, :] if one else c.reshape(-1, c.shape[-1])).to("cuda")
        if c2.dtype not in (torch.complex64, torch.complex128):
            c2 = c2.to(torch.complex128)
        return c2.contiguous(), host, one, c.shape

    def _matrix(self, dtype) -> torch.Tensor:
        return self.matrix.to("cuda", dtype).contiguous()

    def _coupling(self, coupling, dtype) -> torch.Tensor:
        d = self.dij if coupling is None else coupling
        return torch.as_tensor(np.asarray(d) if not isinstance(d, torch.Tensor) else d).to("cuda", dtype)

    def project(self, coefficients):
        """``<beta_i|psi_n>`` for a band or a block of bands, shape ``(..., nproj)``."""
        c2, host, one, shape = self._block(coefficients)
        p = ops.zgemm(c2, self._matrix(c2.dtype), 1)
        p = p[0] if one else p.reshape(*shape[:-1], p.shape[-1])
        return p.cpu().numpy() if host else p

    def apply(self, coefficients, coupling=None):
        """``V_NL |psi>``, same shape as ``coefficients``; ``coupling`` overrides D_ij."""
        c2, host, one, shape = self._block(coefficients)
        if len(self.labels) == 0:
            out = torch.zeros_like(c2)
        else:
            p = ops.zgemm(c2, self._matrix(c2.dtype), 1)
            pd = (p @ self._coupling(coupling, c2.dtype)).contiguous()
            out = ops.zgemm(pd, self._matrix(c2.dtype), 0)
        out = out[0] if one else out.reshape(shape)
        return out.cpu().numpy() if host else out

    def _projections64(self, coefficients) -> torch.Tensor:
        c2, _, _, _ = self._block(coefficients)
        c2 = c2.to(torch.complex128)
        return ops.zgemm(c2, self._matrix(torch.complex128), 1)

    def energy(self, coefficients, weights, coupling=None) -> float:
        """``sum_n w_n <psi_n|V_NL|psi_n>`` in Hartree (FP64)."""
        if len(self.labels) == 0:
            return 0.0
        p = self._projections64(coefficients)
        w = torch.as_tensor(np.asarray(weights, dtype=float) if not isinstance(weights, torch.Tensor) else weights,
                            dtype=torch.float64).to("cuda").reshape(-1)
        quad = (p.conj() * (p @ self._coupling(coupling, torch.complex128))).sum(-1).real
        return float((w * quad).sum().item())

    def occupancies(self, coefficients, weights) -> np.ndarray:
        """``D_ij = Re sum_n w_n <p_i|psi_n><psi_n|p_j>`` (FP64, numpy)."""
        if len(self.labels) == 0:
            return np.zeros((0, 0))
        p = self._projections64(coefficients)
        w = torch.as_tensor(np.asarray(weights, dtype=float) if not isinstance(weights, torch.Tensor) else weights,
                            dtype=torch.float64).to("cuda").reshape(-1)
        occ = (p * w[:, None]).T.contiguous() @ p.conj()
        return occ.real.cpu().numpy()

    def completeness(self) -> np.ndarray:
        """``sum_G |<k+G|beta_i>|^2`` per projector (FP64, numpy)."""
        m = self._matrix(torch.complex128)
        return (m.real**2 + m.imag**2).sum(-1).cpu().numpy()


def build_projectors(sphere, positions, symbols, datasets: dict, volume: float, *, device="cuda") -> Projectors:
    """The projector matrix and D for one k-point sphere (``pseudo/kb.py:260-343``)."""
    g2 = np.asarray(sphere.g2, dtype=float)
    q = np.sqrt(g2)
    # distinct |k+G| (shells): the radial transform depends on |k+G| only
    uq, inv = np.unique(q, return_inverse=True)
    kg = torch.from_numpy(np.asarray(sphere.cartesian, dtype=float)).to(device)
    inv_t = torch.from_numpy(inv.astype(np.int64)).to(device)
    pref = 4.0 * np.pi / np.sqrt(volume)
    radial, ylm = {}, {}
    for sym, ds in datasets.items():
        for ch in ds.projectors:
            f = torch.from_numpy(radial_form_factor(ds, ch, uq)).to(device)
            radial[(sym, ch.l, ch.index)] = f[inv_t]
            if ch.l not in ylm:
                ylm[ch.l] = _ylm(ch.l, kg)
    rows, labels, blocks = [], [], []
    pos = torch.from_numpy(np.asarray(positions, dtype=float)).to(device)
    for atom, sym in enumerate(symbols):
        ds = datasets[sym]
        phase = torch.exp(-1j * (kg @ pos[atom]))
        alab = []
        for ci, ch in enumerate(ds.projectors):
            coef = pref * (-1j) ** ch.l
            base = radial[(sym, ch.l, ch.index)] * phase
            for mi in range(2 * ch.l + 1):
                rows.append(coef * ylm[ch.l][mi] * base)
                alab.append((atom, ch.l, ci, mi - ch.l))
        nb = len(alab)
        blk = np.zeros((nb, nb))
        for a, (_, la, ca, ma) in enumerate(alab):
            for b, (_, lb, cb, mb) in enumerate(alab):
                if la == lb and ma == mb:
                    if ds.dij is not None:
                        blk[a, b] = ds.dij[ca, cb]
                    elif ca == cb:
                        blk[a, b] = ds.projectors[ca].energy
        labels += alab
        blocks.append(blk)
    if not rows:
        return Projectors(torch.zeros((0, len(g2)), dtype=torch.complex128, device=device), np.zeros((0, 0)), [])
    total = sum(len(b) for b in blocks)
    dij = np.zeros((total, total))
    off = 0
    for b in blocks:
        dij[off : off + len(b), off : off + len(b)] = b
        off += len(b)
    return Projectors(matrix=torch.stack(rows).contiguous(), dij=dij, labels=labels)


__all__ = ["Channel", "Dataset", "Projectors", "build_projectors", "radial_form_factor", "read_psp8"]
