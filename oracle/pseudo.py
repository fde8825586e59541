"""Norm-conserving datasets: psp8 reading, radial transforms, real Y_lm and KB projectors.

Restates ``amaterasu/pseudo/base.py``, ``pseudo/readers/psp8.py``,
``pseudo/formfactors.py`` and ``pseudo/kb.py`` (the norm-conserving part only).
Radial integrals use ``scipy.integrate.simpson`` exactly as the reference's
``RadialGrid.integrate`` does (``base.py:130-144``).
"""

from __future__ import annotations

import gzip
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
from scipy.integrate import simpson
from scipy.special import erf, spherical_jn

MAX_L = 3

# ABINIT pspxc codes the reference decodes to a functional name (base.py:42-109).
_PSPXC = {2: "LDA-PZ", 7: "LDA-PW", 11: "PBE", -116133: "PBEsol", -101130: "PBE", -1012: "LDA-PW"}


@dataclass(frozen=True)
class RadialGrid:
    r: np.ndarray

    def __len__(self) -> int:
        return len(self.r)

    def integrate(self, values):
        """Simpson's rule over the last axis with abscissae r (``base.py:130-144``, linear mesh)."""
        values = np.asarray(values, dtype=float)
        total = simpson(values, x=self.r, axis=-1)
        return float(total) if values.ndim == 1 else total


@dataclass(frozen=True)
class Projector:
    l: int
    index: int
    energy: float
    values: np.ndarray  # r beta_l(r)

    def norm(self, grid: RadialGrid) -> float:
        return grid.integrate(self.values**2)


@dataclass
class NCPseudopotential:
    symbol: str
    atomic_number: float
    valence_charge: float
    grid: RadialGrid
    local_potential: np.ndarray
    projectors: list
    l_max: int
    l_local: int
    xc_functional: str | None = None
    core_density: np.ndarray | None = None  # 4 pi r^2 rho_c
    valence_density: np.ndarray | None = None  # 4 pi r^2 rho_v
    dij: np.ndarray | None = None
    kind: str = "NC"
    extra: dict = field(default_factory=dict)

    @property
    def has_core_correction(self) -> bool:
        return self.core_density is not None

    def integrated_valence_charge(self) -> float:
        return self.grid.integrate(self.valence_density)

    def local_potential_tail_error(self, r_min: float = 3.0) -> float:
        out = self.grid.r >= r_min
        expected = -self.valence_charge / self.grid.r[out]
        return float(np.max(np.abs(self.local_potential[out] - expected) / np.abs(expected)))


def _fortran_float(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def _rows(lines, start, count):
    return np.array([[_fortran_float(t) for t in lines[start + i].split()[1:]] for i in range(count)])


def read_psp8(path) -> NCPseudopotential:
    """ABINIT psp8 reader (``readers/psp8.py:90-220``): six header lines, one block per
    non-local l (KB energies + mmax rows of r beta), the local block, then the optional
    core and valence densities stored as 4 pi rho (multiplied by r^2 on load)."""
    path = Path(path)
    opener = gzip.open if path.suffix == ".gz" else open
    with opener(path, "rt", encoding="utf-8", errors="replace") as fh:
        lines = fh.read().splitlines()
    title = lines[0].strip()
    zatom, zion = (_fortran_float(t) for t in lines[1].split()[:2])
    head = lines[2].split()
    pspcod, pspxc, lmax, lloc, mmax = (int(x) for x in head[:5])
    if pspcod != 8:
        raise ValueError(f"{path.name}: pspcod {pspcod} is not 8")
    fchrg = _fortran_float(lines[3].split()[1])
    nproj = [int(x) for x in lines[4].split()[:5]]
    ext = int(lines[5].split()[0])
    if ext in (2, 3):
        raise NotImplementedError("spin-orbit psp8 files are not read")
    cur = 6
    projectors = []
    rref = None
    for l in range(lmax + 1):
        if l == lloc or l >= len(nproj) or nproj[l] == 0:
            continue
        toks = lines[cur].split()
        if int(toks[0]) != l:
            raise ValueError(f"{path.name}: expected l={l} block at line {cur + 1}")
        ekb = [_fortran_float(t) for t in toks[1 : 1 + nproj[l]]]
        tab = _rows(lines, cur + 1, mmax)
        rref = tab[:, 0] if rref is None else rref
        for j in range(nproj[l]):
            projectors.append(Projector(l=l, index=j, energy=ekb[j], values=np.ascontiguousarray(tab[:, 1 + j])))
        cur += 1 + mmax
    if int(lines[cur].split()[0]) != lloc:
        raise ValueError(f"{path.name}: expected the local block at line {cur + 1}")
    tab = _rows(lines, cur + 1, mmax)
    r, vloc = tab[:, 0], tab[:, 1]
    cur += 1 + mmax
    grid = RadialGrid(r=np.ascontiguousarray(r))
    core = None
    if fchrg > 0.0:
        core = np.ascontiguousarray(_rows(lines, cur, mmax)[:, 1] * r**2)
        cur += mmax
    valence = None
    if ext in (1, 3):
        valence = np.ascontiguousarray(_rows(lines, cur, mmax)[:, 1] * r**2)
        cur += mmax
    return NCPseudopotential(
        symbol=title.split()[0],
        atomic_number=zatom,
        valence_charge=zion,
        grid=grid,
        local_potential=np.ascontiguousarray(vloc),
        projectors=projectors,
        l_max=lmax,
        l_local=lloc,
        xc_functional=_PSPXC.get(pspxc),
        core_density=core,
        valence_density=valence,
        extra={"pspxc": pspxc, "fchrg": fchrg},
    )


# ----------------------------------------------------------------------- form factors
def radial_transform(grid: RadialGrid, values, l: int, q) -> np.ndarray:
    """int f(r) j_l(q r) dr at every q (``formfactors.py:55-71``)."""
    q = np.atleast_1d(np.asarray(q, dtype=float))
    values = np.asarray(values, dtype=float)
    if values.shape != grid.r.shape:
        raise ValueError(f"values has shape {values.shape}, grid has {grid.r.shape}")
    return grid.integrate(values * spherical_jn(int(l), q[:, None] * grid.r[None, :]))


def projector_form_factor(grid: RadialGrid, projector: Projector, q) -> np.ndarray:
    """f_l(q) = int r^2 beta j_l(qr) dr (``formfactors.py:106-114``)."""
    return radial_transform(grid, grid.r * projector.values, projector.l, q)


def density_form_factor(grid: RadialGrid, density, q) -> np.ndarray:
    """``formfactors.py:117-123``."""
    return radial_transform(grid, density, 0, q)


def local_potential_form_factor(pseudo: NCPseudopotential, q) -> np.ndarray:
    """v_loc(q) with the erf-screened Coulomb split; alpha term at q=0 (``formfactors.py:74-103``)."""
    q = np.atleast_1d(np.asarray(q, dtype=float))
    g, z = pseudo.grid, float(pseudo.valence_charge)
    r = g.r
    out = np.empty_like(q)
    zero = q < 1e-12
    if np.any(zero):
        out[zero] = 4.0 * np.pi * g.integrate(r * (r * pseudo.local_potential + z))
    if np.any(~zero):
        qq = q[~zero]
        screened = r * pseudo.local_potential + z * erf(r)
        short = 4.0 * np.pi / qq * g.integrate(screened * np.sin(qq[:, None] * r[None, :]))
        out[~zero] = short - 4.0 * np.pi * z * np.exp(-0.25 * qq**2) / qq**2
    return out


# ------------------------------------------------------- form factor derivatives
def radial_transform_derivative(grid: RadialGrid, values, l: int, q) -> np.ndarray:
    """d/dq int f(r) j_l(qr) dr = int f(r) r j_l'(qr) dr (``formfactors.py:133-143``)."""
    q = np.atleast_1d(np.asarray(q, dtype=float))
    values = np.asarray(values, dtype=float)
    if values.shape != grid.r.shape:
        raise ValueError(f"values has shape {values.shape}, grid has {grid.r.shape}")
    x = q[:, None] * grid.r[None, :]
    return grid.integrate(values * grid.r * spherical_jn(int(l), x, derivative=True))


def density_form_factor_derivative(grid: RadialGrid, density, q) -> np.ndarray:
    """``formfactors.py:146-150``."""
    return radial_transform_derivative(grid, density, 0, q)


def projector_form_factor_derivative(grid: RadialGrid, projector: Projector, q) -> np.ndarray:
    """``formfactors.py:153-157``."""
    return radial_transform_derivative(grid, grid.r * projector.values, projector.l, q)


def local_potential_form_factor_derivative(pseudo: NCPseudopotential, q) -> np.ndarray:
    """d v_loc/dq of both pieces of the erf split; 0 at q = 0 (``formfactors.py:160-185``).

    With s(r) = r V_loc + Z erf(r): v(q) = (4 pi/q) int s sin(qr) dr - 4 pi Z e^{-q^2/4}/q^2,
    differentiated term by term."""
    q = np.atleast_1d(np.asarray(q, dtype=float))
    g, z = pseudo.grid, float(pseudo.valence_charge)
    r = g.r
    out = np.zeros_like(q)
    nz = q >= 1e-12
    if not np.any(nz):
        return out
    qq = q[nz]
    s = r * pseudo.local_potential + z * erf(r)
    arg = qq[:, None] * r[None, :]
    i_sin = g.integrate(s * np.sin(arg))
    i_cos = g.integrate(s * r * np.cos(arg))
    gauss = np.exp(-0.25 * qq**2)
    out[nz] = (-4.0 * np.pi / qq**2 * i_sin + 4.0 * np.pi / qq * i_cos
               + 4.0 * np.pi * z * gauss * (0.5 / qq + 2.0 / qq**3))
    return out


# ------------------------------------------------------------------ solid harmonics
def _solid_table():
    """C_lm(q) = |q|^l Y_lm(q_hat) as homogeneous polynomials {(a, b, c): coefficient} of
    x^a y^b z^c, m = -l..l in the order of :func:`real_spherical_harmonics`
    (``kb.py:98-139``)."""
    pi = np.pi
    c1 = np.sqrt(0.75 / pi)
    c2, cz2, cd2 = 0.5 * np.sqrt(15.0 / pi), 0.25 * np.sqrt(5.0 / pi), 0.25 * np.sqrt(15.0 / pi)
    ca, cb, cc = 0.25 * np.sqrt(17.5 / pi), 0.5 * np.sqrt(105.0 / pi), 0.25 * np.sqrt(10.5 / pi)
    cd, ce = 0.25 * np.sqrt(7.0 / pi), 0.25 * np.sqrt(105.0 / pi)
    return {
        0: [{(0, 0, 0): 0.5 / np.sqrt(pi)}],
        1: [{(0, 1, 0): c1}, {(0, 0, 1): c1}, {(1, 0, 0): c1}],
        2: [{(1, 1, 0): c2}, {(0, 1, 1): c2}, {(0, 0, 2): 2 * cz2, (2, 0, 0): -cz2, (0, 2, 0): -cz2},
            {(1, 0, 1): c2}, {(2, 0, 0): cd2, (0, 2, 0): -cd2}],
        3: [{(2, 1, 0): 3 * ca, (0, 3, 0): -ca},
            {(1, 1, 1): cb},
            {(0, 1, 2): 4 * cc, (2, 1, 0): -cc, (0, 3, 0): -cc},
            {(0, 0, 3): 2 * cd, (2, 0, 1): -3 * cd, (0, 2, 1): -3 * cd},
            {(1, 0, 2): 4 * cc, (3, 0, 0): -cc, (1, 2, 0): -cc},
            {(2, 0, 1): ce, (0, 2, 1): -ce},
            {(3, 0, 0): ca, (1, 2, 0): -3 * ca}],
    }


SOLID_HARMONICS = _solid_table()


def _poly(terms: dict, q) -> np.ndarray:
    out = np.zeros(len(q))
    for (a, b, c), coef in terms.items():
        out += coef * q[:, 0] ** a * q[:, 1] ** b * q[:, 2] ** c
    return out


def solid_harmonics(l: int, vectors) -> np.ndarray:
    """C_lm(q), shape (2l+1, n) (``kb.py:98-139``)."""
    if not 0 <= l <= MAX_L:
        raise NotImplementedError(f"solid harmonics are implemented up to l={MAX_L}")
    q = np.atleast_2d(np.asarray(vectors, dtype=float))
    return np.stack([_poly(t, q) for t in SOLID_HARMONICS[l]])


def solid_harmonics_gradient(l: int, vectors) -> np.ndarray:
    """dC_lm/dq_alpha, shape (2l+1, n, 3), by differentiating each monomial
    (``kb.py:142-197``)."""
    if not 0 <= l <= MAX_L:
        raise NotImplementedError(f"solid harmonics are implemented up to l={MAX_L}")
    q = np.atleast_2d(np.asarray(vectors, dtype=float))
    out = np.zeros((2 * l + 1, len(q), 3))
    for m, terms in enumerate(SOLID_HARMONICS[l]):
        for axis in range(3):
            d = {}
            for exps, coef in terms.items():
                if exps[axis] == 0:
                    continue
                lowered = list(exps)
                lowered[axis] -= 1
                d[tuple(lowered)] = d.get(tuple(lowered), 0.0) + coef * exps[axis]
            out[m, :, axis] = _poly(d, q)
    return out


# ------------------------------------------------------------------ real harmonics
def real_spherical_harmonics(l: int, directions) -> np.ndarray:
    """Real Y_lm, m = -l..l, l=1 ordered (y, z, x); zero vector -> +z (``kb.py:52-97``)."""
    if not 0 <= l <= MAX_L:
        raise NotImplementedError(f"real spherical harmonics are implemented up to l={MAX_L}")
    d = np.atleast_2d(np.asarray(directions, dtype=float))
    n = np.linalg.norm(d, axis=1)
    u = np.zeros_like(d)
    u[:, 2] = 1.0
    nz = n > 1e-14
    u[nz] = d[nz] / n[nz, None]
    x, y, z = u[:, 0], u[:, 1], u[:, 2]
    if l == 0:
        return np.full((1, len(u)), 0.5 / np.sqrt(np.pi))
    if l == 1:
        return np.sqrt(0.75 / np.pi) * np.stack([y, z, x])
    if l == 2:
        c = 0.5 * np.sqrt(15.0 / np.pi)
        return np.stack(
            [c * x * y, c * y * z, 0.25 * np.sqrt(5.0 / np.pi) * (3.0 * z**2 - 1.0), c * x * z,
             0.25 * np.sqrt(15.0 / np.pi) * (x**2 - y**2)]
        )
    return np.stack(
        [
            0.25 * np.sqrt(17.5 / np.pi) * y * (3.0 * x**2 - y**2),
            0.5 * np.sqrt(105.0 / np.pi) * x * y * z,
            0.25 * np.sqrt(10.5 / np.pi) * y * (5.0 * z**2 - 1.0),
            0.25 * np.sqrt(7.0 / np.pi) * z * (5.0 * z**2 - 3.0),
            0.25 * np.sqrt(10.5 / np.pi) * x * (5.0 * z**2 - 1.0),
            0.25 * np.sqrt(105.0 / np.pi) * z * (x**2 - y**2),
            0.25 * np.sqrt(17.5 / np.pi) * x * (x**2 - 3.0 * y**2),
        ]
    )


# ---------------------------------------------------------------------- KB projectors
@dataclass
class KBProjectors:
    """Dense <k+G|beta_i> of shape (nproj, npw) and D_ij (``kb.py:199-257``)."""

    matrix: np.ndarray
    dij: np.ndarray
    labels: list

    def __len__(self) -> int:
        return len(self.labels)

    def project(self, c):
        """``kb.py:216-218``: c @ beta^H."""
        return np.asarray(c) @ self.matrix.conj().T

    def apply(self, c, coupling=None):
        """``kb.py:220-230``: (P D) beta."""
        d = self.dij if coupling is None else coupling
        if len(self.labels) == 0:
            return np.zeros_like(c)
        return (self.project(c) @ d) @ self.matrix

    def energy(self, c, weights, coupling=None) -> float:
        """``kb.py:232-241``."""
        d = self.dij if coupling is None else coupling
        if len(self.labels) == 0:
            return 0.0
        p = self.project(c)
        return float(np.dot(weights, np.einsum("ni,ij,nj->n", p.conj(), d, p).real))

    def occupancies(self, c, weights) -> np.ndarray:
        """``kb.py:243-256``: Re sum_n w_n <p_i|psi_n><psi_n|p_j>."""
        if len(self.labels) == 0:
            return np.zeros((0, 0))
        p = self.project(c)
        return np.einsum("n,ni,nj->ij", weights, p, p.conj()).real

    def completeness(self) -> np.ndarray:
        """``kb.py:258-260``: sum_G |<k+G|beta_i>|^2."""
        return np.einsum("ig,ig->i", self.matrix.conj(), self.matrix).real


def build_projectors(sphere, positions, symbols, pseudopotentials, volume) -> KBProjectors:
    """(4 pi/sqrt(Omega)) (-i)^l Y_lm(k+G) f_l(|k+G|) e^{-i(k+G).tau}; atom-major, then
    channel, then m; D block-diagonal per atom coupling equal (l, m) (``kb.py:260-343``)."""
    kg = sphere.cartesian
    q = np.sqrt(sphere.g2)
    pref = 4.0 * np.pi / np.sqrt(volume)
    ff, ylm = {}, {}
    for sym, ps in pseudopotentials.items():
        for p in ps.projectors:
            ff[(sym, p.l, p.index)] = projector_form_factor(ps.grid, p, q)
            if p.l not in ylm:
                ylm[p.l] = real_spherical_harmonics(p.l, kg)
    rows, labels, blocks = [], [], []
    for atom, sym in enumerate(symbols):
        ps = pseudopotentials[sym]
        phase = np.exp(-1j * (kg @ positions[atom]))
        arows, alabels = [], []
        for ch, p in enumerate(ps.projectors):
            for mi in range(2 * p.l + 1):
                arows.append(pref * (-1j) ** p.l * ylm[p.l][mi] * ff[(sym, p.l, p.index)] * phase)
                alabels.append((atom, p.l, ch, mi - p.l))
        nb = len(arows)
        blk = np.zeros((nb, nb))
        for a, (_, la, ca, ma) in enumerate(alabels):
            for b, (_, lb, cb, mb) in enumerate(alabels):
                if la != lb or ma != mb:
                    continue
                if ps.dij is not None:
                    blk[a, b] = ps.dij[ca, cb]
                elif ca == cb:
                    blk[a, b] = ps.projectors[ca].energy
        rows.extend(arows)
        labels.extend(alabels)
        blocks.append(blk)
    if not rows:
        return KBProjectors(np.zeros((0, len(sphere)), dtype=complex), np.zeros((0, 0)), [])
    total = sum(len(b) for b in blocks)
    dij = np.zeros((total, total))
    off = 0
    for b in blocks:
        dij[off : off + len(b), off : off + len(b)] = b
        off += len(b)
    return KBProjectors(matrix=np.asarray(rows), dij=dij, labels=labels)
