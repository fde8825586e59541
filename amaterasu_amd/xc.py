"""Exchange-correlation on the GPU in FP64 (LDA-PZ, LDA-PW, PBE, PBEsol).

The pointwise forms are those of ``amaterasu/xc/lda.py`` and ``amaterasu/xc/gga.py``
(Slater exchange, PZ81, PW92 with and without the modified parametrisation, the PBE
enhancement factor and gradient correction written with log1p/expm1); the grid driver
follows ``amaterasu/xc/driver.py:108-164``: an unpolarised density is evaluated as n/2
in both spin channels, gradients and the divergence of the GGA flux are taken in
reciprocal space with the Nyquist plane blanked (``basis/fft.py:254-304``). Every
elementwise step is a torch FP64 operation on the device; the transforms are
libamaterasu_amd's batched 3D FFT.
"""

from __future__ import annotations

import numpy as np
import torch

from . import ops

X_FACTOR_C = 0.9305257363491000250020102180716672510262
RS_FACTOR = 0.6203504908994000166680068120477781673508
DENSITY_THRESHOLD = 1e-24
X2S = 0.1282782438530421943003109254455883701296
_FZ_DEN = 2.0 ** (4.0 / 3.0) - 2.0
_PW92 = {"a": (0.031091, 0.015545, 0.016887), "alpha1": (0.21370, 0.20548, 0.11125),
         "beta1": (7.5957, 14.1189, 10.357), "beta2": (3.5876, 6.1977, 3.6231),
         "beta3": (1.6382, 3.3662, 0.88026), "beta4": (0.49294, 0.62517, 0.49671), "fz20": 1.709921}
_PW92_MOD = {**_PW92, "a": (0.0310907, 0.01554535, 0.0168869), "fz20": 1.709920934161365617563962776245}
#: Perdew-Zunger 1981 correlation, (unpolarised, polarised) (``lda.py:78-87``)
_PZ81 = {"gamma": (-0.1423, -0.0843), "beta1": (1.0529, 1.3981), "beta2": (0.3334, 0.2611),
         "a": (0.0311, 0.01555), "b": (-0.048, -0.0269), "c": (0.0020, 0.0007), "d": (-0.0116, -0.0048)}
_GGA = {"PBE": {"kappa": 0.8040, "mu": 0.2195149727645171, "beta": 0.06672455060314922,
                "gamma": 0.031090690869654895034},
        "PBESOL": {"kappa": 0.8040, "mu": 0.1234567901234567901234567901234567901235, "beta": 0.046,
                   "gamma": 0.031090690869654895034}}


def _cbrt(x):
    return torch.sign(x) * torch.abs(x) ** (1.0 / 3.0)


def _rs_zeta(up, dn):
    n = up + dn
    safe = torch.clamp(n, min=DENSITY_THRESHOLD)
    return n, RS_FACTOR / _cbrt(safe), torch.clamp((up - dn) / safe, -1.0, 1.0)


def _fz(z):
    return (torch.clamp(1.0 + z, min=0.0) ** (4 / 3) + torch.clamp(1.0 - z, min=0.0) ** (4 / 3) - 2.0) / _FZ_DEN


def _dfz(z):
    return (4 / 3) * (_cbrt(torch.clamp(1.0 + z, min=0.0)) - _cbrt(torch.clamp(1.0 - z, min=0.0))) / _FZ_DEN


def _pw92_g(rs, i, par):
    a, al = par["a"][i], par["alpha1"][i]
    b1, b2, b3, b4 = par["beta1"][i], par["beta2"][i], par["beta3"][i], par["beta4"][i]
    srs = torch.sqrt(rs)
    q = b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs
    dq = 0.5 * b1 / srs + b2 + 1.5 * b3 * srs + 2.0 * b4 * rs
    lg = torch.log1p(1.0 / (2.0 * a * q))
    return -2.0 * a * (1.0 + al * rs) * lg, -2.0 * a * al * lg + 2.0 * a * (1.0 + al * rs) * dq / (2.0 * a * q * q + q)


def _pw92_local(rs, z, par):
    g0, dg0 = _pw92_g(rs, 0, par)
    g1, dg1 = _pw92_g(rs, 1, par)
    g2, dg2 = _pw92_g(rs, 2, par)
    f20 = par["fz20"]
    fz, dfz = _fz(z), _dfz(z)
    z3, z4 = z**3, z**4
    br = g1 - g0 + g2 / f20
    eps = g0 + z4 * fz * br - fz * g2 / f20
    return eps, dg0 + z4 * fz * (dg1 - dg0 + dg2 / f20) - fz * dg2 / f20, (4.0 * z3 * fz + z4 * dfz) * br - dfz * g2 / f20


def _pz81_branch(rs, i):
    """eps_c, d eps_c/d rs of one spin limit: Pade form for rs >= 1, high-density expansion
    below (``lda.py:203-234``)."""
    g, b1, b2 = _PZ81["gamma"][i], _PZ81["beta1"][i], _PZ81["beta2"][i]
    a, b, c, d = _PZ81["a"][i], _PZ81["b"][i], _PZ81["c"][i], _PZ81["d"][i]
    srs = torch.sqrt(rs)
    den = 1.0 + b1 * srs + b2 * rs
    lr = torch.log(torch.clamp(rs, min=1e-300))
    dense = rs < 1.0
    eps = torch.where(dense, a * lr + b + c * rs * lr + d * rs, g / den)
    deps = torch.where(dense, a / rs + c * (lr + 1.0) + d, -g * (0.5 * b1 / srs + b2) / (den * den))
    return eps, deps


def _pz81_local(rs, z):
    e0, d0 = _pz81_branch(rs, 0)
    e1, d1 = _pz81_branch(rs, 1)
    fz = _fz(z)
    return e0 + fz * (e1 - e0), d0 + fz * (d1 - d0), _dfz(z) * (e1 - e0)


def _lda(up, dn, pz81: bool = False):
    """(eps, v_up, v_dn): Slater exchange + PW92 (or PZ81) correlation."""
    n, rs, z = _rs_zeta(up, dn)
    act = n > DENSITY_THRESHOLD
    su, sd = torch.clamp(up, min=0.0), torch.clamp(dn, min=0.0)
    cu, cd = _cbrt(su), _cbrt(sd)
    ex = torch.where(act, -X_FACTOR_C * (cu * su + cd * sd) / torch.where(act, n, torch.ones_like(n)),
                     torch.zeros_like(n))
    vxu = torch.where(act, -(4 / 3) * X_FACTOR_C * cu, torch.zeros_like(n))
    vxd = torch.where(act, -(4 / 3) * X_FACTOR_C * cd, torch.zeros_like(n))
    ec, dr, dz = _pz81_local(rs, z) if pz81 else _pw92_local(rs, z, _PW92)
    common = ec - (rs / 3.0) * dr
    zero = torch.zeros_like(n)
    ec = torch.where(act, ec, zero)
    vcu = torch.where(act, common + (1.0 - z) * dz, zero)
    vcd = torch.where(act, common - (1.0 + z) * dz, zero)
    return ex + ec, vxu + vcu, vxd + vcd


def _gga(up, dn, suu, sud, sdd, par):
    """(eps, v_up, v_dn, vsigma_uu, vsigma_ud, vsigma_dd) of the PBE form."""
    kappa, mu, beta, gamma = par["kappa"], par["mu"], par["beta"], par["gamma"]
    n, rs, z = _rs_zeta(up, dn)
    act = n > DENSITY_THRESHOLD
    zero = torch.zeros_like(n)
    # exchange, per spin channel
    ex_e = zero
    vx, vsx = [], []
    for s, ss in ((up, suu), (dn, sdd)):
        sp = torch.clamp(s, min=0.0)
        safe = torch.clamp(sp, min=DENSITY_THRESHOLD)
        n43 = safe ** (4 / 3)
        s2 = (X2S * torch.sqrt(torch.clamp(ss, min=0.0)) / n43) ** 2
        den = kappa + mu * s2
        fx = 1.0 + kappa * mu * s2 / den
        fxs = 2.0 * kappa * kappa * mu * s2 / den**2
        keep = (sp > DENSITY_THRESHOLD) & act
        ex_e = ex_e + torch.where(sp > DENSITY_THRESHOLD, -X_FACTOR_C * n43 * fx, zero)
        vx.append(torch.where(keep, -(4 / 3) * X_FACTOR_C * _cbrt(safe) * (fx - fxs), zero))
        vsx.append(torch.where(keep, -X_FACTOR_C * kappa * kappa * mu * X2S**2 / (n43 * den**2), zero))
    ex = torch.where(act, ex_e / torch.where(act, n, torch.ones_like(n)), zero)
    # correlation on modified PW92
    safe = torch.clamp(n, min=DENSITY_THRESHOLD)
    elda, dl_drs, dl_dz = _pw92_local(rs, z, _PW92_MOD)
    opz, omz = torch.clamp(1.0 + z, min=DENSITY_THRESHOLD), torch.clamp(1.0 - z, min=DENSITY_THRESHOLD)
    phi = 0.5 * (opz ** (2 / 3) + omz ** (2 / 3))
    dphi = (1 / 3) * (opz ** (-1 / 3) - omz ** (-1 / 3))
    phi3 = phi**3
    stot = torch.clamp(suu + 2.0 * sud + sdd, min=0.0)
    tpx = 1.0 / (4.0 * 2.0 ** (1 / 3) * phi * torch.sqrt(rs))
    xt = torch.sqrt(stot) / safe ** (4 / 3)
    t = tpx * xt
    t2 = t * t
    et = torch.expm1(-elda / (gamma * phi3))
    es = torch.where(torch.abs(et) < 1e-300, torch.full_like(et, 1e-300), et)
    A = beta / (gamma * es)
    u = t2 + A * t2 * t2
    d = 1.0 + A * u
    w = beta * u / (gamma * d)
    lg = torch.log1p(w)
    H = gamma * phi3 * lg
    pre = gamma * phi3 / (1.0 + w)
    dh_dt = pre * beta / (gamma * d**2) * (2.0 * t + 4.0 * A * t2 * t)
    dh_da = pre * beta * (t2 * t2 - u * u) / (gamma * d**2)
    da_de = beta * (et + 1.0) / (gamma**2 * phi3 * es**2)
    da_dphi = -3.0 * beta * elda * (et + 1.0) / (gamma**2 * phi**4 * es**2)
    dh_drs = dh_da * da_de * dl_drs + dh_dt * (-t / (2.0 * rs))
    dh_dz = 3.0 * gamma * phi * phi * lg * dphi + dh_da * (da_dphi * dphi + da_de * dl_dz) + dh_dt * (-t / phi) * dphi
    ec = elda + H
    de_drs = dl_drs + dh_drs
    de_dz = dl_dz + dh_dz
    common = ec - (rs / 3.0) * de_drs - (4 / 3) * xt * dh_dt * tpx
    vcu = torch.where(act, common + (1.0 - z) * de_dz, zero)
    vcd = torch.where(act, common - (1.0 + z) * de_dz, zero)
    vst = torch.where(act, safe * phi3 * beta * (tpx**2 / safe ** (8 / 3)) * (2.0 + 4.0 * A * t2)
                      / (2.0 * (1.0 + w) * d**2), zero)
    ec = torch.where(act, ec, zero)
    return ex + ec, vx[0] + vcu, vx[1] + vcd, vsx[0] + vst, 2.0 * vst, vsx[1] + vst


class GridOps:
    """Reciprocal-space derivatives on one grid (FP64, device)."""

    def __init__(self, shape, cell, device="cuda"):
        self.shape = tuple(shape)
        self.cell = np.asarray(cell, dtype=float)
        recip = 2.0 * np.pi * np.linalg.inv(self.cell).T
        freqs = []
        for n in self.shape:
            f = np.fft.fftfreq(n, d=1.0 / n)
            freqs.append(f)
        m = np.stack(np.meshgrid(*freqs, indexing="ij"), axis=-1)
        g = m @ recip
        fz = []
        for n, f in zip(self.shape, freqs):
            f = f.copy()
            if n % 2 == 0:
                f[n // 2] = 0.0
            fz.append(f)
        mz = np.stack(np.meshgrid(*fz, indexing="ij"), axis=-1)
        self.g = torch.from_numpy(g).to(device)
        self.gz = torch.from_numpy(mz @ recip).to(device)  # Nyquist planes blanked
        self.g2 = (self.g**2).sum(-1)
        self.size = int(np.prod(self.shape))
        self.volume = abs(float(np.linalg.det(self.cell)))

    def to_reciprocal(self, f):
        return ops.fft3d(f.to(torch.complex128).contiguous(), -1, 1.0 / self.size)

    def to_real(self, c):
        return ops.fft3d(c.to(torch.complex128).contiguous(), +1, 1.0)

    def gradient(self, f):
        t = self.to_reciprocal(f)
        return torch.stack([self.to_real(1j * self.gz[..., a] * t).real for a in range(3)], dim=-1)

    def divergence(self, v):
        acc = torch.zeros(self.shape, dtype=torch.complex128, device=v.device)
        for a in range(3):
            acc = acc + 1j * self.gz[..., a] * self.to_reciprocal(v[..., a])
        return self.to_real(acc).real


class XCFunctional:
    """Grid evaluation of LDA-PW / PBE / PBEsol on the device (``xc/driver.py:108-164``)."""

    def __init__(self, name: str):
        key = name.strip().upper().replace("_", "-")
        if key not in ("LDA-PZ", "LDA-PW", "PBE", "PBESOL"):
            raise ValueError(f"unknown functional {name!r}; available: LDA-PZ, LDA-PW, PBE, PBEsol")
        self.key = key
        self.name = {"PBESOL": "PBEsol"}.get(key, key)

    @property
    def is_gga(self) -> bool:
        return self.key not in ("LDA-PZ", "LDA-PW")

    def evaluate_on_grid(self, density: torch.Tensor, gops: GridOps):
        """(energy, potential (nspin, *grid)) for density (nspin, *grid) FP64 on the device."""
        energy, pot, _, _ = self.evaluate_on_grid_details(density, gops)
        return energy, pot

    def evaluate_on_grid_details(self, density: torch.Tensor, gops: GridOps):
        """(energy, potential, gradients, vsigma): for a GGA also the spin-channel gradients
        (2, *grid, 3) and vsigma (3, *grid) (up-up, up-down, down-down), which the stress
        needs (``xc/driver.py:47-64``); ``None`` for an LDA."""
        d = density if density.dim() == 4 else density[None]
        nspin = d.shape[0]
        up, dn = (d[0] / 2, d[0] / 2) if nspin == 1 else (d[0], d[1])
        dv = gops.volume / gops.size
        if not self.is_gga:
            eps, vu, vd = _lda(up, dn, pz81=self.key == "LDA-PZ")
            energy = float((eps * (up + dn)).sum().item()) * dv
            pot = vu[None] if nspin == 1 else torch.stack([vu, vd])
            return energy, pot, None, None
        gu = gops.gradient(up)
        gd = gu if nspin == 1 else gops.gradient(dn)
        suu = (gu * gu).sum(-1)
        sud = (gu * gd).sum(-1)
        sdd = (gd * gd).sum(-1)
        eps, vu, vd, vsuu, vsud, vsdd = _gga(up, dn, suu, sud, sdd, _GGA[self.key])
        energy = float((eps * (up + dn)).sum().item()) * dv
        vu = vu - gops.divergence(2.0 * vsuu[..., None] * gu + vsud[..., None] * gd)
        if nspin == 1:
            pot = vu[None]
        else:
            vd = vd - gops.divergence(2.0 * vsdd[..., None] * gd + vsud[..., None] * gu)
            pot = torch.stack([vu, vd])
        return energy, pot, (gu, gd), (vsuu, vsud, vsdd)


__all__ = ["GridOps", "XCFunctional"]
