"""Exchange-correlation functionals and their grid evaluation (FP64, CPU).

Restates ``amaterasu/xc/lda.py`` (Slater exchange, PW92 / modified PW92 correlation),
``amaterasu/xc/gga.py`` (PBE, PBEsol) and ``amaterasu/xc/driver.py:108-164``
(spin-polarised evaluation of the unpolarised case as n/2 per channel, GGA gradients
and the divergence term by FFT). The parameter tables are the published PW92/PBE
constants in the digits the reference uses (``lda.py:49-76``, ``gga.py:60-76``).
"""

from __future__ import annotations

import numpy as np

X_FACTOR_C = 0.9305257363491000250020102180716672510262
RS_FACTOR = 0.6203504908994000166680068120477781673508
DENSITY_THRESHOLD = 1e-24
FZ_DEN = 2.0 ** (4.0 / 3.0) - 2.0
X2S = 0.1282782438530421943003109254455883701296

PW92 = {"a": (0.031091, 0.015545, 0.016887), "alpha1": (0.21370, 0.20548, 0.11125),
        "beta1": (7.5957, 14.1189, 10.357), "beta2": (3.5876, 6.1977, 3.6231),
        "beta3": (1.6382, 3.3662, 0.88026), "beta4": (0.49294, 0.62517, 0.49671), "fz20": 1.709921}
PW92_MOD = {**PW92, "a": (0.0310907, 0.01554535, 0.0168869), "fz20": 1.709920934161365617563962776245}
PBE = {"kappa": 0.8040, "mu": 0.2195149727645171, "beta": 0.06672455060314922, "gamma": 0.031090690869654895034}
PBESOL = {"kappa": 0.8040, "mu": 0.1234567901234567901234567901234567901235, "beta": 0.046,
          "gamma": 0.031090690869654895034}


def rs_and_zeta(rho):
    """``lda.py:98-104``."""
    n = rho[:, 0] + rho[:, 1]
    safe = np.maximum(n, DENSITY_THRESHOLD)
    return n, RS_FACTOR / np.cbrt(safe), np.clip((rho[:, 0] - rho[:, 1]) / safe, -1.0, 1.0)


def f_zeta(z):
    """``lda.py:107-111``."""
    return (np.maximum(1.0 + z, 0.0) ** (4 / 3) + np.maximum(1.0 - z, 0.0) ** (4 / 3) - 2.0) / FZ_DEN


def f_zeta_derivative(z):
    """``lda.py:114-118``."""
    return (4 / 3) * (np.cbrt(np.maximum(1.0 + z, 0.0)) - np.cbrt(np.maximum(1.0 - z, 0.0))) / FZ_DEN


def lda_exchange(rho):
    """e_x = -X_FACTOR_C sum_s n_s^(4/3), v_s = -(4/3) X_FACTOR_C n_s^(1/3) (``lda.py:122-142``)."""
    rho = np.atleast_2d(np.asarray(rho, dtype=float))
    n = rho[:, 0] + rho[:, 1]
    act = n > DENSITY_THRESHOLD
    spin = np.maximum(rho, 0.0)
    cr = np.cbrt(spin)
    e = -X_FACTOR_C * np.sum(cr * spin, axis=1)
    eps = np.zeros_like(n)
    np.divide(e, n, out=eps, where=act)
    return eps, np.where(act[:, None], -(4 / 3) * X_FACTOR_C * cr, 0.0)


def pw92_g(rs, i, par):
    """PW92 eq. (10) G(rs) and dG/drs with log1p (``lda.py:146-166``)."""
    a, al = par["a"][i], par["alpha1"][i]
    b1, b2, b3, b4 = par["beta1"][i], par["beta2"][i], par["beta3"][i], par["beta4"][i]
    srs = np.sqrt(rs)
    q = b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs
    dq = 0.5 * b1 / srs + b2 + 1.5 * b3 * srs + 2.0 * b4 * rs
    lg = np.log1p(1.0 / (2.0 * a * q))
    g = -2.0 * a * (1.0 + al * rs) * lg
    dg = -2.0 * a * al * lg + 2.0 * a * (1.0 + al * rs) * dq / (2.0 * a * q * q + q)
    return g, dg


def _assemble(eps, d_drs, d_dz, rs, z, act):
    """``lda.py:247-262``: v_s = eps - (rs/3) deps/drs + (s - zeta) deps/dzeta."""
    common = eps - (rs / 3.0) * d_drs
    v = np.stack([common + (1.0 - z) * d_dz, common - (1.0 + z) * d_dz], axis=-1)
    return np.where(act, eps, 0.0), np.where(act[:, None], v, 0.0)


def _pw92_local(rs, z, par):
    g0, dg0 = pw92_g(rs, 0, par)
    g1, dg1 = pw92_g(rs, 1, par)
    g2, dg2 = pw92_g(rs, 2, par)
    f20 = par["fz20"]
    fz, dfz = f_zeta(z), f_zeta_derivative(z)
    z3, z4 = z**3, z**4
    br = g1 - g0 + g2 / f20
    eps = g0 + z4 * fz * br - fz * g2 / f20
    d_drs = dg0 + z4 * fz * (dg1 - dg0 + dg2 / f20) - fz * dg2 / f20
    d_dz = (4.0 * z3 * fz + z4 * dfz) * br - dfz * g2 / f20
    return eps, d_drs, d_dz


#: Perdew-Zunger 1981 (PRB 23, 5048) correlation parameters, (unpolarised, polarised):
#: the published table (``lda.py:78-87``).
PZ81 = {"gamma": (-0.1423, -0.0843), "beta1": (1.0529, 1.3981), "beta2": (0.3334, 0.2611),
        "a": (0.0311, 0.01555), "b": (-0.048, -0.0269), "c": (0.0020, 0.0007), "d": (-0.0116, -0.0048)}


def pz81_branch(rs, i):
    """eps_c and d eps_c / d rs of one spin limit: the Pade form for rs >= 1, the
    high-density expansion below (``lda.py:203-234``)."""
    g, b1, b2 = PZ81["gamma"][i], PZ81["beta1"][i], PZ81["beta2"][i]
    a, b, c, d = PZ81["a"][i], PZ81["b"][i], PZ81["c"][i], PZ81["d"][i]
    srs = np.sqrt(rs)
    den = 1.0 + b1 * srs + b2 * rs
    low, dlow = g / den, -g * (0.5 * b1 / srs + b2) / den**2
    lr = np.log(np.maximum(rs, 1e-300))
    high, dhigh = a * lr + b + c * rs * lr + d * rs, a / rs + c * (lr + 1.0) + d
    dense = rs < 1.0
    return np.where(dense, high, low), np.where(dense, dhigh, dlow)


def correlation_pz81(rho):
    """``lda.py:203-244``: the two limits interpolated with f(zeta)."""
    rho = np.atleast_2d(np.asarray(rho, dtype=float))
    n, rs, z = rs_and_zeta(rho)
    (e0, d0), (e1, d1) = pz81_branch(rs, 0), pz81_branch(rs, 1)
    fz, dfz = f_zeta(z), f_zeta_derivative(z)
    return _assemble(e0 + fz * (e1 - e0), d0 + fz * (d1 - d0), dfz * (e1 - e0), rs, z, n > DENSITY_THRESHOLD)


def correlation_pw92(rho, *, modified=False):
    """``lda.py:169-199``."""
    rho = np.atleast_2d(np.asarray(rho, dtype=float))
    n, rs, z = rs_and_zeta(rho)
    eps, d_drs, d_dz = _pw92_local(rs, z, PW92_MOD if modified else PW92)
    return _assemble(eps, d_drs, d_dz, rs, z, n > DENSITY_THRESHOLD)


def gga_exchange(rho, sigma, par):
    """PBE-form exchange, spin-scaled, F_x = 1 + kappa mu s^2/(kappa + mu s^2) (``gga.py:112-156``)."""
    kappa, mu = par["kappa"], par["mu"]
    n = rho[:, 0] + rho[:, 1]
    act = n > DENSITY_THRESHOLD
    spin = np.maximum(rho, 0.0)
    ssig = np.stack([np.maximum(sigma[:, 0], 0.0), np.maximum(sigma[:, 2], 0.0)], axis=-1)
    safe = np.maximum(spin, DENSITY_THRESHOLD)
    n43 = safe ** (4 / 3)
    s2 = (X2S * np.sqrt(ssig) / n43) ** 2
    den = kappa + mu * s2
    fx = 1.0 + kappa * mu * s2 / den
    fxs = 2.0 * kappa * kappa * mu * s2 / den**2
    ce = np.where(spin > DENSITY_THRESHOLD, -X_FACTOR_C * n43 * fx, 0.0)
    eps = np.zeros_like(n)
    np.divide(ce.sum(axis=1), n, out=eps, where=act)
    vrho = -(4 / 3) * X_FACTOR_C * np.cbrt(safe) * (fx - fxs)
    vss = -X_FACTOR_C * kappa * kappa * mu * X2S**2 / (n43 * den**2)
    keep = (spin > DENSITY_THRESHOLD) & act[:, None]
    vrho = np.where(keep, vrho, 0.0)
    vss = np.where(keep, vss, 0.0)
    return np.where(act, eps, 0.0), vrho, np.stack([vss[:, 0], np.zeros_like(n), vss[:, 1]], axis=-1)


def gga_correlation(rho, sigma, par):
    """PBE-form correlation on the modified PW92 local part (``gga.py:160-254``)."""
    beta, gamma = par["beta"], par["gamma"]
    n, rs, z = rs_and_zeta(rho)
    act = n > DENSITY_THRESHOLD
    safe = np.maximum(n, DENSITY_THRESHOLD)
    elda, delda_drs, delda_dz = _pw92_local(rs, z, PW92_MOD)
    opz, omz = np.maximum(1.0 + z, DENSITY_THRESHOLD), np.maximum(1.0 - z, DENSITY_THRESHOLD)
    phi = 0.5 * (opz ** (2 / 3) + omz ** (2 / 3))
    dphi = (1 / 3) * (opz ** (-1 / 3) - omz ** (-1 / 3))
    phi3 = phi**3
    stot = np.maximum(sigma[:, 0] + 2.0 * sigma[:, 1] + sigma[:, 2], 0.0)
    tpx = 1.0 / (4.0 * 2.0 ** (1 / 3) * phi * np.sqrt(rs))
    xt = np.sqrt(stot) / safe ** (4 / 3)
    t = tpx * xt
    t2 = t * t
    et = np.expm1(-elda / (gamma * phi3))
    es = np.where(np.abs(et) < 1e-300, 1e-300, et)
    A = beta / (gamma * es)
    u = t2 + A * t2 * t2
    d = 1.0 + A * u
    w = beta * u / (gamma * d)
    lg = np.log1p(w)
    H = gamma * phi3 * lg
    pre = gamma * phi3 / (1.0 + w)
    dh_dt = pre * beta / (gamma * d**2) * (2.0 * t + 4.0 * A * t2 * t)
    dh_da = pre * beta * (t2 * t2 - u * u) / (gamma * d**2)
    dh_dphi = 3.0 * gamma * phi * phi * lg
    da_de = beta * (et + 1.0) / (gamma**2 * phi3 * es**2)
    da_dphi = -3.0 * beta * elda * (et + 1.0) / (gamma**2 * phi**4 * es**2)
    dh_drs = dh_da * da_de * delda_drs + dh_dt * (-t / (2.0 * rs))
    dh_dz = dh_dphi * dphi + dh_da * (da_dphi * dphi + da_de * delda_dz) + dh_dt * (-t / phi) * dphi
    eps = elda + H
    deps_drs = delda_drs + dh_drs
    deps_dz = delda_dz + dh_dz
    common = eps - (rs / 3.0) * deps_drs - (4 / 3) * xt * dh_dt * tpx
    vrho = np.stack([common + (1.0 - z) * deps_dz, common - (1.0 + z) * deps_dz], axis=-1)
    t2s = tpx**2 / safe ** (8 / 3)
    vst = safe * phi3 * beta * t2s * (2.0 + 4.0 * A * t2) / (2.0 * (1.0 + w) * d**2)
    eps = np.where(act, eps, 0.0)
    vrho = np.where(act[:, None], vrho, 0.0)
    vst = np.where(act, vst, 0.0)
    return eps, vrho, np.stack([vst, 2.0 * vst, vst], axis=-1)


class XCFunctional:
    """``driver.py:67-164`` for LDA-PZ, LDA-PW, PBE and PBEsol."""

    def __init__(self, name: str):
        key = name.strip().upper().replace("_", "-")
        table = {"LDA-PZ": "lda", "LDA-PW": "lda", "PBE": "gga", "PBESOL": "gga"}
        if key not in table:
            raise ValueError(f"unknown functional {name!r}")
        self.kind = table[key]
        self.key = key
        self.name = {"PBESOL": "PBEsol"}.get(key, key)
        self.params = PBESOL if key == "PBESOL" else PBE

    @property
    def is_gga(self) -> bool:
        return self.kind == "gga"

    def evaluate(self, rho, sigma=None):
        rho = np.atleast_2d(np.asarray(rho, dtype=float))
        if self.kind == "lda":
            ex, vx = lda_exchange(rho)
            ec, vc = correlation_pz81(rho) if self.key == "LDA-PZ" else correlation_pw92(rho)
            return ex + ec, vx + vc
        sigma = np.atleast_2d(np.asarray(sigma, dtype=float))
        ex, vx, sx = gga_exchange(rho, sigma, self.params)
        ec, vc, sc = gga_correlation(rho, sigma, self.params)
        return ex + ec, vx + vc, sx + sc

    def evaluate_on_grid(self, density, grid):
        """Returns ``(energy, potential (nspin, *grid))`` (``driver.py:108-164``)."""
        energy, pot, _, _ = self.evaluate_on_grid_details(density, grid)
        return energy, pot

    def evaluate_on_grid_details(self, density, grid):
        """``(energy, potential, gradients (2, *grid, 3) | None, vsigma (3, *grid) | None)``:
        the fields of the reference's ``XCResult`` the stress reads (``driver.py:47-64``)."""
        d = np.asarray(density, dtype=float)
        if d.ndim == 3:
            d = d[None]
        nspin = d.shape[0]
        sd = np.stack([d[0] / 2, d[0] / 2]) if nspin == 1 else d
        npts = int(np.prod(grid.shape))
        rho = sd.reshape(2, npts).T
        grads = None
        if self.is_gga:
            grads = np.stack([grid.gradient(sd[s]) for s in range(2)])
            f = grads.reshape(2, npts, 3)
            sigma = np.stack([np.einsum("pi,pi->p", f[0], f[0]), np.einsum("pi,pi->p", f[0], f[1]),
                              np.einsum("pi,pi->p", f[1], f[1])], axis=-1)
            eps, vrho, vsig = self.evaluate(rho, sigma)
        else:
            eps, vrho = self.evaluate(rho)
            vsig = None
        energy = float(np.sum(eps * (rho[:, 0] + rho[:, 1])) * grid.cell_volume_per_point)
        pot = vrho.T.reshape(2, *grid.shape)
        if vsig is not None:
            vg = vsig.T.reshape(3, *grid.shape)
            for s in range(2):
                flux = 2.0 * vg[2 * s][..., None] * grads[s] + vg[1][..., None] * grads[1 - s]
                pot[s] = pot[s] - grid.divergence(flux)
        if nspin == 1:
            pot = pot[:1]
        vsig_grid = None if vsig is None else vsig.T.reshape(3, *grid.shape)
        return energy, pot, grads, vsig_grid
