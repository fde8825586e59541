"""Pins of the oracle's PBE / PBEsol gradient terms (``xc/gga.py:112-254``).

Every check here is built from a definition that does not go through the oracle's own
variables: the reduced gradients s and t are formed from the textbook expressions
(Perdew, Burke & Ernzerhof, PRL 77, 3865 (1996)), not from the oracle's X2S or its
t-prefactor, and the expected values come from the functional's published limits:

* the parameter relations the reference asserts (``tests/unit/test_xc_functionals.py:139-150``):
  mu = beta pi^2 / 3, gamma = (1 - ln 2) / pi^2, PBEsol mu = 10/81, beta = 0.046, kappa = 0.804;
* exchange: F_x(s) = 1 + mu s^2 + O(s^4) at small s, F_x(s) -> 1 + kappa at large s, and the
  exact spin-scaling relation E_x[n_up, n_dn] = (E_x[2 n_up] + E_x[2 n_dn]) / 2;
* correlation: H = beta phi^3 t^2 + O(t^4) at small t (with phi of the spin polarisation),
  eps_c^unif + H -> 0 as t -> infinity, and spin-flip symmetry;
* the hydrogen atom, n = exp(-2r)/pi, fully polarised: PBE exchange -0.3059 Ha and
  correlation -0.0060 Ha, LSD -0.2680 / -0.0222 Ha (PBE 1996, Table I, to the four
  decimals published).

A planted 1 % error in mu, beta, gamma or kappa makes at least one of them fail
(``test_a_planted_parameter_error_is_caught``).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import xc

PARAMS = {"PBE": xc.PBE, "PBEsol": xc.PBESOL}


def _sigma_for_s(n_total: np.ndarray, s: np.ndarray) -> np.ndarray:
    """|grad n|^2 of an unpolarised density n with reduced gradient s = |grad n| / (2 k_F n),
    k_F = (3 pi^2 n)^(1/3)."""
    kf = np.cbrt(3.0 * np.pi**2 * n_total)
    return (2.0 * kf * n_total * s) ** 2


def _sigma_for_t(n_total: np.ndarray, zeta: float, t: np.ndarray) -> np.ndarray:
    """|grad n|^2 for the correlation reduced gradient t = |grad n| / (2 phi k_s n),
    k_s = sqrt(4 k_F / pi), phi = ((1+zeta)^(2/3) + (1-zeta)^(2/3)) / 2."""
    kf = np.cbrt(3.0 * np.pi**2 * n_total)
    ks = np.sqrt(4.0 * kf / np.pi)
    phi = 0.5 * ((1.0 + zeta) ** (2 / 3) + (1.0 - zeta) ** (2 / 3))
    return (2.0 * phi * ks * n_total * t) ** 2


def _spin_sigma(total_sigma: np.ndarray, zeta: float) -> np.ndarray:
    """(sigma_uu, sigma_ud, sigma_dd) for spin densities that are fixed fractions of n:
    grad n_up = (1+zeta)/2 grad n, grad n_dn = (1-zeta)/2 grad n."""
    a, b = 0.5 * (1.0 + zeta), 0.5 * (1.0 - zeta)
    return np.stack([a * a * total_sigma, a * b * total_sigma, b * b * total_sigma], axis=-1)


def _rho(n_total: np.ndarray, zeta: float) -> np.ndarray:
    return np.stack([0.5 * (1.0 + zeta) * n_total, 0.5 * (1.0 - zeta) * n_total], axis=-1)


@pytest.mark.parametrize("name", ["PBE", "PBEsol"])
def test_parameters_match_the_published_relations(name):
    p = PARAMS[name]
    assert p["kappa"] == 0.804
    assert p["gamma"] == pytest.approx((1.0 - np.log(2.0)) / np.pi**2, rel=1e-15)
    if name == "PBE":
        assert p["mu"] == pytest.approx(p["beta"] * np.pi**2 / 3.0, rel=1e-15)
    else:
        assert p["mu"] == pytest.approx(10.0 / 81.0, rel=1e-15)
        assert p["beta"] == 0.046
    # gamma is the same constant as the modified PW92 A_0 (to the digits PW92 publishes)
    assert xc.PW92_MOD["a"][0] == pytest.approx(p["gamma"], rel=1e-6)


@pytest.mark.parametrize("name", ["PBE", "PBEsol"])
@pytest.mark.parametrize("n", [0.003, 0.1, 2.0])
def test_exchange_small_gradient_slope_is_mu(name, n):
    """(F_x - 1) / s^2 -> mu, with the O(s^4) term -mu^2/kappa s^4 the only correction."""
    p = PARAMS[name]
    nt = np.full(4, n)
    s = np.array([1e-3, 3e-3, 1e-2, 3e-2])
    sig = _spin_sigma(_sigma_for_s(nt, s), 0.0)
    rho = _rho(nt, 0.0)
    fx = xc.gga_exchange(rho, sig, p)[0] / xc.lda_exchange(rho)[0]
    slope = (fx - 1.0) / s**2
    assert slope == pytest.approx(p["mu"] - p["mu"] ** 2 / p["kappa"] * s**2, rel=1e-7)


@pytest.mark.parametrize("name", ["PBE", "PBEsol"])
def test_exchange_large_gradient_limit_is_one_plus_kappa(name):
    p = PARAMS[name]
    nt = np.array([0.01, 0.5])
    s = np.array([1e4, 1e4])
    rho = _rho(nt, 0.0)
    fx = xc.gga_exchange(rho, _spin_sigma(_sigma_for_s(nt, s), 0.0), p)[0] / xc.lda_exchange(rho)[0]
    # F_x = 1 + kappa - kappa^2 / (kappa + mu s^2)
    assert fx == pytest.approx(1.0 + p["kappa"], rel=1e-7)


@pytest.mark.parametrize("name", ["PBE", "PBEsol"])
def test_exchange_obeys_spin_scaling(name):
    """E_x[n_up, n_dn] = (E_x[2 n_up, unpolarised] + E_x[2 n_dn, unpolarised]) / 2, with
    the gradients scaled alike (exact for exchange)."""
    p = PARAMS[name]
    rng = np.random.default_rng(5)
    nu, nd = rng.uniform(0.01, 1.0, 6), rng.uniform(0.01, 1.0, 6)
    gu, gd = rng.uniform(0.0, 0.6, 6), rng.uniform(0.0, 0.6, 6)  # |grad n_up|^2, |grad n_dn|^2
    rho = np.stack([nu, nd], axis=-1)
    e_pol = xc.gga_exchange(rho, np.stack([gu, np.zeros(6), gd], axis=-1), p)[0] * (nu + nd)

    def unpol(ns, gs):  # density 2 n_s, |grad(2 n_s)|^2 = 4 gs, as two equal channels
        r = np.stack([ns, ns], axis=-1)
        return xc.gga_exchange(r, np.stack([gs, gs, gs], axis=-1), p)[0] * (2.0 * ns)

    assert e_pol == pytest.approx(0.5 * (unpol(nu, gu) + unpol(nd, gd)), rel=1e-13)


@pytest.mark.parametrize("name", ["PBE", "PBEsol"])
@pytest.mark.parametrize("zeta", [0.0, 0.4, -0.7])
def test_correlation_small_t_is_beta_phi3_t2(name, zeta):
    p = PARAMS[name]
    nt = np.full(4, 0.2)
    t = np.array([1e-4, 3e-4, 1e-3, 3e-3])
    rho = _rho(nt, zeta)
    sig = _spin_sigma(_sigma_for_t(nt, zeta, t), zeta)
    h = xc.gga_correlation(rho, sig, p)[0] - xc.correlation_pw92(rho, modified=True)[0]
    phi = 0.5 * ((1.0 + zeta) ** (2 / 3) + (1.0 - zeta) ** (2 / 3))
    ratio = h / (phi**3 * t**2)
    # H = beta phi^3 t^2 (1 + O(t^2)): the first two points agree with beta to 1e-5
    assert ratio[:2] == pytest.approx(p["beta"], rel=2e-5)
    assert np.all(np.diff(np.abs(ratio - p["beta"])) > 0)  # the deviation grows with t


@pytest.mark.parametrize("name", ["PBE", "PBEsol"])
@pytest.mark.parametrize("zeta", [0.0, 0.5])
def test_correlation_vanishes_at_large_t(name, zeta):
    """H -> -eps_c^unif as t -> infinity (PBE's rapidly-varying limit)."""
    p = PARAMS[name]
    nt = np.array([0.01, 0.3, 3.0])
    rho = _rho(nt, zeta)
    sig = _spin_sigma(_sigma_for_t(nt, zeta, np.full(3, 1e5)), zeta)
    eps = xc.gga_correlation(rho, sig, p)[0]
    lda = xc.correlation_pw92(rho, modified=True)[0]
    assert np.abs(eps).max() < 1e-9 * np.abs(lda).max()


@pytest.mark.parametrize("name", ["PBE", "PBEsol"])
def test_correlation_is_spin_flip_symmetric(name):
    p = PARAMS[name]
    rng = np.random.default_rng(9)
    rho = rng.uniform(0.01, 1.0, size=(8, 2))
    sig = rng.uniform(0.0, 0.5, size=(8, 3))
    sig[:, 1] = 0.4 * np.sqrt(sig[:, 0] * sig[:, 2])
    e, v, vs = xc.gga_correlation(rho, sig, p)
    e2, v2, vs2 = xc.gga_correlation(rho[:, ::-1].copy(), sig[:, ::-1].copy(), p)
    assert e2 == pytest.approx(e, rel=1e-13)
    assert v2[:, ::-1] == pytest.approx(v, rel=1e-12, abs=1e-14)
    assert vs2[:, ::-1] == pytest.approx(vs, rel=1e-12, abs=1e-14)


def _hydrogen(p):
    """Exchange and correlation energies of n = exp(-2r)/pi, fully polarised, by radial
    quadrature (|grad n| = 2n)."""
    r = np.linspace(1e-6, 40.0, 400001)
    n = np.exp(-2.0 * r) / np.pi
    rho = np.stack([n, np.zeros_like(n)], axis=-1)
    g2 = (2.0 * n) ** 2
    sig = np.stack([g2, np.zeros_like(n), np.zeros_like(n)], axis=-1)
    w = 4.0 * np.pi * r**2 * n
    ex = np.trapezoid(xc.gga_exchange(rho, sig, p)[0] * w, r)
    ec = np.trapezoid(xc.gga_correlation(rho, sig, p)[0] * w, r)
    ex_lda = np.trapezoid(xc.lda_exchange(rho)[0] * w, r)
    ec_lda = np.trapezoid(xc.correlation_pw92(rho, modified=True)[0] * w, r)
    return ex, ec, ex_lda, ec_lda


def test_hydrogen_atom_matches_the_published_table():
    """PBE 1996 Table I (four decimals): H atom E_x LSD -0.2680, PBE -0.3059 (exact -0.3125);
    E_c LSD -0.0222, PBE -0.0060 (exact 0)."""
    ex, ec, ex_lda, ec_lda = _hydrogen(xc.PBE)
    assert ex_lda == pytest.approx(-0.2680, abs=6e-5)
    assert ec_lda == pytest.approx(-0.0222, abs=6e-5)
    assert ex == pytest.approx(-0.3059, abs=6e-5)
    assert ec == pytest.approx(-0.0060, abs=6e-5)


@pytest.mark.parametrize("key", ["mu", "beta", "gamma", "kappa"])
def test_a_planted_parameter_error_is_caught(key):
    """The checks above fail when one parameter is off by 1 % (the planted error lives in a
    copy of the oracle's table, passed in place of the real one)."""
    bad = dict(xc.PBE)
    bad[key] *= 1.01
    failed = []
    nt = np.full(2, 0.1)
    rho = _rho(nt, 0.0)
    s = np.array([1e-3, 3e-3])
    fx = xc.gga_exchange(rho, _spin_sigma(_sigma_for_s(nt, s), 0.0), bad)[0] / xc.lda_exchange(rho)[0]
    if not np.allclose((fx - 1.0) / s**2, xc.PBE["mu"] - xc.PBE["mu"] ** 2 / xc.PBE["kappa"] * s**2, rtol=1e-7):
        failed.append("exchange slope")
    big = xc.gga_exchange(rho, _spin_sigma(_sigma_for_s(nt, np.full(2, 1e4)), 0.0), bad)[0]
    if not np.allclose(big / xc.lda_exchange(rho)[0], 1.0 + xc.PBE["kappa"], rtol=1e-7):
        failed.append("exchange bound")
    t = np.array([1e-4, 3e-4])
    h = xc.gga_correlation(rho, _spin_sigma(_sigma_for_t(nt, 0.0, t), 0.0), bad)[0]
    h = h - xc.correlation_pw92(rho, modified=True)[0]
    if not np.allclose(h / t**2, xc.PBE["beta"], rtol=2e-5):
        failed.append("correlation slope")
    if not (bad["mu"] == pytest.approx(bad["beta"] * np.pi**2 / 3, rel=1e-15)
            and bad["gamma"] == pytest.approx((1 - np.log(2)) / np.pi**2, rel=1e-15) and bad["kappa"] == 0.804):
        failed.append("parameter relations")
    ex, ec, _, _ = _hydrogen(bad)
    if abs(ex + 0.3059) > 6e-5 or abs(ec + 0.0060) > 6e-5:
        failed.append("hydrogen")
    assert failed, f"a 1 % error in {key} passed every check"
