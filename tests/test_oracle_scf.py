"""Oracle XC, Hartree, occupations, mixing and SCF loop, pinned by closed forms, exact
derivative identities and the reference's SCF invariants (tests/unit/test_scf.py:59-89,
test_hamiltonian.py:27-65, test_xc_functionals.py, test_occupations.py)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import xc


def test_lda_exchange_of_the_uniform_gas():
    n = np.array([1e-3, 0.01, 0.3, 2.0])
    eps, v = xc.lda_exchange(np.stack([n / 2, n / 2], axis=-1))
    assert eps == pytest.approx(-0.75 * (3 / np.pi) ** (1 / 3) * n ** (1 / 3), rel=1e-14)
    assert v[:, 0] == pytest.approx((4 / 3) * eps, rel=1e-14)


@pytest.mark.parametrize("name", ["LDA-PZ", "LDA-PW", "PBE", "PBEsol"])
def test_potential_is_the_derivative_of_the_energy(name):
    """v_s = d(n eps)/dn_s and vsigma = d(n eps)/dsigma by central differences: pins every
    derivative chain in the restatement independently of the energy formula."""
    f = xc.XCFunctional(name)
    rng = np.random.default_rng(0)
    rho = rng.uniform(0.01, 1.0, size=(20, 2))
    sig = rng.uniform(0.0, 0.5, size=(20, 3))
    sig[:, 1] = 0.3 * np.sqrt(sig[:, 0] * sig[:, 2])

    def energy(r, s):
        out = f.evaluate(r, s) if f.is_gga else f.evaluate(r)
        return out[0] * (r[:, 0] + r[:, 1])

    out = f.evaluate(rho, sig) if f.is_gga else f.evaluate(rho)
    h = 1e-6
    for s in range(2):
        dp, dm = rho.copy(), rho.copy()
        dp[:, s] += h
        dm[:, s] -= h
        fd = (energy(dp, sig) - energy(dm, sig)) / (2 * h)
        assert out[1][:, s] == pytest.approx(fd, rel=1e-6, abs=1e-8)
    if f.is_gga:
        for c in range(3):
            sp, sm = sig.copy(), sig.copy()
            sp[:, c] += h
            sm[:, c] -= h
            fd = (energy(rho, sp) - energy(rho, sm)) / (2 * h)
            assert out[2][:, c] == pytest.approx(fd, rel=1e-5, abs=1e-8)


def test_pz81_against_pw92_and_its_own_branches():
    """PZ81 (lda.py:203-244): its two branches meet at rs = 1 (the published coefficients
    make the value continuous to 3.2e-5 Ha); it agrees with PW92 -- an independent fit of the
    same Ceperley-Alder data -- to 1.5 mHa over rs 0.5..6 in both spin limits; and for
    large rs it approaches the Pade limit gamma / (beta2 rs), with the relative correction
    beta1 / (beta2 sqrt(rs))."""
    rs = np.linspace(0.5, 6.0, 23)
    n = 3.0 / (4.0 * np.pi * rs**3)
    for pol in (0, 1):
        rho = np.stack([n, np.zeros_like(n)], axis=1) if pol else np.stack([n / 2, n / 2], axis=1)
        pz, _ = xc.correlation_pz81(rho)
        pw, _ = xc.correlation_pw92(rho)
        assert np.abs(pz - pw).max() < 1.5e-3
        lo, _ = xc.pz81_branch(np.array([1.0 + 1e-12]), pol)
        hi, _ = xc.pz81_branch(np.array([1.0 - 1e-12]), pol)
        assert abs(lo[0] - hi[0]) < 4e-5
    big = np.array([1e8])
    e, _ = xc.pz81_branch(big, 0)
    assert e[0] == pytest.approx(xc.PZ81["gamma"][0] / (xc.PZ81["beta2"][0] * big[0]), rel=5e-4)


def test_pbe_reduces_to_lda_without_gradient():
    rng = np.random.default_rng(1)
    rho = rng.uniform(0.01, 1.0, size=(10, 2))
    e_gga = xc.gga_exchange(rho, np.zeros((10, 3)), xc.PBE)[0] + xc.gga_correlation(rho, np.zeros((10, 3)), xc.PBE)[0]
    e_lda = xc.lda_exchange(rho)[0] + xc.correlation_pw92(rho, modified=True)[0]
    assert e_gga == pytest.approx(e_lda, rel=1e-13)


def test_pbe_exchange_enhancement_bound():
    """F_x -> 1 + kappa as s -> infinity (the Lieb-Oxford-motivated bound)."""
    n = 0.1
    rho = np.array([[n / 2, n / 2]])
    big = np.array([[1e8, 0.0, 1e8]])
    eps = xc.gga_exchange(rho, big, xc.PBE)[0]
    lda = xc.lda_exchange(rho)[0]
    assert eps / lda == pytest.approx(1.804, rel=1e-6)


def test_gga_grid_potential_is_the_functional_derivative(si2_system):
    """dE_xc/dn(r) by finite differences of the grid energy equals v_xc(r) (the divergence
    term of xc/driver.py:147-154 included)."""
    f = xc.XCFunctional("PBE")
    grid = si2_system.grid
    rho = si2_system.initial_density()
    e0, v = f.evaluate_on_grid(rho, grid)
    dv = grid.cell_volume_per_point
    for idx in [(0, 0, 0), (3, 5, 7), (10, 10, 10)]:
        h = 1e-5 * rho[(0, *idx)]
        p, m = rho.copy(), rho.copy()
        p[(0, *idx)] += h
        m[(0, *idx)] -= h
        fd = (f.evaluate_on_grid(p, grid)[0] - f.evaluate_on_grid(m, grid)[0]) / (2 * h * dv)
        assert v[(0, *idx)] == pytest.approx(fd, rel=1e-6)


def test_hartree_of_gaussians_is_the_ewald_reciprocal_sum():
    """test_hamiltonian.py:27-65: E_H of normalised Gaussians = E_Ewald + eta/sqrt(pi) sum q^2."""
    from oracle.basis import FFTGrid, GSphere
    from oracle.ewald import ewald_energy
    from oracle.scf import hartree

    side = 12.0
    cell = side * np.eye(3)
    pos = np.array([[3.0, 3.0, 3.0], [9.0, 3.0, 3.0]])
    q = np.array([1.0, -1.0])
    alpha = 2.0
    eta = np.sqrt(alpha / 2.0)
    grid = FFTGrid.for_cutoff(cell, 120.0)
    sphere = GSphere.build(cell, 120.0)
    pts = grid.points_cartesian()
    dens = np.zeros(grid.shape)
    for shift in np.stack(np.meshgrid(*[np.arange(-1, 2) * side] * 3, indexing="ij"), axis=-1).reshape(-1, 3):
        for c, p in zip(q, pos):
            dens += c * (alpha / np.pi) ** 1.5 * np.exp(-alpha * np.sum((pts - p - shift) ** 2, axis=-1))
    e, v = hartree(grid, sphere, dens)
    assert e == pytest.approx(ewald_energy(cell, pos, q, eta=eta) + eta / np.sqrt(np.pi) * np.sum(q**2), abs=1e-8)
    assert 0.5 * grid.integrate(v * dens) == pytest.approx(e, rel=1e-9)


@pytest.mark.parametrize("scheme", ["fermi-dirac", "gaussian", "marzari-vanderbilt", "methfessel-paxton"])
def test_occupations_hold_the_electron_count(scheme):
    from oracle.scf import compute_occupations

    rng = np.random.default_rng(2)
    eig = np.sort(rng.uniform(-1, 1, size=(1, 4, 12)), axis=-1)
    w = np.array([0.1, 0.2, 0.3, 0.4])
    r = compute_occupations(eig, w, 9.3, scheme=scheme, smearing=0.02)
    assert 2.0 * np.einsum("k,skn->", w, r.occupations) == pytest.approx(9.3, abs=1e-9)
    fixed = compute_occupations(eig, w, 8.0, scheme="fixed")
    assert 2.0 * np.einsum("k,skn->", w, fixed.occupations) == pytest.approx(8.0, abs=1e-12)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_methfessel_paxton_moments_and_entropy(order):
    """Methfessel-Paxton of order N (PRB 40, 3616; occupations.py:47-111): the smeared delta
    -f' integrates to one and its moments x^2 .. x^{2N} vanish (x^{2N+2} does not); N = 0
    would be the Gaussian; the generalised entropy obeys S' = x f' and vanishes far from
    the Fermi level. A wrong sign of A_n or a shifted Hermite index breaks the moments."""
    from oracle.scf import entropy_function, occupation_function

    x = np.linspace(-7.0, 7.0, 28001)
    h = x[1] - x[0]
    f = occupation_function(x, "methfessel-paxton", order)
    ent = entropy_function(x, "methfessel-paxton", order)
    df = np.gradient(f, h)
    assert np.trapezoid(-df, x) == pytest.approx(1.0, abs=1e-9)
    for k in range(1, order + 1):
        assert abs(np.trapezoid(-df * x ** (2 * k), x)) < 1e-6
    assert abs(np.trapezoid(-df * x ** (2 * order + 2), x)) > 0.1
    assert np.abs(np.gradient(ent, h) - x * df).max() < 1e-6
    assert abs(ent[0]) < 1e-12 and abs(ent[-1]) < 1e-12
    gauss = occupation_function(x, "gaussian")
    assert np.abs(occupation_function(x, "methfessel-paxton", 0) - gauss).max() < 1e-15


def test_pulay_reaches_the_fixed_point_of_a_linear_map():
    """A contraction x -> A x + b: Pulay must land on (I - A)^-1 b (test_mixing.py:21-60)."""
    from oracle.scf import PulayMixer

    rng = np.random.default_rng(3)
    n = 30
    q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    a = q @ np.diag(rng.uniform(-0.8, 0.8, n)) @ q.T
    b = rng.normal(size=n)
    star = np.linalg.solve(np.eye(n) - a, b)
    x = np.zeros((1, n))
    mix = PulayMixer(beta=0.4, history=8)
    for _ in range(60):
        x = mix(x, a @ x[0] + b)
        x = np.asarray(x).reshape(1, n)
    assert np.abs(x[0] - star).max() < 1e-8


def test_scf_invariants(si2_system):
    """test_scf.py:59-89: converges, electron count exact, eigenvalue-sum identity < 1e-9 eV."""
    import copy

    from oracle.scf import SCFSettings, run_scf
    from oracle.units import HARTREE_IN_EV

    s = copy.copy(si2_system)
    s.xc = xc.XCFunctional("PBE")
    r = run_scf(s, SCFSettings(tol_energy=1e-9, tol_density=1e-8))
    assert r.converged and r.iterations < 30
    assert r.electron_count == pytest.approx(8.0, abs=1e-9)
    assert abs(r.history[-1][2]) * HARTREE_IN_EV < 1e-9
    energies = [h[0] for h in r.history]
    assert energies[-1] == pytest.approx(min(energies), abs=1e-12)


def test_spin_kerker_filters_the_charge_only(si2_system):
    """mixing.py:49-106 for two spin channels: equal channels (no moment) are filtered as the
    total charge, opposite channels (pure moment) pass through unchanged, and the G = 0
    component of the charge is removed."""
    from oracle.scf import Kerker

    grid = si2_system.grid
    rng = np.random.default_rng(21)
    r = rng.normal(size=(1, *grid.shape))
    k = Kerker(grid, 1.5)
    same = k(np.concatenate([r, r]))
    one = k(2 * r)[0] / 2
    assert np.abs(same[0] - one).max() < 1e-13 and np.abs(same[1] - one).max() < 1e-13
    opposite = np.concatenate([r, -r])
    assert np.abs(k(opposite) - opposite).max() < 1e-13
    assert abs(k(r)[0].mean()) < 1e-13


def test_broyden_and_anderson_reach_the_fixed_point_of_a_linear_map():
    """Johnson's Broyden (mixing.py:217-268) and Anderson (Pulay over two iterations) on
    x -> b + A x with a contraction A: both converge to (I - A)^{-1} b, Broyden in fewer
    steps than linear mixing."""
    from oracle.scf import make_mixer

    rng = np.random.default_rng(6)
    n = 40
    q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    a = (q * rng.uniform(-0.8, 0.8, size=n)) @ q.T
    b = rng.normal(size=n)
    fixed = np.linalg.solve(np.eye(n) - a, b)

    def run(name, steps):
        m = make_mixer(name, beta=0.4, history=8)
        x = np.zeros(n)
        for _ in range(steps):
            x = m(x, b + a @ x)
        return np.abs(x - fixed).max()

    assert run("broyden", 40) < 1e-8
    assert run("anderson", 80) < 1e-5
    assert run("broyden", 25) < run("linear", 25)
    with pytest.raises(ValueError):
        make_mixer("simplex")
