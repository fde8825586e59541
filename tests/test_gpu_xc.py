"""The device grid-side operators against the oracle on the Si2 grid: exchange-correlation
energies and potentials (LDA-PZ, LDA-PW, PBE, PBEsol; spin-unpolarised and polarised;
``xc/driver.py:108-164``) and the Hartree energy and potential (``hartree.py:28-50``)."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _density(shape, nspin, seed):
    rng = np.random.default_rng(seed)
    i, j, k = np.meshgrid(*[np.arange(n) / n for n in shape], indexing="ij")
    base = 0.05 + 0.04 * np.cos(2 * np.pi * (i + j)) ** 2 + 0.03 * np.sin(2 * np.pi * k) ** 2
    base = base + 0.002 * rng.random(shape)
    if nspin == 1:
        return base[None]
    return np.stack([0.6 * base, 0.4 * base])


@pytest.mark.parametrize("name", ["LDA-PZ", "LDA-PW", "PBE", "PBEsol"])
@pytest.mark.parametrize("nspin", [1, 2])
def test_xc_on_the_grid_matches_the_oracle(si2_system, name, nspin):
    import torch

    from amaterasu_amd.xc import GridOps, XCFunctional
    from oracle.xc import XCFunctional as OXC

    grid = si2_system.grid
    rho = _density(grid.shape, nspin, 3)
    e_ref, v_ref = OXC(name).evaluate_on_grid(rho, grid)
    e, v = XCFunctional(name).evaluate_on_grid(torch.from_numpy(rho).cuda(), GridOps(grid.shape, grid.cell))
    v = v.cpu().numpy()
    assert e == pytest.approx(e_ref, rel=1e-12)
    assert v.shape == v_ref.shape
    assert np.abs(v - v_ref).max() < 1e-11 * np.abs(v_ref).max()


def test_hartree_matches_the_oracle(si2_system):
    import torch

    from amaterasu_amd.scf import GridSide
    from oracle.scf import hartree

    rho = _density(si2_system.grid.shape, 1, 4)[0]
    e_ref, v_ref = hartree(si2_system.grid, si2_system.density_sphere, rho)
    e, v = GridSide(si2_system).hartree(torch.from_numpy(rho).cuda())
    assert e == pytest.approx(e_ref, rel=1e-12)
    assert np.abs(v.cpu().numpy() - v_ref).max() < 1e-12 * np.abs(v_ref).max()
