"""Host logic of the SCF loop on CPU tensors against the oracle: every mixer without a
preconditioner (the Kerker filter needs the library transforms) and the occupation solver
for every smearing scheme. The same code runs on device tensors in the GPU tests."""

from __future__ import annotations

import numpy as np
import pytest
import torch


@pytest.mark.parametrize("name", ["linear", "anderson", "pulay", "broyden"])
def test_mixers_on_host_tensors_match_the_oracle(name):
    from amaterasu_amd.scf import make_mixer
    from oracle.scf import make_mixer as omake

    rng = np.random.default_rng(41)
    shape = (2, 6, 5, 7)
    dm, om = make_mixer(name, 0.4, 4), omake(name, 0.4, 4)
    target = 0.05 + np.abs(rng.normal(size=shape))
    x = np.abs(rng.normal(size=shape))
    for _ in range(8):
        out = target + 0.5 * np.tanh(x - target) + 0.05 * rng.normal(size=shape)
        got = dm(torch.from_numpy(x), torch.from_numpy(out)).numpy()
        want = om(x, out)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
        x = want


@pytest.mark.parametrize("scheme,order", [("fixed", 1), ("fermi-dirac", 1), ("gaussian", 1),
                                          ("marzari-vanderbilt", 1), ("methfessel-paxton", 1),
                                          ("methfessel-paxton", 3)])
def test_occupations_on_host_tensors_match_the_oracle(scheme, order):
    from amaterasu_amd.scf import occupations
    from oracle.scf import compute_occupations

    rng = np.random.default_rng(9)
    eig = np.sort(rng.normal(scale=0.3, size=(1, 3, 10)), axis=-1)
    w = np.array([0.2, 0.3, 0.5])
    smearing = 0.0 if scheme == "fixed" else 0.015
    occ, mu, ent, gap = occupations(torch.from_numpy(eig), torch.from_numpy(w), 11.0, scheme, smearing, 1,
                                    order=order)
    ref = compute_occupations(eig, w, 11.0, scheme=scheme, smearing=smearing, order=order)
    assert np.abs(occ.numpy() - ref.occupations).max() < 1e-9
    assert mu == pytest.approx(ref.fermi_level, abs=1e-11)
    assert ent == pytest.approx(ref.entropy_term, abs=1e-12)
