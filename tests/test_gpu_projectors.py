"""The device projector object's operator methods (the surface of ``pseudo/kb.py:216-260``:
project, apply with and without a coupling override, energy, occupancies, completeness)
against the oracle's on the same Si2 k-point, for numpy input (numpy out) and device input
(device out, FP64 and FP32)."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import si_primitive_bohr

pytestmark = pytest.mark.gpu


def _pair(si_pseudo):
    from amaterasu_amd.system import build_system, load_dataset
    from oracle.system import build_system as obuild

    cell, pos = si_primitive_bohr()
    kpts, wts = [[0.1, -0.2, 0.3]], [1.0]
    o = obuild(cell, pos, ["Si", "Si"], {"Si": si_pseudo}, 8.0, 32.0, kpts, wts)
    s = build_system(cell, pos, ["Si", "Si"], {"Si": load_dataset()}, 32.0, kpts, 8.0, weights=wts,
                     grid_shape=o.grid.shape)
    return o.projectors[0], s.projectors[0]


def test_projector_methods_match_the_oracle(si_pseudo):
    import torch

    ok, dk = _pair(si_pseudo)
    assert len(ok) == len(dk) and [tuple(x) for x in dk.labels] == [tuple(x) for x in ok.labels]
    npw = ok.matrix.shape[1]
    rng = np.random.default_rng(12)
    b = rng.normal(size=(7, npw)) + 1j * rng.normal(size=(7, npw))
    w = rng.uniform(0.0, 2.0, size=7)
    coupling = ok.dij + 0.1 * np.diag(rng.normal(size=len(ok)))

    def close(got, want, tol=1e-12):
        assert got.shape == want.shape
        assert np.abs(got - want).max() <= tol * max(np.abs(want).max(), 1e-300)

    close(dk.project(b), ok.project(b))
    close(dk.project(b[3]), ok.project(b[3]))
    close(dk.apply(b), ok.apply(b))
    close(dk.apply(b, coupling), ok.apply(b, coupling))
    assert dk.energy(b, w) == pytest.approx(ok.energy(b, w), rel=1e-12)
    assert dk.energy(b, w, coupling) == pytest.approx(ok.energy(b, w, coupling), rel=1e-12)
    close(dk.occupancies(b, w), ok.occupancies(b, w))
    close(dk.completeness(), ok.completeness())
    # device tensors stay on the device
    tb = torch.from_numpy(b).cuda()
    got = dk.apply(tb)
    assert isinstance(got, torch.Tensor) and got.is_cuda and got.dtype == torch.complex128
    close(got.cpu().numpy(), ok.apply(b))
    got32 = dk.apply(tb.to(torch.complex64))
    assert got32.dtype == torch.complex64
    close(got32.cpu().numpy().astype(complex), ok.apply(b), tol=5e-6)
    with pytest.raises(ValueError):
        dk.project(b[:, :-1])
