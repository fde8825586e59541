"""What the trained network gains, on cells it was not trained on (the held-out seeds
256..319 of the training set): the SCF from the learned start converges to the same state
as from the atomic superposition, and needs fewer iterations, starting from a smaller
density residual. Skipped when no trained model is present in the package."""

from __future__ import annotations

import pytest

pytestmark = pytest.mark.gpu


def _model():
    from amaterasu_amd.learned import MODEL_PATH, DensityModel

    if not MODEL_PATH.exists():
        pytest.skip("no trained model in amaterasu_amd/data")
    return DensityModel.load()


@pytest.mark.parametrize("cell", [300, 311])
def test_learned_start_same_answer_fewer_iterations(cell):
    from amaterasu_amd.learned import learned_density
    from amaterasu_amd.learned_data import TOL, training_system
    from amaterasu_amd.scf import run_scf

    model = _model()
    s = training_system(cell)
    atomic = run_scf(s, TOL, density=s.initial_density())
    learned = run_scf(s, TOL, density=learned_density(s, model))
    assert atomic.converged and learned.converged
    assert abs(atomic.energies.total - learned.energies.total) / 8 < 1e-6
    assert learned.history[0][1] < atomic.history[0][1]  # smaller first density residual
    assert learned.iterations <= atomic.iterations
