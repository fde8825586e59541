"""What the trained network gains, on cells it was not trained on: the first eight held-out
seeds of the training set (256..263; the network saw 0..255). From the learned start the SCF
converges to the same state as from the atomic superposition (1e-6 Ha/atom), starts from a
smaller density residual on every cell, never needs more iterations, and needs strictly fewer
on average. (All 64 held-out cells, scored once on an MI355X: 13.88 -> 12.77 iterations on
average, 53 cells one fewer, 9 two fewer, 2 equal, none more; wall time -10.7 %.) Skipped when
no trained model is present in the package."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CELLS = list(range(256, 264))


def test_learned_start_same_answer_fewer_iterations():
    from amaterasu_amd.learned import MODEL_PATH, DensityModel, learned_density
    from amaterasu_amd.learned_data import TOL, training_system
    from amaterasu_amd.scf import run_scf

    if not MODEL_PATH.exists():
        pytest.skip("no trained model in amaterasu_amd/data")
    model = DensityModel.load()
    it_a, it_l = [], []
    for cell in CELLS:
        s = training_system(cell)
        atomic = run_scf(s, TOL, density=s.initial_density())
        learned = run_scf(s, TOL, density=learned_density(s, model))
        print(f"cell {cell}: atomic {atomic.iterations} learned {learned.iterations} iterations, first residual "
              f"{atomic.history[0][1]:.3e} -> {learned.history[0][1]:.3e}", flush=True)
        assert atomic.converged and learned.converged
        assert abs(atomic.energies.total - learned.energies.total) / 8 < 1e-6
        assert learned.history[0][1] < atomic.history[0][1]  # smaller first density residual
        assert learned.iterations <= atomic.iterations
        it_a.append(atomic.iterations)
        it_l.append(learned.iterations)
    assert np.mean(it_l) < np.mean(it_a)
