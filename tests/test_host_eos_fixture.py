"""CPU checks of the equation-of-state fixture and of the fit the GPU gate uses: the
bundled dataset is the one the DojoReport used (its md5), and the Birch-Murnaghan fit of the
report's own energies returns the report's published V0, B0 and B1
(``tests/physics/test_dojo_eos.py:140-170`` in the reference)."""

from __future__ import annotations

import gzip
import hashlib
import json

import numpy as np
import pytest

from conftest import DATA, GOLDEN
from test_gpu_scf_configs import EV_PER_A3_IN_GPA, delta_gauge, fit_eos


def _ref():
    return json.loads((GOLDEN / "deltafactor_si.json").read_text())["Si"]


def test_bundled_dataset_is_the_one_the_report_used():
    with gzip.open(DATA / "Si.pbe.psp8.gz", "rb") as fh:
        assert hashlib.md5(fh.read()).hexdigest() == _ref()["md5"]


def test_fit_reproduces_the_published_parameters():
    ref = _ref()
    v = np.asarray(ref["volumes_A3"], float) / ref["num_sites"]
    e = np.asarray(ref["etotals_eV"], float) / ref["num_sites"]
    p, residual = fit_eos(v, e)
    assert residual * 1000 < 0.05
    assert p[1] == pytest.approx(ref["v0_A3_per_atom"], rel=1e-4)
    assert p[2] * EV_PER_A3_IN_GPA == pytest.approx(ref["b0_GPa"], rel=1e-3)
    assert p[3] == pytest.approx(ref["b1"], rel=1e-2)
    assert delta_gauge(p, p) == 0.0
    # a rigid shift leaves the delta gauge unchanged; a 1 % softer curve does not
    q = p.copy()
    q[2] *= 0.99
    assert delta_gauge(p, q) > 0.05
