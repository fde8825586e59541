"""Pins of the learned-start oracle: features against closed forms and brute force, the
backward pass against central finite differences and against torch autograd (float64),
and the bfloat16 rounding against torch's conversion."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import mlp


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(size=4000) * 10.0 ** rng.integers(-6, 6, size=4000),
                        [0.0, -0.0, 1.0, 1.0 + 2.0**-8, 1.0 + 3 * 2.0**-9, 65504.0]])
    want = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mlp.bf16_round(x), want)


def test_single_atom_features_closed_form():
    """One atom in a cubic cell wider than 2 r_c: only the home image contributes."""
    a = 3.0 * mlp.R_CUT
    cell = a * np.eye(3)
    pts = np.array([[0.3, 0.0, 0.0], [1.7, 2.1, -0.4], [0.0, 0.0, mlp.R_CUT * 0.999], [0.0, 0.0, mlp.R_CUT * 1.001]])
    got = mlp.rbf_features(cell, [[0.0, 0.0, 0.0]], pts)
    mu, gamma = mlp.rbf_centres()
    for p, row in zip(pts, got):
        d = np.linalg.norm(p)
        fc = 0.5 * (np.cos(np.pi * d / mlp.R_CUT) + 1.0) if d < mlp.R_CUT else 0.0
        np.testing.assert_allclose(row, fc * np.exp(-gamma * (d - mu) ** 2), rtol=1e-13, atol=1e-300)


def test_features_brute_force_and_invariance():
    """A small skewed cell against an explicit wide image sum; translating every atom and
    point together, or swapping atoms, leaves the features unchanged."""
    rng = np.random.default_rng(5)
    cell = np.array([[6.0, 0.4, 0.0], [1.1, 5.5, 0.3], [0.2, -0.6, 6.3]])
    pos = rng.uniform(size=(3, 3)) @ cell
    pts = rng.uniform(size=(7, 3)) @ cell
    got = mlp.rbf_features(cell, pos, pts)
    mu, gamma = mlp.rbf_centres()
    want = np.zeros_like(got)
    for i in range(-4, 5):
        for j in range(-4, 5):
            for k in range(-4, 5):
                for t in pos:
                    d = np.linalg.norm(pts - t - np.array([i, j, k]) @ cell, axis=1)
                    fc = np.where(d < mlp.R_CUT, 0.5 * (np.cos(np.pi * d / mlp.R_CUT) + 1.0), 0.0)
                    want += fc[:, None] * np.exp(-gamma * (d[:, None] - mu) ** 2)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
    shift = np.array([0.7, -2.2, 9.1])
    np.testing.assert_allclose(mlp.rbf_features(cell, pos[::-1] + shift, pts + shift), got, rtol=1e-12, atol=1e-14)


def _random_case(seed=1, npts=37):
    rng = np.random.default_rng(seed)
    p = mlp.init_params(seed)
    for key in ("b1", "b2", "b3", "b4"):
        p[key] = rng.normal(scale=0.1, size=p[key].shape)
    p["mean"] = rng.normal(size=mlp.N_RBF) * 0.1
    p["std"] = 0.5 + rng.uniform(size=mlp.N_RBF)
    feats = rng.uniform(0, 2, size=(npts, mlp.N_RBF))
    target = rng.normal(size=npts)
    return p, feats, target


def test_backward_matches_finite_differences():
    p, feats, target = _random_case()
    y, cache = mlp.forward(p, feats, bf16=False)
    loss, grads = mlp.backward(p, cache, target, bf16=False)
    rng = np.random.default_rng(11)
    h = 1e-6
    for key in ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4"):
        for _ in range(4):
            idx = tuple(rng.integers(0, n) for n in p[key].shape)
            q = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in p.items()}
            q[key][idx] += h
            lp = np.mean((mlp.forward(q, feats, bf16=False)[0] - target) ** 2)
            q[key][idx] -= 2 * h
            lm = np.mean((mlp.forward(q, feats, bf16=False)[0] - target) ** 2)
            assert grads[key][idx] == pytest.approx((lp - lm) / (2 * h), rel=1e-6, abs=1e-10)


def test_forward_backward_match_torch_autograd():
    p, feats, target = _random_case(seed=2)
    y, cache = mlp.forward(p, feats, bf16=False)
    loss, grads = mlp.backward(p, cache, target, bf16=False)
    tp = {k: torch.tensor(v, dtype=torch.float64, requires_grad=k[0] in "wb") for k, v in p.items()
          if isinstance(v, np.ndarray)}
    h = (torch.from_numpy(feats) - tp["mean"]) / tp["std"]
    for i, (w, b, _, _) in enumerate(mlp.LAYERS):
        h = h @ tp[w].T + tp[b]
        if i < 3:
            h = torch.nn.functional.silu(h)
    tl = torch.mean((h[:, 0] - torch.from_numpy(target)) ** 2)
    tl.backward()
    np.testing.assert_allclose(y, h[:, 0].detach().numpy(), rtol=1e-13, atol=1e-13)
    assert loss == pytest.approx(float(tl.detach()), rel=1e-13)
    for key in ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4"):
        np.testing.assert_allclose(grads[key], tp[key].grad.numpy(), rtol=1e-11, atol=1e-13)


def test_bf16_forward_close_to_fp64():
    """bfloat16 operands perturb the output at the bfloat16 unit round-off scale, not more."""
    p, feats, _ = _random_case(seed=4, npts=200)
    y64, _ = mlp.forward(p, feats, bf16=False)
    y16, _ = mlp.forward(p, feats, bf16=True)
    assert np.abs(y16 - y64).max() / np.abs(y64).max() < 3e-2


def test_density_start_counts_electrons():
    rng = np.random.default_rng(0)
    atomic = rng.uniform(0.01, 0.1, size=(6, 6, 6))
    corr = rng.normal(scale=0.05, size=atomic.shape)
    rho = mlp.density_start(atomic, corr, 8.0, 0.37)
    assert rho.min() > 0
    assert rho.sum() * 0.37 == pytest.approx(8.0, rel=1e-14)
