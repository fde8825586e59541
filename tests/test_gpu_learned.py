"""HIP kernels of the learned start against the FP64 oracle (oracle/mlp.py).

Features are FP64 on the device and stored as float32, so they match to float32 rounding.
The network kernels round weights, layer inputs and back-propagated gradients to bfloat16
exactly where the oracle's ``bf16=True`` form does and sum in FP32, so they agree with it up
to FP32 summation order (and the rare bfloat16 rounding flip that order can cause): the
tolerances below are written for that. Adam is checked against torch.optim.Adam."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _model_and_data(seed=1, npts=1000 + 7):
    from amaterasu_amd.learned import DensityModel
    from oracle import mlp

    rng = np.random.default_rng(seed)
    p = mlp.init_params(seed)
    for key in ("b1", "b2", "b3", "b4"):
        p[key] = rng.normal(scale=0.1, size=p[key].shape)
    p["mean"] = rng.normal(size=mlp.N_RBF) * 0.1
    p["std"] = 0.5 + rng.uniform(size=mlp.N_RBF)
    theta = np.concatenate([p[k].ravel() for k in ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4")])
    model = DensityModel(torch.from_numpy(theta.astype(np.float32)).cuda(),
                         torch.from_numpy(p["mean"].astype(np.float32)).cuda(),
                         torch.from_numpy(p["std"].astype(np.float32)).cuda())
    # the oracle sees exactly the float32 values the device sees
    for k in ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4", "mean", "std"):
        p[k] = p[k].astype(np.float32).astype(np.float64)
    feats = rng.uniform(0.0, 2.0, size=(npts, mlp.N_RBF)).astype(np.float32)
    target = rng.normal(size=npts).astype(np.float32)
    return model, p, feats, target


@pytest.mark.parametrize("case", ["skewed", "si8"])
def test_features_match_oracle(case):
    from amaterasu_amd.learned import grid_features
    from amaterasu_amd.workloads import si_conventional
    from oracle import mlp

    if case == "skewed":
        rng = np.random.default_rng(5)
        cell = np.array([[6.0, 0.4, 0.0], [1.1, 5.5, 0.3], [0.2, -0.6, 6.3]])
        pos = rng.uniform(size=(3, 3)) @ cell
        shape = (6, 5, 7)
    else:
        cell, pos = si_conventional(1, sigma_ang=0.1, seed=3)
        shape = (12, 12, 12)
    got = grid_features(cell, pos, shape).cpu().numpy().astype(np.float64)
    idx = np.stack(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), -1).reshape(-1, 3)
    pts = (idx / np.array(shape)) @ cell
    want = mlp.rbf_features(cell, pos, pts)
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=2e-7, atol=1e-30)


def test_forward_matches_oracle():
    from oracle import mlp

    model, p, feats, _ = _model_and_data()
    y = model.forward(torch.from_numpy(feats).cuda()).cpu().numpy().astype(np.float64)
    want, _ = mlp.forward(p, feats.astype(np.float64), bf16=True)
    err = np.abs(y - want)
    assert np.linalg.norm(y - want) / np.linalg.norm(want) < 1e-4
    assert err.max() < 5e-3 * np.abs(want).max()
    # and the bf16 form stays close to the full FP64 network
    y64, _ = mlp.forward(p, feats.astype(np.float64), bf16=False)
    assert np.linalg.norm(y - y64) / np.linalg.norm(y64) < 3e-2


@pytest.mark.parametrize("npts", [16, 1000 + 7, 40000])
def test_gradient_matches_oracle(npts):
    from amaterasu_amd.learned import loss_and_grad
    from oracle import mlp

    model, p, feats, target = _model_and_data(seed=2, npts=npts)
    loss, grad = loss_and_grad(model, torch.from_numpy(feats).cuda(), torch.from_numpy(target).cuda())
    y, cache = mlp.forward(p, feats.astype(np.float64), bf16=True)
    wl, wg = mlp.backward(p, cache, target.astype(np.float64), bf16=True)
    assert float(loss.item()) == pytest.approx(wl, rel=1e-5)
    g = grad.cpu().numpy().astype(np.float64)
    off = 0
    for key in ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4"):
        n = wg[key].size
        got, want = g[off:off + n], wg[key].ravel()
        off += n
        assert np.linalg.norm(got - want) <= 2e-3 * np.linalg.norm(want) + 1e-12, key
    assert off == g.size


def test_gradient_is_deterministic_and_batch_split_adds_up():
    """Fixed reduction order: bit-identical repeats. Splitting a batch over two calls with
    the full count gives the same gradient up to FP32 rounding."""
    from amaterasu_amd.learned import loss_and_grad

    model, _, feats, target = _model_and_data(seed=3, npts=5000)
    f, t = torch.from_numpy(feats).cuda(), torch.from_numpy(target).cuda()
    _, g1 = loss_and_grad(model, f, t)
    _, g2 = loss_and_grad(model, f, t)
    assert torch.equal(g1, g2)
    _, ga = loss_and_grad(model, f[:2000].contiguous(), t[:2000].contiguous(), count=5000)
    _, gb = loss_and_grad(model, f[2000:].contiguous(), t[2000:].contiguous(), count=5000)
    assert torch.linalg.vector_norm(ga + gb - g1) <= 1e-5 * torch.linalg.vector_norm(g1)


def test_adam_matches_torch():
    from amaterasu_amd.learned import Adam, DensityModel

    model = DensityModel.initial(seed=4)
    ref = torch.nn.Parameter(model.theta.detach().cpu().clone())
    opt_ref = torch.optim.Adam([ref], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    opt = Adam(model, lr=3e-3)
    rng = np.random.default_rng(0)
    for _ in range(5):
        g = torch.from_numpy(rng.normal(size=ref.numel()).astype(np.float32))
        ref.grad = g.clone()
        opt_ref.step()
        opt.step(g.cuda())
    assert torch.allclose(model.theta.cpu(), ref.detach(), rtol=1e-5, atol=1e-6)


def test_training_fits_a_smooth_target():
    """A few epochs on a smooth function of the features lower the loss by an order of magnitude."""
    from amaterasu_amd.learned import DensityModel, train

    rng = np.random.default_rng(7)
    feats = rng.uniform(0.0, 2.0, size=(60000, 16)).astype(np.float32)
    target = (np.sin(feats[:, 0]) + 0.5 * feats[:, 3] * feats[:, 7] - 0.3 * feats[:, 12]).astype(np.float32)
    model = DensityModel.initial(seed=0, mean=feats.mean(0), std=feats.std(0))
    hist = train(model, torch.from_numpy(feats).cuda(), torch.from_numpy(target).cuda(), epochs=12, batch=4096,
                 lr=3e-3)
    assert hist[-1] < 0.1 * hist[0]


def test_learned_density_keeps_the_electron_count(si2_system):
    from amaterasu_amd.learned import DensityModel, learned_density
    from amaterasu_amd.system import attach_ionic_fields, build_system, load_dataset

    ds = load_dataset()
    o = si2_system
    s = build_system(o.cell, o.positions, o.symbols, {"Si": ds}, 32.0, o.kpoints.points, 8.0,
                     weights=o.kpoints.weights, grid_shape=o.grid.shape)
    attach_ionic_fields(s, {"Si": ds}, 32.0)
    model = DensityModel.initial(seed=1, out_scale=1e-3)
    rho = learned_density(s, model)
    assert rho.shape == (1, *o.grid.shape)
    assert float(rho.min()) > 0
    assert float(rho.sum()) * s.grid.cell_volume_per_point == pytest.approx(8.0, rel=1e-12)
