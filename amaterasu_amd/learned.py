"""Learned starting density: an MLP correction over the atomic superposition.

No counterpart in the reference. It produces the ``density=`` argument of ``run_scf``
(``scf/driver.py:267-299``) as

    rho_0(r) = clip(rho_atomic(r) + out_scale * y(f(r)), 1e-12) rescaled to N_el,

where ``rho_atomic`` is ``PlaneWaveSystem.initial_density`` (``hamiltonian/system.py:147-190``,
``local.py:77-106``), f(r) are 16 Gaussian radial features of the atoms within 4 A of the
grid point and y is a 16 -> 64 -> 64 -> 64 -> 1 SiLU network. The atomic superposition
stays the fallback (no model, or a model whose start has a larger density residual) and
the reference point: the SCF converges to the same answer from either start; only the iteration
count and the wall time change.

Features, inference and training (loss, gradient, Adam) are HIP kernels of
libamaterasu_amd (bf16 matrix-core operands, FP32 sums); torch holds the device buffers.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from ._lib import check, lib

BOHR_IN_ANG = 0.5291772105638411
N_RBF = 16
R_CUT = 4.0 / BOHR_IN_ANG
GAMMA = 1.0 / (2.0 * (R_CUT / (N_RBF - 1)) ** 2)
MODEL_PATH = Path(__file__).resolve().parent / "data" / "density_mlp.npz"


def _ptr(t: torch.Tensor):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def param_count() -> int:
    return int(lib().amt_mlp_param_count())


def image_range(cell, r_cut: float = R_CUT) -> np.ndarray:
    """Lattice images to search per axis: ceil(r_c / plane spacing) + 1."""
    cell = np.asarray(cell, dtype=float)
    vol = abs(np.linalg.det(cell))
    spacing = np.array([vol / np.linalg.norm(np.cross(cell[(a + 1) % 3], cell[(a + 2) % 3])) for a in range(3)])
    return (np.ceil(r_cut / spacing).astype(int) + 1).astype(np.int32)


def grid_features(cell, positions, shape) -> torch.Tensor:
    """(N, 16) float32 features of every point of the grid (C order), on the device."""
    cell = np.asarray(cell, dtype=float)
    frac = np.mod(np.atleast_2d(np.asarray(positions, dtype=float)) @ np.linalg.inv(cell), 1.0)
    n1, n2, n3 = (int(x) for x in shape)
    imgs = image_range(cell)
    if np.any(imgs > 8):
        raise ValueError(f"cell too thin for a {R_CUT:.3f} Bohr cutoff: {imgs} images per axis")
    d_cell = torch.from_numpy(np.ascontiguousarray(cell)).to("cuda")
    d_frac = torch.from_numpy(np.ascontiguousarray(frac)).to("cuda")
    out = torch.empty((n1 * n2 * n3, N_RBF), dtype=torch.float32, device="cuda")
    im = (ctypes.c_int * 3)(*imgs.tolist())
    check(lib().amt_mlp_features(_ptr(d_cell), _ptr(d_frac), len(frac), n1, n2, n3, im, R_CUT, GAMMA, _ptr(out),
                                 _stream()), "mlp_features")
    return out


@dataclass
class DensityModel:
    """Parameters (FP32 master vector), feature normalisation and output scale."""

    theta: torch.Tensor
    mean: torch.Tensor
    std: torch.Tensor
    out_scale: float = 1.0

    @classmethod
    def initial(cls, seed: int = 0, mean=None, std=None, out_scale: float = 1.0) -> "DensityModel":
        """He-normal weights, zero biases (host numpy draws, so the start is reproducible)."""
        rng = np.random.default_rng(seed)
        parts = []
        for fin, fout in ((N_RBF, 64), (64, 64), (64, 64), (64, 1)):
            parts.append(rng.normal(scale=np.sqrt(2.0 / fin), size=fout * fin))
            parts.append(np.zeros(fout))
        theta = np.concatenate(parts).astype(np.float32)
        if theta.size != param_count():
            raise RuntimeError("parameter layout disagrees with the library")
        m = np.zeros(N_RBF) if mean is None else np.asarray(mean)
        s = np.ones(N_RBF) if std is None else np.asarray(std)
        return cls(torch.from_numpy(theta).to("cuda"), torch.as_tensor(m, dtype=torch.float32, device="cuda"),
                   torch.as_tensor(s, dtype=torch.float32, device="cuda"), float(out_scale))

    def save(self, path) -> None:
        np.savez(path, theta=self.theta.cpu().numpy(), mean=self.mean.cpu().numpy(), std=self.std.cpu().numpy(),
                 out_scale=np.array(self.out_scale))

    @classmethod
    def load(cls, path=MODEL_PATH) -> "DensityModel":
        with np.load(path, allow_pickle=False) as z:
            theta = z["theta"].astype(np.float32)
            if theta.size != param_count():
                raise ValueError(f"{path}: {theta.size} parameters, the library expects {param_count()}")
            return cls(torch.from_numpy(theta).to("cuda"), torch.from_numpy(z["mean"].astype(np.float32)).to("cuda"),
                       torch.from_numpy(z["std"].astype(np.float32)).to("cuda"), float(z["out_scale"]))

    def forward(self, feats: torch.Tensor) -> torch.Tensor:
        """Raw network output y for (N, 16) float32 features."""
        feats = feats.contiguous()
        if feats.dtype != torch.float32 or feats.dim() != 2 or feats.shape[1] != N_RBF:
            raise ValueError(f"features must be (N, {N_RBF}) float32, got {tuple(feats.shape)} {feats.dtype}")
        y = torch.empty(feats.shape[0], dtype=torch.float32, device=feats.device)
        check(lib().amt_mlp_forward(_ptr(feats), feats.shape[0], _ptr(self.theta), _ptr(self.mean), _ptr(self.std),
                                    _ptr(y), _stream()), "mlp_forward")
        return y

    def correction(self, cell, positions, shape) -> torch.Tensor:
        """out_scale * y on the grid, float64, shape ``shape``."""
        y = self.forward(grid_features(cell, positions, shape))
        return (self.out_scale * y.to(torch.float64)).reshape(tuple(shape))


def loss_and_grad(model: DensityModel, feats: torch.Tensor, target: torch.Tensor, count=None, *, ws=None):
    """(loss, grad) of mean (y - target)^2 over the rows (divided by ``count`` if given)."""
    if feats.dtype != torch.float32 or feats.dim() != 2 or feats.shape[1] != N_RBF or not feats.is_contiguous():
        raise ValueError(f"features must be contiguous (N, {N_RBF}) float32, got {tuple(feats.shape)} {feats.dtype}")
    if target.dtype != torch.float32 or target.shape != (feats.shape[0],) or not target.is_contiguous():
        raise ValueError(f"target must be contiguous ({feats.shape[0]},) float32, got {tuple(target.shape)}")
    n = feats.shape[0]
    need = int(lib().amt_mlp_workspace(n))
    if ws is None or ws.numel() * ws.element_size() < need:
        ws = torch.empty(need // 4 + 1, dtype=torch.float32, device="cuda")
    grad = torch.empty(param_count(), dtype=torch.float32, device="cuda")
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    check(lib().amt_mlp_loss_grad(_ptr(feats), _ptr(target), n, float(count or n), _ptr(model.theta),
                                  _ptr(model.mean), _ptr(model.std), _ptr(grad), _ptr(loss), _ptr(ws),
                                  ws.numel() * ws.element_size(), _stream()), "mlp_loss_grad")
    return loss, grad


class Adam:
    """Adam on the model's FP32 master parameters, one HIP kernel per step."""

    def __init__(self, model: DensityModel, lr=1e-3, betas=(0.9, 0.999), eps=1e-8):
        self.model, self.lr, self.betas, self.eps = model, lr, betas, eps
        self.m = torch.zeros_like(model.theta)
        self.v = torch.zeros_like(model.theta)
        self.t = 0

    def step(self, grad: torch.Tensor, lr=None) -> None:
        self.t += 1
        check(lib().amt_adam(_ptr(self.model.theta), _ptr(grad), _ptr(self.m), _ptr(self.v), self.model.theta.numel(),
                             float(self.lr if lr is None else lr), self.betas[0], self.betas[1], self.eps, self.t,
                             _stream()), "adam")


def train(model: DensityModel, feats: torch.Tensor, target: torch.Tensor, *, epochs: int = 20,
          batch: int = 65536, lr: float = 2e-3, seed: int = 0, log=None) -> list:
    """Minibatch Adam over (feats, target) rows with a cosine learning-rate decay; returns
    the mean training loss per epoch. Rows are shuffled on the device each epoch."""
    opt = Adam(model, lr=lr)
    n = feats.shape[0]
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    ws = torch.empty(int(lib().amt_mlp_workspace(min(batch, n))) // 4 + 1, dtype=torch.float32, device="cuda")
    steps = epochs * ((n + batch - 1) // batch)
    hist, step = [], 0
    for ep in range(epochs):
        perm = torch.randperm(n, device="cuda", generator=gen)
        tot = torch.zeros(1, dtype=torch.float64, device="cuda")
        for b0 in range(0, n, batch):
            idx = perm[b0:b0 + batch]
            f, t = feats.index_select(0, idx).contiguous(), target.index_select(0, idx).contiguous()
            loss, grad = loss_and_grad(model, f, t, ws=ws)
            rate = 0.5 * lr * (1.0 + np.cos(np.pi * step / max(steps, 1)))
            opt.step(grad, lr=rate)
            tot += loss.to(torch.float64) * f.shape[0]
            step += 1
        hist.append(float(tot.item()) / n)
        if log:
            log(f"epoch {ep + 1:3d} loss {hist[-1]:.6e}")
    return hist


def learned_density(system, model: DensityModel | None = None, atomic=None) -> torch.Tensor:
    """The learned start (nspin, n1, n2, n3) float64 on the device: atomic + correction,
    floored at 1e-12 and rescaled to the electron count (the floor and rescale of
    ``local.py:102-105``). Spin channels share the correction in proportion."""
    if atomic is None:
        atomic = system.initial_density()
    atomic = torch.as_tensor(atomic, dtype=torch.float64, device="cuda")
    if model is None:
        model = DensityModel.load()
    corr = model.correction(system.cell, system.positions, system.grid.shape)
    total = atomic.sum(0)
    rho = torch.clamp(total + corr, min=1e-12)
    rho *= system.nelectrons / (float(rho.sum().item()) * system.grid.cell_volume_per_point)
    share = atomic / torch.clamp(total, min=1e-300)[None]
    return (share * rho[None]).contiguous()


__all__ = ["Adam", "DensityModel", "grid_features", "learned_density", "loss_and_grad", "param_count", "train"]
