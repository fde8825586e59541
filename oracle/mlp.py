"""FP64 restatement of the learned starting-density model (test infrastructure only).

The model has no counterpart in the reference: it produces the ``density=`` argument of
``run_scf`` (``scf/driver.py:267-299``) as a correction over the atomic superposition of
``PlaneWaveSystem.initial_density`` (``hamiltonian/system.py:147-190``,
``hamiltonian/local.py:77-106``), which stays the fallback. This module is the checker of
the HIP feature, forward and backward kernels:

* ``rbf_features``: for a grid point r, f_k(r) = sum over atoms a and lattice images L with
  d = |r - tau_a - L| < r_c of exp(-gamma (d - mu_k)^2) * 0.5 (cos(pi d / r_c) + 1),
  mu_k = r_c k / (K - 1), k = 0..K-1, gamma = 1 / (2 (r_c / (K - 1))^2).
* ``forward``: x = (f - mean) / std; three hidden layers of width 64 with SiLU; a linear
  output y; the density correction is ``out_scale * y``.
* ``backward``: gradients of L = mean_p (y_p - t_p)^2 with respect to every parameter.

With ``bf16=True`` the operands the matrix cores see are rounded to bfloat16
(round-to-nearest-even) at the same points as the HIP kernels -- the weights, the layer
inputs and the back-propagated pre-activation gradients -- while every sum is carried in
FP64 here (FP32 on the device). With ``bf16=False`` everything is FP64; the tests pin that
form by central finite differences and by torch's autograd in float64.
"""

from __future__ import annotations

import numpy as np

from .units import ANG_IN_BOHR

N_RBF = 16
WIDTH = 64
R_CUT = 4.0 * ANG_IN_BOHR  # 4 Angstrom in Bohr
LAYERS = (("w1", "b1", N_RBF, WIDTH), ("w2", "b2", WIDTH, WIDTH), ("w3", "b3", WIDTH, WIDTH), ("w4", "b4", WIDTH, 1))


def bf16_round(x) -> np.ndarray:
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def rbf_centres(n_rbf: int = N_RBF, r_cut: float = R_CUT):
    mu = r_cut * np.arange(n_rbf) / (n_rbf - 1)
    gamma = 1.0 / (2.0 * (r_cut / (n_rbf - 1)) ** 2)
    return mu, gamma


def image_range(cell, r_cut: float = R_CUT) -> np.ndarray:
    """Images needed along each lattice vector: ceil(r_c / plane spacing) + 1 (points and
    atoms are both wrapped into the cell, so one extra image covers the wrap)."""
    cell = np.asarray(cell, dtype=float)
    vol = abs(np.linalg.det(cell))
    spacing = np.array([vol / np.linalg.norm(np.cross(cell[(a + 1) % 3], cell[(a + 2) % 3])) for a in range(3)])
    return np.ceil(r_cut / spacing).astype(int) + 1


def rbf_features(cell, positions, points, n_rbf: int = N_RBF, r_cut: float = R_CUT) -> np.ndarray:
    """(P, n_rbf) features of the Cartesian points (Bohr)."""
    cell = np.asarray(cell, dtype=float)
    pos = np.atleast_2d(np.asarray(positions, dtype=float))
    pts = np.atleast_2d(np.asarray(points, dtype=float))
    inv = np.linalg.inv(cell)
    pos = (np.mod(pos @ inv, 1.0)) @ cell
    pts = (np.mod(pts @ inv, 1.0)) @ cell
    mu, gamma = rbf_centres(n_rbf, r_cut)
    n = image_range(cell, r_cut)
    out = np.zeros((len(pts), n_rbf))
    for i in range(-n[0], n[0] + 1):
        for j in range(-n[1], n[1] + 1):
            for k in range(-n[2], n[2] + 1):
                shift = np.array([i, j, k], dtype=float) @ cell
                d = np.linalg.norm(pts[:, None, :] - pos[None, :, :] - shift, axis=-1)  # (P, A)
                inside = d < r_cut
                fc = np.where(inside, 0.5 * (np.cos(np.pi * d / r_cut) + 1.0), 0.0)
                out += np.einsum("pa,pak->pk", fc, np.exp(-gamma * (d[..., None] - mu) ** 2))
    return out


def init_params(seed: int = 0) -> dict:
    """He-style normal initialisation, zero biases."""
    rng = np.random.default_rng(seed)
    p = {}
    for w, b, fin, fout in LAYERS:
        p[w] = rng.normal(scale=np.sqrt(2.0 / fin), size=(fout, fin))
        p[b] = np.zeros(fout)
    p["mean"] = np.zeros(N_RBF)
    p["std"] = np.ones(N_RBF)
    p["out_scale"] = 1.0
    return p


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params: dict, feats, *, bf16: bool = True):
    """y (P,) and the cache the backward pass needs."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    x = r((np.asarray(feats, dtype=np.float64) - params["mean"]) / params["std"])
    cache = {"in": [], "z": []}
    h = x
    for idx, (w, b, _, _) in enumerate(LAYERS):
        z = h @ r(params[w]).T + params[b]
        cache["in"].append(h)
        cache["z"].append(z)
        if idx < len(LAYERS) - 1:
            h = r(_silu(z))
    y = cache["z"][-1][:, 0]
    return y, cache


def backward(params: dict, cache: dict, target, *, bf16: bool = True):
    """(loss, grads) for L = mean (y - t)^2."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    y = cache["z"][-1][:, 0]
    t = np.asarray(target, dtype=np.float64)
    npt = len(y)
    loss = float(np.mean((y - t) ** 2))
    grads = {}
    dz = (2.0 / npt) * (y - t)[:, None]  # (P, 1)
    for idx in range(len(LAYERS) - 1, -1, -1):
        w, b, _, _ = LAYERS[idx]
        dzr = r(dz)
        grads[w] = dzr.T @ cache["in"][idx]
        grads[b] = dz.sum(axis=0)
        if idx > 0:
            dh = dzr @ r(params[w])
            dz = dh * _dsilu(cache["z"][idx - 1])
    return loss, grads


def density_start(atomic, correction, nelectrons: float, dv: float):
    """atomic + correction, clipped at zero and rescaled to the electron count
    (the floor and rescale of ``local.py:102-105``)."""
    rho = np.maximum(np.asarray(atomic, dtype=float) + np.asarray(correction, dtype=float), 1e-12)
    return rho * (nelectrons / (rho.sum() * dv))
