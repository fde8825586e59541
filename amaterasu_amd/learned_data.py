"""Training data, training and scoring of the learned starting density.

``python -m amaterasu_amd.learned_data generate --first 0 --count 80 --out DIR``
    converges seeded strained / displaced Si8 cells with the plain FP64 GPU SCF and writes
    the density correction rho_scf - rho_atomic (float32) per cell;
``python -m amaterasu_amd.learned_data train --data DIR [DIR ...] --out model.npz``
    computes the grid features, trains the network with the HIP kernels on the first 256
    cells and reports the held-out loss on the rest;
``python -m amaterasu_amd.learned_data score --model model.npz --first 256 --count 64``
    reruns the held-out cells from the atomic and from the learned start and reports SCF
    iterations and wall time to the same threshold (1e-6 Ha/atom).

Cells: conventional Si8 (a = 5.431 A), a symmetric strain with every component drawn from
U[-0.03, 0.03], Gaussian displacements sigma = 0.1 A (seed 1000 + cell index); ecut 15 Ha,
the 2x2x2 shifted Monkhorst-Pack mesh reduced by time reversal (4 k-points), LDA (PW92)
exchange-correlation with the repository's norm-conserving Si dataset.
"""

from __future__ import annotations

import argparse
import dataclasses
import json
import time
from pathlib import Path

import numpy as np
import torch

from .learned import DensityModel, grid_features, learned_density, train
from .scf import HARTREE_IN_EV, SCFSettings, run_scf
from .system import attach_ionic_fields, build_system, load_dataset
from .workloads import KMESH_SHIFTED_222, si_conventional

ECUT = 15.0
TOL = SCFSettings(tol_energy=1e-6 * HARTREE_IN_EV, tol_density=1e-6, max_steps=60)


def training_cell(index: int):
    """Cell (Bohr rows) and Cartesian positions of training cell ``index``."""
    rng = np.random.default_rng(1000 + index)
    cell, pos = si_conventional(1)
    e = rng.uniform(-0.03, 0.03, size=(3, 3))
    strain = np.eye(3) + 0.5 * (e + e.T)
    frac = pos @ np.linalg.inv(cell)
    cell = cell @ strain
    pos = frac @ cell + rng.normal(scale=0.1 / 0.5291772105638411, size=pos.shape)
    return cell, pos


def training_system(index: int, ds=None):
    ds = ds or load_dataset()
    cell, pos = training_cell(index)
    s = build_system(cell, pos, ["Si"] * 8, {"Si": ds}, 4 * ECUT, KMESH_SHIFTED_222, ECUT)
    return attach_ionic_fields(s, {"Si": ds}, 4 * ECUT, xc="LDA-PW")


def generate(first: int, count: int, out: Path, log=print, *, batched_k: bool = True) -> None:
    out.mkdir(parents=True, exist_ok=True)
    ds = load_dataset()
    settings = dataclasses.replace(TOL, batched_k=batched_k)
    for c in range(first, first + count):
        t0 = time.perf_counter()
        s = training_system(c, ds)
        atomic = s.initial_density()
        r = run_scf(s, settings)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        corr = (torch.from_numpy(r.density).to("cuda") - atomic).sum(0)
        np.savez(out / f"cell{c:04d}.npz", cell=s.cell, positions=s.positions, shape=np.array(s.grid.shape),
                 correction=corr.to(torch.float32).cpu().numpy(), iterations=np.array(r.iterations),
                 converged=np.array(r.converged), energy=np.array(r.energies.total), seconds=np.array(dt))
        log(f"cell {c:4d} grid {s.grid.shape} iterations {r.iterations:3d} converged {r.converged} "
            f"E {r.energies.total:.8f} {dt:.2f}s")


def _load(dirs, first: int, count: int):
    rows = []
    for c in range(first, first + count):
        for d in dirs:
            f = Path(d) / f"cell{c:04d}.npz"
            if f.exists():
                with np.load(f, allow_pickle=False) as z:
                    row = {k: z[k] for k in z.files}
                row["correction"] = row["correction"].astype(np.float32)
                rows.append(row)
                break
        else:
            raise FileNotFoundError(f"cell {c} not found in {dirs}")
    return rows


def _features_targets(rows, scale: float):
    feats, targets = [], []
    for r in rows:
        feats.append(grid_features(r["cell"], r["positions"], tuple(int(x) for x in r["shape"])))
        targets.append(torch.from_numpy(r["correction"].reshape(-1) / scale).to("cuda", torch.float32))
    return torch.cat(feats), torch.cat(targets)


def train_model(dirs, out: Path, *, ntrain: int = 256, nheld: int = 64, epochs: int = 30, batch: int = 65536,
                lr: float = 2e-3, log=print) -> dict:
    rows = _load(dirs, 0, ntrain)
    held = _load(dirs, ntrain, nheld)
    scale = float(np.sqrt(np.mean(np.concatenate([r["correction"].ravel() ** 2 for r in rows]))))
    f, t = _features_targets(rows, scale)
    fh, th = _features_targets(held, scale)
    mean, std = f.mean(0), f.std(0).clamp(min=1e-6)
    model = DensityModel.initial(seed=0, mean=mean.cpu().numpy(), std=std.cpu().numpy(), out_scale=scale)
    t0 = time.perf_counter()
    hist = train(model, f, t, epochs=epochs, batch=batch, lr=lr, log=log)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    held_loss = float(torch.mean((model.forward(fh) - th) ** 2).item())
    base_loss = float(torch.mean(th**2).item())
    model.save(out)
    res = dict(train_points=int(f.shape[0]), epochs=epochs, seconds=dt, train_loss=hist[-1], held_loss=held_loss,
               held_loss_zero_model=base_loss, scale=scale)
    log(json.dumps(res))
    return res


def score(model_path: Path, first: int, count: int, log=print) -> dict:
    model = DensityModel.load(model_path)
    ds = load_dataset()
    out = []
    for c in range(first, first + count):
        s = training_system(c, ds)
        res = {}
        for name in ("atomic", "learned"):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rho = s.initial_density() if name == "atomic" else learned_density(s, model)
            r = run_scf(s, TOL, density=rho)
            torch.cuda.synchronize()
            res[name] = (r.iterations, time.perf_counter() - t0, r.energies.total, r.converged)
        out.append(res)
        log(f"cell {c:4d} atomic {res['atomic'][0]:3d} it {res['atomic'][1]:.2f}s  learned {res['learned'][0]:3d} it "
            f"{res['learned'][1]:.2f}s  dE {abs(res['atomic'][2] - res['learned'][2]):.2e} Ha")
    ia = np.array([r["atomic"][0] for r in out])
    il = np.array([r["learned"][0] for r in out])
    ta = np.array([r["atomic"][1] for r in out])
    tl = np.array([r["learned"][1] for r in out])
    de = np.array([abs(r["atomic"][2] - r["learned"][2]) for r in out]) / 8
    summary = dict(cells=count, iterations_atomic=float(ia.mean()), iterations_learned=float(il.mean()),
                   seconds_atomic=float(ta.sum()), seconds_learned=float(tl.sum()),
                   max_energy_difference_per_atom=float(de.max()),
                   all_converged=bool(all(r["atomic"][3] and r["learned"][3] for r in out)))
    log(json.dumps(summary))
    return summary


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m amaterasu_amd.learned_data")
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate")
    g.add_argument("--first", type=int, default=0)
    g.add_argument("--count", type=int, default=320)
    g.add_argument("--out", type=Path, required=True)
    g.add_argument("--per-k", action="store_true", help="k-by-k Davidson instead of the k-batched one")
    t = sub.add_parser("train")
    t.add_argument("--data", type=Path, nargs="+", required=True)
    t.add_argument("--out", type=Path, required=True)
    t.add_argument("--epochs", type=int, default=30)
    t.add_argument("--ntrain", type=int, default=256)
    t.add_argument("--nheld", type=int, default=64)
    s = sub.add_parser("score")
    s.add_argument("--model", type=Path, required=True)
    s.add_argument("--first", type=int, default=256)
    s.add_argument("--count", type=int, default=64)
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)  # noqa: E731
    if a.cmd == "generate":
        generate(a.first, a.count, a.out, log, batched_k=not a.per_k)
    elif a.cmd == "train":
        train_model(a.data, a.out, ntrain=a.ntrain, nheld=a.nheld, epochs=a.epochs, log=log)
    else:
        score(a.model, a.first, a.count, log)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
