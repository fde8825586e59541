"""Block Davidson over all k-points of a spin channel at once (``solve_batched``).

Each k-point runs exactly the iteration of ``scf/eigen/davidson.py:48-111`` -- its own
active set, its own dropped directions, its own subspace size and collapse, its own
convergence -- but the k-points advance together so that every step is one launch per
kernel over all of them: one H.psi over the padded block ``[nk][rows][npw_pad]``, one
batched GEMM per Gram / projected matrix / rotation, one batched Cholesky and one
batched Hermitian eigensolve.

Ragged shapes are padded, never masked by branching on the device:
* plane waves: every row is ``npw_pad`` long with zeros past the k-point's own count
  (H.psi, the GEMMs and the preconditioner keep the padding zero);
* subspace rows: k-point k uses the first m_k rows of a ``[nk][M][npw_pad]`` basis; the
  unused rows are zero. Each small eigensolve runs on the k-point's own m_k x m_k block
  and is padded afterwards (unit eigenvectors on the unused coordinates, eigenvalues that
  sort them out of the way); before a Cholesky the padded diagonal is set to one;
* new directions: k-point k contributes n_k <= nb rows of a zero-padded block.

A k-point that has converged (or that stopped adding directions) is frozen: its
eigenpairs and iteration count are those of the step at which the per-k loop would
have returned.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import ops
from .eigen import _lowest_eigenpairs, orthonormalize


@dataclass
class BatchedResult:
    """Per k-point ``DavidsonResult`` fields, stacked: eigenvalues [nk][nb], coefficients
    [nk][nb][npw_pad] (zero past each k's npw), residual norms, iterations, converged."""

    eigenvalues: torch.Tensor
    coefficients: torch.Tensor
    residual_norms: torch.Tensor
    iterations: np.ndarray
    converged: np.ndarray


def _herm(m: torch.Tensor) -> torch.Tensor:
    m = m.to(torch.complex128)
    return 0.5 * (m + m.conj().transpose(-1, -2))


def _bgram(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """a @ b^H per batch entry on the matrix cores."""
    return ops.zgemm(a, b, 1)


def _brot(rot: torch.Tensor, block: torch.Tensor) -> torch.Tensor:
    return ops.zgemm(rot.to(block.dtype).contiguous(), block, 0)


def _pad_diag(m: torch.Tensor, used: torch.Tensor, value: float, step: float = 0.0) -> torch.Tensor:
    """Set the diagonal of rows/cols j >= used[k] (zero rows/cols) to ``value + step * j``:
    distinct values keep the padded block free of repeated eigenvalues, which the batched
    Hermitian eigensolver does not always resolve."""
    n = m.shape[-1]
    idx = torch.arange(n, device=m.device)
    mask = idx[None, :] >= used[:, None]  # [nk][n]
    diag = torch.diagonal(m, dim1=-2, dim2=-1)
    diag += mask.to(m.dtype) * (value + step * idx.to(m.dtype))[None, :]
    return m


def _cholesky_rows(block: torch.Tensor, used: torch.Tensor):
    """L^{-1} block per k (rows >= used[k] are zero and stay zero); None where it fails."""
    s = _pad_diag(_herm(_bgram(block, block)), used, 1.0)
    low, info = torch.linalg.cholesky_ex(s)
    eye = torch.eye(s.shape[-1], dtype=s.dtype, device=s.device).repeat(s.shape[0], 1, 1)
    linv = torch.linalg.solve_triangular(low, eye, upper=False)
    return _brot(linv, block), info


def _eigh(m: torch.Tensor, used, fill: float):
    """Eigen-decomposition of the leading used[k] x used[k] block of each matrix, one solver
    call per matrix at its own size, returned padded to the batch size: eigenvalue ``fill``
    and unit eigenvectors on the padded coordinates, so padded pairs sort after (fill
    large) or before the physical ones and never mix with them. (The batched solver's
    Jacobi sweeps do not always converge on ill-conditioned Gram matrices, and the
    divide-and-conquer solver has returned non-finite eigenvectors for blocks padded with
    large diagonal values.)"""
    nk, n = m.shape[0], m.shape[-1]
    used = [int(u) for u in (used.tolist() if isinstance(used, torch.Tensor) else used)]
    w = torch.full((nk, n), float(fill), dtype=torch.float64, device=m.device)
    u = torch.zeros_like(m)
    for k in range(nk):
        c = used[k]
        if c < n:
            idx = torch.arange(c, n, device=m.device)
            u[k, idx, idx] = 1.0
        if c > 0:
            wk, uk = torch.linalg.eigh(m[k, :c, :c])
            w[k, :c] = wk
            u[k, :c, :c] = uk
    order = torch.argsort(w, dim=-1, stable=True)
    return torch.gather(w, 1, order), torch.gather(u, 2, order[:, None, :].expand(-1, n, -1))


def _gram_rows(block: torch.Tensor, used: torch.Tensor, threshold: float):
    """Per k: keep directions with singular value > threshold * s_max, orthonormal
    (the routes of ``eigen.orthogonalize_against``). Returns (rows, kept)."""
    g = _herm(_bgram(block, block))
    gp = _pad_diag(g.clone(), used, 1.0)
    low, info = torch.linalg.cholesky_ex(gp)
    eye = torch.eye(g.shape[-1], dtype=g.dtype, device=g.device).repeat(g.shape[0], 1, 1)
    linv = torch.linalg.solve_triangular(low, eye, upper=False)
    # the padded identity block adds at most sqrt(pad) to both Frobenius norms; use the
    # used block alone: zero the padded rows/cols before measuring
    n = g.shape[-1]
    live = (torch.arange(n, device=g.device)[None, :] < used[:, None]).to(g.dtype)
    mask = live[:, :, None] * live[:, None, :]
    lo = 1.0 / torch.linalg.matrix_norm(linv * mask)
    hi = torch.linalg.matrix_norm(low * mask)
    ok = (used == 0) | ((info == 0) & torch.isfinite(lo) & (lo > threshold * hi))
    if bool(ok.all().item()):
        v = _brot(linv, block)
        again, info2 = _cholesky_rows(v, used)
        good = info2 == 0
        return torch.where(good[:, None, None], again, v), used
    # padded rows get negative eigenvalues: they sort last after the flip, and
    # s = sqrt(max(w, 0)) = 0 drops them
    w, u = _eigh(g, used, -1.0)
    w = torch.flip(w, dims=[-1])
    u = torch.flip(u, dims=[-1])
    s = torch.sqrt(torch.clamp(w, min=0.0))
    s0 = torch.clamp(s[:, :1], min=1e-300)
    keep = (s > threshold * s0) & (torch.arange(s.shape[-1], device=s.device)[None, :] < used[:, None])
    kept = keep.sum(-1)
    inv = torch.where(keep, 1.0 / torch.where(keep, s, torch.ones_like(s)), torch.zeros_like(s))
    rot = (u.conj().transpose(-1, -2) * inv[:, :, None]).to(block.dtype)  # rows beyond kept are zero
    v = _brot(rot, block)
    again, info = _cholesky_rows(v, kept)
    ok = info == 0
    if bool(ok.all()):
        return again, kept
    return torch.where(ok[:, None, None], again, v), kept


_SOLVE_STREAMS: dict = {}


def _solve_streams(count: int) -> list:
    """Side streams (default priority) for the per-k eigensolves, created once per device."""
    dev = torch.cuda.current_device()
    have = _SOLVE_STREAMS.setdefault(dev, [])
    while len(have) < count:
        have.append(torch.cuda.Stream())
    return have[:count]


def _lowest(proj: torch.Tensor, used, nbands: int):
    """Lowest ``nbands`` eigenpairs of the leading used[k] x used[k] block of each projected
    matrix, at its own size with the per-k Davidson's solver (the library's for 96..2048 rows):
    eigenvalues (nk, nbands) and eigenvector rows (nk, nbands, n), zero on padded coordinates.

    The solves of different k-points are independent chains of small, latency-bound launches,
    so they run concurrently, one side stream and one workspace per k, ordered against the
    caller's stream by events."""
    nk, n = proj.shape[0], proj.shape[-1]
    used = [int(u) for u in (used.tolist() if isinstance(used, torch.Tensor) else used)]
    w = torch.empty((nk, nbands), dtype=torch.float64, device=proj.device)
    rot = torch.zeros((nk, nbands, n), dtype=proj.dtype, device=proj.device)
    main = torch.cuda.current_stream()
    if nk == 1:
        c = used[0]
        wk, zk = _lowest_eigenpairs(proj[0, :c, :c], nbands)
        w[0] = wk
        rot[0, :, :c] = zk
        return w, rot
    ready = torch.cuda.Event()
    ready.record(main)
    results = []
    for k, side in enumerate(_solve_streams(nk)):
        c = used[k]
        side.wait_event(ready)
        with torch.cuda.stream(side):
            blk = proj[k, :c, :c].contiguous()
            wk, zk = _lowest_eigenpairs(blk, nbands, workspace_tag=f"heev{k}")
            done = torch.cuda.Event()
            done.record(side)
        proj.record_stream(side)
        results.append((k, c, wk, zk, done))
    for k, c, wk, zk, done in results:
        main.wait_event(done)
        wk.record_stream(main)
        zk.record_stream(main)
        w[k] = wk
        rot[k, :, :c] = zk
    return w, rot


def _projected(basis, h_basis):
    """The projected matrices <b_r|H|b_c> (Hermitian part, FP64) of whole blocks."""
    return _herm(_bgram(basis, h_basis).conj())


def _rayleigh_ritz(proj, basis, h_basis, used, nbands):
    w, rot = _lowest(proj, used, nbands)
    return w, _brot(rot, basis), _brot(rot, h_basis)


def solve_batched(apply_block, guess: torch.Tensor, kinetic: torch.Tensor, npw, *, tolerance: float = 1e-8,
                  max_iterations: int = 40, subspace_factor: int = 4) -> BatchedResult:
    """Lowest ``nb`` eigenpairs at every k-point of ``guess`` ``[nk][nb][npw_pad]``.

    ``apply_block(x)`` returns H x for a zero-padded ``[nk][rows][npw_pad]`` block (the
    device Hamiltonian's ``apply_batched``); ``kinetic`` is ``[nk][npw_pad]`` (zero past
    each k's npw) in the precision of ``guess``; ``npw`` the per-k plane-wave counts.
    """
    nk, nb, npad = guess.shape
    dev = guess.device
    cdt = guess.dtype
    maxsub = [min(max(subspace_factor, 2) * nb, int(n)) for n in npw]
    cap = max(maxsub)
    full = torch.full((nk,), nb, dtype=torch.long, device=dev)
    v0, info = _cholesky_rows(guess.contiguous(), full)
    for k in torch.nonzero(info).flatten().tolist():  # rank-deficient guess: the SVD-equivalent route
        v0[k, :, : int(npw[k])] = orthonormalize(guess[k, :, : int(npw[k])].contiguous())
    basis = torch.zeros((nk, cap, npad), dtype=cdt, device=dev)
    hbasis = torch.zeros_like(basis)
    basis[:, :nb] = v0
    hbasis[:, :nb] = apply_block(v0)
    m = np.full(nk, nb)
    # the projected matrices are kept between steps: a step adds only the columns of its new
    # directions (8 m n_new npw flops instead of 8 m^2 npw), the matrix the reference rebuilds
    # from scratch (subspace.py:72) up to rounding; a collapse recomputes the Ritz block
    proj = torch.zeros((nk, cap, cap), dtype=torch.complex128, device=dev)
    proj[:, :nb, :nb] = _projected(basis[:, :nb], hbasis[:, :nb])
    eps, vec, hvec = _rayleigh_ritz(proj[:, :nb, :nb], basis[:, :nb], hbasis[:, :nb], full, nb)
    # frozen outputs of the k-points that have returned
    out_eps, out_vec = eps, vec
    out_norm = torch.full((nk, nb), float("inf"), dtype=torch.float64, device=dev)
    running = np.ones(nk, dtype=bool)
    iters = np.zeros(nk, dtype=int)
    kin = kinetic.to(dev).contiguous()
    for step in range(1, max_iterations + 1):
        iters[running] = step
        res, norms = ops.residual(hvec.reshape(nk * nb, npad), vec.reshape(nk * nb, npad), eps.reshape(-1))
        norms = norms.reshape(nk, nb)
        if running.all():
            out_norm = norms
        else:
            out_norm[torch.as_tensor(running, device=dev)] = norms[torch.as_tensor(running, device=dev)]
        active = (norms > tolerance).cpu().numpy() & running[:, None]
        finished = running & ~active.any(axis=1)
        running &= ~finished
        if not running.any():
            break
        nnew = active.sum(axis=1)
        width = int(nnew.max())
        # rows past a k-point's own count stay zero (empty only when every k fills the width)
        full_width = bool((nnew[running] == width).all()) and running.all()
        corr = (torch.empty if full_width else torch.zeros)((nk, width, npad), dtype=cdt, device=dev)
        for k in np.flatnonzero(running):
            rows = torch.as_tensor(np.flatnonzero(active[k]) + k * nb, dtype=torch.int32, device=dev)
            corr[k, : nnew[k]] = ops.precondition(res, vec.reshape(nk * nb, npad), kin[k], rows,
                                                  rows_checked=True)  # host-built from this k's bands
        # orthogonalise against the current basis (its padded rows are zero), twice
        cur = int(m.max())
        for _ in range(2):
            c = _bgram(corr, basis[:, :cur])
            ops.zgemm(c, basis[:, :cur], 0, alpha=-1.0, beta=1.0, out=corr)
        rn = torch.linalg.vector_norm(corr, dim=-1)
        keep = (rn > 1e-7).cpu().numpy() & (np.arange(width)[None, :] < nnew[:, None])
        nkeep = keep.sum(axis=1)
        if (nkeep != nnew).any():
            # compact the kept rows to the front of each k's block and zero the dropped ones
            order = np.argsort(~keep, axis=1, kind="stable")
            corr = torch.gather(corr, 1, torch.as_tensor(order, device=dev)[:, :, None].expand(-1, -1, npad))
            corr = corr * torch.as_tensor(np.arange(width)[None, :] < nkeep[:, None], device=dev)[:, :, None]
        new, kept_t = _gram_rows(corr.contiguous(), torch.as_tensor(nkeep, device=dev), 1e-7)
        kept = kept_t.cpu().numpy()
        stopped = running & (kept == 0)
        running &= ~stopped
        if not running.any():
            break
        kept = np.where(running, kept, 0)
        new = new.contiguous()
        hnew = apply_block(new)
        for k in np.flatnonzero(running):
            basis[k, m[k] : m[k] + kept[k]] = new[k, : kept[k]]
            hbasis[k, m[k] : m[k] + kept[k]] = hnew[k, : kept[k]]
        cur = int((m + kept).max())
        cols = _bgram(basis[:, :cur], hnew)  # conj of <b_r|H|new_j>, rows of new past kept are zero
        for k in np.flatnonzero(running):
            a, b = int(m[k]), int(m[k] + kept[k])
            ck = cols[k, :b, : kept[k]].conj().to(torch.complex128)
            proj[k, :a, a:b] = ck[:a]
            proj[k, a:b, :a] = ck[:a].conj().T
            proj[k, a:b, a:b] = _herm(ck[a:b])
        m = m + kept
        eps_n, vec_n, hvec_n = _rayleigh_ritz(proj[:, :cur, :cur], basis[:, :cur], hbasis[:, :cur],
                                              torch.as_tensor(m, device=dev), nb)
        if running.all():
            eps, vec, hvec = eps_n, vec_n, hvec_n
            out_eps, out_vec = eps, vec
        else:
            run_t = torch.as_tensor(running, device=dev)
            eps = torch.where(run_t[:, None], eps_n, eps)
            vec = torch.where(run_t[:, None, None], vec_n, vec)
            hvec = torch.where(run_t[:, None, None], hvec_n, hvec)
            out_eps = torch.where(run_t[:, None], eps, out_eps)
            out_vec = torch.where(run_t[:, None, None], vec, out_vec)
        for k in np.flatnonzero(running & (m + nb > np.asarray(maxsub))):
            if nk > 1:  # rows nb .. m[k] take part in the other k-points' padded products
                basis[k, nb : m[k]].zero_()
                hbasis[k, nb : m[k]].zero_()
            basis[k, :nb] = vec[k]
            hbasis[k, :nb] = hvec[k]
            proj[k].zero_()
            proj[k, :nb, :nb] = _projected(basis[k : k + 1, :nb], hbasis[k : k + 1, :nb])[0]
            m[k] = nb
    conv = (out_norm <= tolerance).all(dim=1).cpu().numpy()
    return BatchedResult(out_eps, out_vec, out_norm, iters, conv)


__all__ = ["BatchedResult", "solve_batched"]
