"""Block Davidson on the GPU with the call surface of ``amaterasu.scf.eigen``.

Mirrors ``scf/eigen/__init__.py:31-46`` (``solve``), ``scf/eigen/davidson.py:48-111``
(``davidson``, ``DavidsonResult``), ``scf/eigen/subspace.py:17-88`` and
``scf/eigen/precond.py:27-46``. The block-sized work runs in libamaterasu_amd:
residuals and the Teter-Payne-Allan step are fused kernels, every Gram/projected
matrix and every rotation is an MFMA complex GEMM. The small dense factorisations
(Cholesky, Hermitian eigensolve) stay in FP64 on the GPU (torch.linalg on ROCm).

One step differs in method, not in result: ``orthogonalize_against`` orthonormalises
the projected block by CholeskyQR2 when the Cholesky factor of its Gram matrix proves
that every singular value exceeds ``threshold * s_max``, and otherwise through the
eigen-decomposition of the Gram matrix followed by a Cholesky re-orthonormalisation,
where the reference takes a thin SVD of the (rows x npw) block (``subspace.py:56-60``).
All routes keep exactly the directions whose singular value exceeds
``threshold * s_max`` and return an orthonormal basis of their span (the Rayleigh-Ritz
step that follows depends on the span only); none needs a tall-skinny SVD on the device.

``apply_hamiltonian`` may be the reference's closure over a :class:`Hamiltonian` of this
package (device tensors flow through) or any numpy callable (blocks go through the host).
numpy in gives numpy out.
"""

from __future__ import annotations

import os
from collections import OrderedDict
from dataclasses import dataclass

import numpy as np
import torch

from . import ops
from ._lib import FP64, AmtError

_IMPLEMENTED = ("davidson",)


@dataclass
class DavidsonResult:
    """``davidson.py:33-45``."""

    eigenvalues: object
    coefficients: object
    residual_norms: object
    iterations: int
    converged: bool

    @property
    def nbands(self) -> int:
        return len(self.eigenvalues)


def available_eigensolvers() -> tuple:
    return _IMPLEMENTED


def _to_device(x, dtype=torch.complex128) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(device="cuda", dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x), dtype=np.complex128)).to("cuda", dtype)


def _gram(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """a @ b^H on the matrix cores."""
    if a.shape[0] == 0 or b.shape[0] == 0:
        return torch.zeros((a.shape[0], b.shape[0]), dtype=a.dtype, device=a.device)
    return ops.zgemm(a.contiguous(), b.contiguous(), 1)


def _rotate(rot: torch.Tensor, block: torch.Tensor) -> torch.Tensor:
    """rot @ block on the matrix cores."""
    return ops.zgemm(rot.to(block.dtype).contiguous(), block.contiguous(), 0)


def _hermitian(m: torch.Tensor) -> torch.Tensor:
    m = m.to(torch.complex128)
    return 0.5 * (m + m.conj().T)


def _cholesky_orthonormalize(block: torch.Tensor):
    """L^{-1} block with L L^H = block block^H; None if the Cholesky fails. One host
    synchronisation: the factorisation's status and the finiteness of L^{-1} (a finite
    L^{-1} times a finite block is finite) are read together."""
    s = _hermitian(_gram(block, block))
    lower, info = torch.linalg.cholesky_ex(s)
    eye = torch.eye(s.shape[0], dtype=s.dtype, device=s.device)
    linv = torch.linalg.solve_triangular(lower, eye, upper=False)
    if not bool(((info == 0) & torch.isfinite(torch.view_as_real(linv)).all()).item()):
        return None
    return _rotate(linv, block)


#: Largest |(V V^H - I)_ij| the Loewdin refinement accepts, per block precision: its error is
#: O(E^3), below the rounding of the block's own arithmetic for E under these bounds.
_LOWDIN_MAX = {torch.complex64: 2e-3, torch.complex128: 1e-5}


def _lowdin_refine(v: torch.Tensor):
    """Rows of ``v`` (already orthonormal up to E = v v^H - I) made orthonormal to O(E^3) by
    the series (I + E)^{-1/2} = I - E/2 + 3 E^2 / 8, the second CholeskyQR pass's job without
    a factorisation; None (one host read, as the Cholesky status would need) when E is too
    large or not finite, and the caller factorises."""
    g = _hermitian(_gram(v, v))
    eye = torch.eye(g.shape[0], dtype=g.dtype, device=g.device)
    e = g - eye
    rot = eye - 0.5 * e + 0.375 * (e @ e)
    err = float(torch.view_as_real(e).abs().max().item())
    if not err <= _LOWDIN_MAX[v.dtype]:  # also false for NaN
        return None
    return _rotate(rot, v)


def _cholesky_keeps_all(g: torch.Tensor, threshold: float):
    """L^{-1} when the Cholesky factor proves that every singular value of the block
    exceeds ``threshold`` times the largest (then every direction is kept), else None.

    With G = B B^H = L L^H the singular values of B are those of L, and
    s_min(L) >= 1 / ||L^{-1}||_F, s_max(L) <= ||L||_F: the test is rigorous, not a guess.
    """
    lower, info = torch.linalg.cholesky_ex(g)
    eye = torch.eye(g.shape[-1], dtype=g.dtype, device=g.device)
    linv = torch.linalg.solve_triangular(lower, eye, upper=False)
    lo = 1.0 / torch.linalg.matrix_norm(linv)
    ok = (info == 0) & torch.isfinite(lo) & (lo > threshold * torch.linalg.matrix_norm(lower))
    return linv if bool(ok.item()) else None


def _gram_orthonormalize(block: torch.Tensor, threshold: float) -> torch.Tensor:
    """Rows spanning the directions with singular value > threshold * s_max, orthonormal.

    When the Cholesky factor of the Gram matrix proves that no singular value falls
    below the threshold, the block is orthonormalised by one Cholesky pass and a Loewdin
    refinement (or a second Cholesky pass when the first left the rows far from
    orthonormal: CholeskyQR2) and spans the same subspace the SVD route keeps. Otherwise s^2 and the
    right singular vectors come from the eigen-decomposition of the Gram matrix
    (descending, as an SVD orders them), followed by one CholeskyQR pass.
    """
    g = _hermitian(_gram(block, block))
    linv = _cholesky_keeps_all(g, threshold)
    if linv is not None:
        v = _rotate(linv, block)
        again = _lowdin_refine(v)
        if again is None:
            again = _cholesky_orthonormalize(v)
        return v if again is None else again
    w, u = torch.linalg.eigh(g)
    w = torch.flip(w, dims=[0])
    u = torch.flip(u, dims=[1])
    s = torch.sqrt(torch.clamp(w, min=0.0))
    s0 = float(s[0].item()) if s.numel() else 0.0
    keep = s > threshold * max(s0, 1e-300)
    nkeep = int(keep.sum().item())
    if nkeep == 0:
        return block[:0]
    rot = (u[:, :nkeep].conj().T / s[:nkeep, None]).to(block.dtype)
    v = _rotate(rot, block)
    again = _cholesky_orthonormalize(v)
    return v if again is None else again


def orthonormalize(block):
    """``subspace.py:17-32``: Cholesky of block block^H, SVD-equivalent fallback."""
    host = not isinstance(block, torch.Tensor)
    b = _to_device(block) if host else block.contiguous()
    out = _cholesky_orthonormalize(b)
    if out is None:
        out = _gram_orthonormalize(b, 1e-10)
    return out.cpu().numpy() if host else out


def orthogonalize_against(block, basis, *, threshold: float = 1e-7, overwrite: bool = False):
    """``subspace.py:41-60``: two projections off ``basis``, drop rows with norm <= threshold,
    orthonormalise keeping singular values > threshold * s_max. ``overwrite=True`` lets the
    projections work in a contiguous device ``block`` in place (the caller's temporary)."""
    host = not isinstance(block, torch.Tensor)
    if host:
        b = _to_device(block)
    else:
        b = block.contiguous()  # a copy already when block was not contiguous
        if not overwrite and b.data_ptr() == block.data_ptr():
            b = b.clone()
    q = _to_device(basis, b.dtype) if host else basis.contiguous()
    for _ in range(2):
        if q.shape[0] and b.shape[0]:
            c = _gram(b, q)  # b q^H
            ops.zgemm(c, q, 0, alpha=-1.0, beta=1.0, out=b)
    norms = torch.linalg.vector_norm(b, dim=1)
    keep = torch.nonzero(norms > threshold).flatten()
    if keep.numel() != b.shape[0]:
        b = b.index_select(0, keep).contiguous()
    if b.shape[0] == 0:
        return b.cpu().numpy() if host else b
    out = _gram_orthonormalize(b, threshold)
    return out.cpu().numpy() if host else out


def rayleigh_ritz(basis, h_basis, nbands: int):
    """``subspace.py:63-76``: projected = conj(V) HV^T, Hermitised, eigh, lowest nbands."""
    host = not isinstance(basis, torch.Tensor)
    v = _to_device(basis) if host else basis
    hv = _to_device(h_basis, v.dtype) if host else h_basis
    m = _gram(v, hv)  # m[i][j] = sum_G v_i hv_j^*  =  conj(projected[i][j])
    proj = _hermitian(m.conj())
    w, vec = torch.linalg.eigh(proj)
    rot = vec[:, :nbands].T.contiguous()
    values, vv, hvv = w[:nbands], _rotate(rot, v), _rotate(rot, hv)
    if host:
        return values.cpu().numpy(), vv.cpu().numpy(), hvv.cpu().numpy()
    return values, vv, hvv


def teter_payne_allan(residual, kinetic, reference):
    """``precond.py:27-46`` elementwise (host or device, no normalisation)."""
    if isinstance(residual, torch.Tensor):
        ref = torch.clamp(torch.as_tensor(reference, dtype=torch.float64, device=residual.device).reshape(-1),
                          min=1e-8)
        kin = torch.as_tensor(kinetic, dtype=torch.float64, device=residual.device)
        x = kin[None, :] / ref[:, None]
        p = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x))
        return residual * (p / (p + 16.0 * x**4)).to(residual.dtype)
    ref = np.maximum(np.atleast_1d(np.asarray(reference, dtype=float)), 1e-8)
    x = np.asarray(kinetic)[None, :] / ref[:, None]
    p = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x))
    return residual * (p / (p + 16.0 * x**4))


def random_bands(npw: int, nbands: int, kinetic, seed: int = 0, *, device: bool = False):
    """``subspace.py:79-88``: the same numpy default_rng draws (bit-identical start), then the
    orthonormalisation on the GPU."""
    rng = np.random.default_rng(seed)
    block = rng.normal(size=(nbands, npw)) + 1j * rng.normal(size=(nbands, npw))
    block = block / (1.0 + np.asarray(kinetic))
    out = orthonormalize(_to_device(block))
    return out if device else out.cpu().numpy()


class _Operator:
    """Calls ``apply_hamiltonian`` with device tensors when it accepts them, else via numpy."""

    def __init__(self, fn, dtype):
        self.fn = fn
        self.dtype = dtype
        self.host = None

    def __call__(self, block: torch.Tensor) -> torch.Tensor:
        if self.host is None:
            try:
                out = self.fn(block)
            except AmtError:
                raise
            except (TypeError, RuntimeError, AttributeError):  # a host-only callable given a device tensor
                out = None
            if isinstance(out, torch.Tensor):
                self.host = False
                return out.to(self.dtype).contiguous()
            self.host = True
        if self.host:
            res = self.fn(block.to(torch.complex128).cpu().numpy())
            return _to_device(res, self.dtype)
        return self.fn(block).to(self.dtype).contiguous()


def graphs_enabled() -> bool:
    """Whether the SCF replays Davidson passes from captured graphs (``AMT_DAVIDSON_GRAPHS=1``)."""
    return os.environ.get("AMT_DAVIDSON_GRAPHS", "0") == "1"


def graphs_per_operator() -> int:
    """Graphs kept per operator identity (``AMT_DAVIDSON_GRAPHS_MAX``, default 512); the least
    recently replayed one goes first. A cached graph holds host-side nodes only."""
    return max(int(os.environ.get("AMT_DAVIDSON_GRAPHS_MAX", "512")), 1)

#: Operator identities whose buffers and graphs are kept at once.
GRAPH_OPERATORS = 8


class StepGraphs:
    """The expansion half of Davidson passes captured as HIP graphs, for one operator.

    One graph per (m, nn) -- m basis rows before the pass, nn new directions -- holds every
    launch between the two host reads of a pass (the active set, the orthogonalisation's
    Cholesky status): H on the nn new rows, the new columns of the projected matrix, the
    Rayleigh-Ritz eigensolve and the two rotations, and the residuals of the new Ritz pairs.
    The basis, H.basis and projected-matrix buffers belong to this object, so their
    addresses, baked into every graph, never change; the operator's own device inputs
    (potential, projector tables) must stay in place while its identity is unchanged.
    A shape is run eagerly the first time it is met (that run sizes every cached workspace
    the graph will use) and captured the second time; a capture during which a cached
    workspace had to grow is dropped, because the graph would hold the old buffer's
    address. Replays and eager runs issue the same kernels in the same order, so both give
    the same numbers. Graphs hold references to the workspaces they were captured with, so
    those buffers outlive any later growth of the cache.

    Every graph ends by copying its Ritz values, vectors, H.vectors, residuals and residual
    norms into this object's output buffers, so the graphs of one object share one memory
    pool of temporaries and a cached graph holds no device memory of its own. A pass
    consumes those outputs (preconditioner, collapse, the returned copies) before the next
    replay of any graph.
    """

    def __init__(self, cap: int, npw: int, nbands: int, dtype):
        self.basis = torch.zeros((cap, npw), dtype=dtype, device="cuda")
        self.h_basis = torch.zeros_like(self.basis)
        self.proj = torch.zeros((cap, cap), dtype=torch.complex128, device="cuda")
        blk = torch.empty((nbands, npw), dtype=dtype, device="cuda")
        real = torch.empty(nbands, dtype=torch.float64, device="cuda")
        self.out = (real, blk, torch.empty_like(blk), torch.empty_like(blk), torch.empty_like(real))
        self.graphs: OrderedDict = OrderedDict()
        self.seen: set = set()
        self.pool = torch.cuda.graph_pool_handle()
        self.captures = 0
        self.replays = 0

    def _capture(self, body):
        graph = torch.cuda.CUDAGraph()
        main = torch.cuda.current_stream()
        side = _capture_stream()
        side.wait_stream(main)
        with torch.cuda.stream(side):
            graph.capture_begin(pool=self.pool)
            try:
                for dst, src in zip(self.out, body()):
                    dst.copy_(src)
            finally:
                graph.capture_end()
        main.wait_stream(side)
        return graph, self.out

    def run(self, key, body):
        entry = self.graphs.get(key)
        if entry is not None:
            self.graphs.move_to_end(key)
            entry[0].replay()
            self.replays += 1
            return entry[1]
        if key not in self.seen:
            self.seen.add(key)
            return body()
        before = dict(ops._WORKSPACES)
        graph, out = self._capture(body)
        after = ops._WORKSPACES
        if len(after) != len(before) or any(after.get(k) is not v for k, v in before.items()):
            del graph
            return body()
        self.graphs[key] = (graph, out, list(after.values()))
        if len(self.graphs) > graphs_per_operator():
            self.graphs.popitem(last=False)
        self.captures += 1
        graph.replay()
        self.replays += 1
        return out


_STEP_GRAPHS: OrderedDict = OrderedDict()
_CAPTURE_STREAMS: dict = {}


def _capture_stream() -> torch.cuda.Stream:
    """A side stream per device (default priority) for captures: the null stream cannot capture."""
    dev = torch.cuda.current_device()
    st = _CAPTURE_STREAMS.get(dev)
    if st is None:
        st = _CAPTURE_STREAMS[dev] = torch.cuda.Stream()
    return st


def step_graphs(graph_key, cap: int, npw: int, nbands: int, dtype) -> StepGraphs:
    """The (cached) :class:`StepGraphs` of one operator identity and block shape."""
    key = (graph_key, cap, npw, nbands, dtype, torch.cuda.current_device())
    sg = _STEP_GRAPHS.get(key)
    if sg is None:
        sg = StepGraphs(cap, npw, nbands, dtype)
        _STEP_GRAPHS[key] = sg
        if len(_STEP_GRAPHS) > GRAPH_OPERATORS:
            _STEP_GRAPHS.popitem(last=False)
    else:
        _STEP_GRAPHS.move_to_end(key)
    return sg


def graph_stats() -> dict:
    """Captures and replays of every cached :class:`StepGraphs` (for tests and traces)."""
    return {"captures": sum(g.captures for g in _STEP_GRAPHS.values()),
            "replays": sum(g.replays for g in _STEP_GRAPHS.values()),
            "graphs": sum(len(g.graphs) for g in _STEP_GRAPHS.values())}


def clear_graphs() -> None:
    _STEP_GRAPHS.clear()


def _expand(op, basis, h_basis, proj, m: int, nn: int, nbands: int):
    """H on the new rows m..m+nn, the projected matrix's new columns, Rayleigh-Ritz and the
    residuals of the new Ritz pairs (``davidson.py:80-82,99-101``)."""
    h_basis[m : m + nn] = op(basis[m : m + nn])
    cols = _gram(basis[: m + nn], h_basis[m : m + nn]).conj()  # proj[:, new]
    proj[:m, m : m + nn] = cols[:m]
    proj[m : m + nn, :m] = cols[:m].conj().T
    proj[m : m + nn, m : m + nn] = _hermitian(cols[m:])
    eps, vec, hvec = _ritz(proj[: m + nn, : m + nn], basis[: m + nn], h_basis[: m + nn], nbands)
    res, norms = ops.residual(hvec, vec, eps)
    return eps, vec, hvec, res, norms


def davidson(apply_hamiltonian, guess, kinetic, *, tolerance: float = 1e-8, max_iterations: int = 40,
             subspace_factor: int = 4, precision="fp64", graph_key=None) -> DavidsonResult:
    """Lowest eigenpairs of a Hermitian operator from a starting block (``davidson.py:48-111``).

    The basis and H.basis live in preallocated ``[max_subspace][npw]`` device buffers, and
    the projected matrix of the Rayleigh-Ritz step is kept between passes: each pass adds
    only the columns of the new directions (8 m n_new npw flops instead of 8 m^2 npw) --
    the same matrix the reference rebuilds from scratch (``subspace.py:72``), up to
    rounding. After a collapse the block of the Ritz vectors is recomputed.

    ``graph_key`` (any hashable) asks for the expansion half of each pass to be replayed
    from HIP graphs (:class:`StepGraphs`): the caller promises that every device input of
    ``apply_hamiltonian`` stays at the same address with the same contents while it passes
    the same key. The iteration, and every number it produces, is the eager one's.
    """
    host = not isinstance(guess, torch.Tensor)
    cdt = torch.complex128 if precision in ("fp64", FP64, None) else torch.complex64
    op = _Operator(apply_hamiltonian, cdt)
    kin = torch.as_tensor(np.asarray(kinetic, dtype=float) if not isinstance(kinetic, torch.Tensor) else kinetic,
                          dtype=torch.float64).to("cuda")
    v0 = orthonormalize(_to_device(guess, cdt))
    nbands, npw = v0.shape
    max_subspace = min(max(subspace_factor, 2) * nbands, npw)
    cap = max(max_subspace, 2 * nbands)
    sg = None
    if graph_key is not None:
        sg = step_graphs(graph_key, cap, npw, nbands, cdt)
        basis, h_basis, proj = sg.basis, sg.h_basis, sg.proj
    else:
        basis = torch.zeros((cap, npw), dtype=cdt, device="cuda")
        h_basis = torch.zeros_like(basis)
        proj = torch.zeros((cap, cap), dtype=torch.complex128, device="cuda")
    basis[:nbands] = v0
    h_basis[:nbands] = op(v0)
    if sg is not None and op.host:
        raise AmtError("davidson: graph replay needs an operator that takes device blocks")
    m = nbands
    proj[:m, :m] = _hermitian(_gram(basis[:m], h_basis[:m]).conj())
    eps, vec, hvec = _ritz(proj[:m, :m], basis[:m], h_basis[:m], nbands)
    res, next_norms = ops.residual(hvec, vec, eps)
    norms = torch.full((nbands,), float("inf"), dtype=torch.float64, device="cuda")
    iteration = 0
    for step in range(1, max_iterations + 1):
        iteration = step
        # the residuals of this pass's Ritz pairs (a graph's replay overwrites its outputs)
        norms = next_norms if sg is None else next_norms.clone()
        active = torch.nonzero(norms > tolerance).flatten()
        if active.numel() == 0:
            break
        corr = ops.precondition(res, vec, kin, active, rows_checked=True)  # rows of nonzero(norms)
        new = orthogonalize_against(corr, basis[:m], overwrite=True)
        nn = new.shape[0]
        if nn == 0:
            break
        nn = min(nn, cap - m)
        basis[m : m + nn] = new[:nn]
        body = (lambda m=m, nn=nn: _expand(op, basis, h_basis, proj, m, nn, nbands))
        if sg is not None and HEEV_MIN <= m + nn <= HEEV_MAX:
            eps, vec, hvec, res, next_norms = sg.run((m, nn), body)
        else:
            eps, vec, hvec, res, next_norms = body()
        m += nn
        if m + nbands > max_subspace:
            basis[:nbands] = vec
            h_basis[:nbands] = hvec
            m = nbands
            proj[:m, :m] = _hermitian(_gram(basis[:m], h_basis[:m]).conj())
    if sg is not None:  # graph outputs are overwritten by the next replay
        eps, vec = eps.clone(), vec.clone()
    converged = bool((norms <= tolerance).all().item())
    if host:
        return DavidsonResult(eps.cpu().numpy(), vec.to(torch.complex128).cpu().numpy(), norms.cpu().numpy(),
                              iteration, converged)
    return DavidsonResult(eps, vec, norms, iteration, converged)


#: Projected matrices at least this large go to the library's own Hermitian eigensolver
#: (device tridiagonalisation); smaller ones to torch.linalg.eigh.
HEEV_MIN = 96
HEEV_MAX = 2048


def _lowest_eigenpairs(proj: torch.Tensor, nbands: int, workspace_tag: str = "heev"):
    """(w[:nbands], rows of eigenvectors) of a Hermitian FP64 matrix."""
    m = proj.shape[0]
    if HEEV_MIN <= m <= HEEV_MAX:
        return ops.heev_lowest(proj.contiguous(), nbands, workspace_tag=workspace_tag)
    w, vec = torch.linalg.eigh(proj)
    return w[:nbands], vec[:, :nbands].T.contiguous()


def _ritz(proj: torch.Tensor, basis: torch.Tensor, h_basis: torch.Tensor, nbands: int):
    """Lowest ``nbands`` eigenpairs of the projected matrix, rotated into the basis."""
    w, rot = _lowest_eigenpairs(proj, nbands)
    return w, _rotate(rot, basis), _rotate(rot, h_basis)


def solve(name, apply_hamiltonian, guess, kinetic, *, apply_overlap=None, **kwargs) -> DavidsonResult:
    """``eigen/__init__.py:31-46``."""
    if name != "davidson":
        raise NotImplementedError(
            f"eigensolver {name!r} is not implemented yet; use "
            f"{' or '.join(repr(x) for x in _IMPLEMENTED)} (scf.eigensolver)"
        )
    if apply_overlap is not None:
        raise NotImplementedError("the generalised (PAW) Davidson is not part of the HIP path")
    return davidson(apply_hamiltonian, guess, kinetic, **kwargs)


__all__ = ["DavidsonResult", "StepGraphs", "available_eigensolvers", "clear_graphs", "davidson", "graph_stats",
           "graphs_enabled", "orthogonalize_against", "orthonormalize", "random_bands", "rayleigh_ritz", "solve",
           "step_graphs", "teter_payne_allan"]
