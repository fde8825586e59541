"""Device-resident state for one FFT grid and a set of k-point spheres.

A :class:`DevicePlan` owns the library plan (sphere->grid and stick maps, |k+G|^2/2,
twiddles) and the per-k Kleinman-Bylander tables in the padded k-batch layout
``[k][row][npw_pad]``. It is built once per geometry from the sphere and projector
objects the caller already has (amaterasu's ``PlaneWaveSystem.wave_spheres`` and
``.projectors`` -- ``hamiltonian/system.py:273`` and ``:339-342``).
"""

from __future__ import annotations

import ctypes
import itertools

import numpy as np
import torch

from . import autotune, ops
from ._lib import FP32, FP64, AmtError, check, lib

_SERIAL = itertools.count(1)


def _blocks_from_labels(labels, nproj: int):
    """Atom block (start, length) of every projector row; rows are laid out atom by atom."""
    start = np.zeros(nproj, dtype=np.int32)
    length = np.zeros(nproj, dtype=np.int32)
    if nproj == 0:
        return start, length
    atoms = np.array([lab[0] for lab in labels]) if labels else np.zeros(nproj, dtype=int)
    i = 0
    while i < nproj:
        j = i
        while j < nproj and atoms[j] == atoms[i]:
            j += 1
        start[i:j] = i
        length[i:j] = j - i
        i = j
    return start, length


def _is_block_diagonal(d: np.ndarray, start: np.ndarray, length: np.ndarray) -> bool:
    mask = np.zeros_like(d, dtype=bool)
    for s, n in set(zip(start.tolist(), length.tolist())):
        mask[s : s + n, s : s + n] = True
    return not np.any(d[~mask])


class DevicePlan:
    """Grid + spheres + projectors on the device, for FP64 and FP32 work."""

    def __init__(self, grid_shape, spheres, projectors=None, *, npw_pad: int | None = None, device="cuda"):
        self.shape = tuple(int(n) for n in grid_shape)
        self.nk = len(spheres)
        self.npw = [len(s.miller) for s in spheres]
        self.npw_pad = int(npw_pad or max(self.npw))
        if self.npw_pad < max(self.npw):
            raise AmtError("npw_pad smaller than the widest sphere")
        self.device = torch.device(device)
        miller = np.ascontiguousarray(np.concatenate([np.asarray(s.miller) for s in spheres]), dtype=np.int32)
        kin = np.ascontiguousarray(np.concatenate([0.5 * np.asarray(s.g2, dtype=float) for s in spheres]))
        npw = np.asarray(self.npw, dtype=np.int32)
        handle = ctypes.c_void_p()
        n1, n2, n3 = self.shape
        check(lib().amt_plan_create(ctypes.byref(handle), n1, n2, n3, self.nk,
                                    npw.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), self.npw_pad,
                                    miller.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                    kin.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "plan_create")
        self._handle = handle
        info = (ctypes.c_int * 16)()
        check(lib().amt_plan_info(handle, info, 16), "plan_info")
        self.info = list(info)
        self.smax = self.info[5]
        self.plane_ok = {FP32: self.info[10] > 0, FP64: self.info[11] > 0}
        # |k+G|^2/2 per k, padded with zeros, on the device
        kpad = np.zeros((self.nk, self.npw_pad))
        off = 0
        for k, n in enumerate(self.npw):
            kpad[k, :n] = kin[off : off + n]
            off += n
        self.kinetic = {FP64: torch.from_numpy(kpad).to(self.device)}
        self.kinetic[FP32] = self.kinetic[FP64].to(torch.float32)
        self._kb = {}  # precision -> (matrix [nk][nproj][npw_pad], dij [nk][nproj][nproj], start, length)
        self._attached = {}  # precision -> the kb tables whose pointers the library plan holds
        self._mat64 = None
        self._proj_src = projectors
        self.nproj = 0 if projectors is None else len(projectors[0].labels)
        self._shared_veff = None
        #: unique over the process; renewed whenever a device table a captured graph could
        #: hold (projector tables, shared potential buffers) is replaced
        self.serial = next(_SERIAL)

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                lib().amt_plan_destroy(h)
            except Exception:
                pass
            self._handle = None

    @property
    def handle(self):
        return self._handle

    def nsticks(self):
        out = (ctypes.c_int * self.nk)()
        check(lib().amt_plan_nsticks(self._handle, out), "plan_nsticks")
        return list(out)

    # ---------------------------------------------------------------- projectors
    def set_projectors(self, projectors) -> None:
        for prec in list(self._attached):
            check(lib().amt_plan_set_projectors(self._handle, prec, 0, None, None, None, None), "plan_set_projectors")
        self._attached = {}
        self._proj_src = projectors
        self._kb = {}
        self._mat64 = None
        self.nproj = 0 if projectors is None else len(projectors[0].labels)
        self.serial = next(_SERIAL)

    def _attach(self, prec: int):
        """Hand this precision's projector tables to the library plan (amt_plan_set_projectors);
        the tensors stay cached here, so the pointers the plan holds stay valid."""
        kb = self.kb_tables(prec)
        if kb is None:
            return
        if self._attached.get(prec) is not kb:
            mat, dij, start, length = kb
            check(lib().amt_plan_set_projectors(self._handle, prec, self.nproj, mat.data_ptr(), dij.data_ptr(),
                                                start.data_ptr(), length.data_ptr()), "plan_set_projectors")
            self._attached[prec] = kb

    def _matrix64(self) -> torch.Tensor:
        if self._mat64 is None:
            mat = torch.zeros((self.nk, self.nproj, self.npw_pad), dtype=torch.complex128, device=self.device)
            for k, p in enumerate(self._proj_src):
                m = p.matrix
                src = m if isinstance(m, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(m))
                mat[k, :, : self.npw[k]] = src.to(self.device, torch.complex128)
            self._mat64 = mat
        return self._mat64

    def kb_tables(self, prec: int, coupling=None):
        """Projector matrices, D and atom blocks at precision ``prec`` (built lazily).

        ``coupling`` (a PAW-style D_ij override, ``pseudo/kb.py:220-230``) is not cached.
        """
        if self._proj_src is None or self.nproj == 0:
            return None
        if coupling is None and prec in self._kb:
            return self._kb[prec]
        mat = self._matrix64()
        if prec != FP64:
            cached = self._kb.get(prec)
            mat = cached[0] if cached is not None else mat.to(ops.complex_dtype(prec)).contiguous()
        nproj = self.nproj
        start, length = _blocks_from_labels(self._proj_src[0].labels, nproj)
        ds = []
        for p in self._proj_src:
            d = np.asarray(p.dij if coupling is None else coupling, dtype=float)
            if not _is_block_diagonal(d, start, length):
                start[:] = 0
                length[:] = nproj
            ds.append(d)
        dij = torch.from_numpy(np.stack(ds)).to(self.device, ops.real_dtype(prec)).contiguous()
        tables = (mat, dij, torch.from_numpy(start).to(self.device), torch.from_numpy(length).to(self.device))
        if coupling is None:
            self._kb[prec] = tables
        return tables

    def shared_potential(self, pot: torch.Tensor) -> dict:
        """Copy a (nspin, *grid) potential into this plan's persistent FP64 and FP32 device
        buffers (allocated on first use, reallocated only when nspin changes) and return
        them as {FP64: ..., FP32: ...}. Graphs captured over H.psi keep reading them."""
        buf = self._shared_veff
        if buf is None or tuple(buf[FP64].shape) != tuple(pot.shape):
            buf = {FP64: torch.empty(tuple(pot.shape), dtype=torch.float64, device=self.device),
                   FP32: torch.empty(tuple(pot.shape), dtype=torch.float32, device=self.device)}
            self._shared_veff = buf
            self.serial = next(_SERIAL)
        buf[FP64].copy_(pot)
        buf[FP32].copy_(buf[FP64])
        return buf

    # ---------------------------------------------------------------- workspaces
    def sticks(self, prec: int, nrows: int) -> torch.Tensor:
        nbytes = lib().amt_plan_stick_bytes(self._handle, prec, nrows)
        return ops.workspace(nbytes, f"sticks{id(self)}")

    # ---------------------------------------------------------------- H.psi
    def hpsi(self, psi: torch.Tensor, veff: torch.Tensor, k0: int = 0, coupling=None,
             out: torch.Tensor | None = None) -> torch.Tensor:
        """H.psi for a [nkk][nb][npw_pad] block (or [nb][npw_pad] at k-point k0)."""
        single = psi.dim() == 2
        p3 = psi.unsqueeze(0) if single else psi
        if not p3.is_contiguous():
            p3 = p3.contiguous()
        nkk, nb, npw = p3.shape
        if npw != self.npw_pad:
            raise AmtError(f"band rows have {npw} columns, plan expects npw_pad={self.npw_pad}")
        if k0 < 0 or k0 + nkk > self.nk:
            raise AmtError("k-point range outside the plan")
        prec = ops.precision_of(p3)
        v = veff.to(device=self.device, dtype=ops.real_dtype(prec)).contiguous()
        if v.numel() != self.shape[0] * self.shape[1] * self.shape[2]:
            raise AmtError("effective potential does not match the grid")
        if out is None:
            h = torch.zeros_like(p3)
        else:
            h = out.unsqueeze(0) if out.dim() == 2 else out
            if tuple(h.shape) != tuple(p3.shape) or h.dtype != p3.dtype or not h.is_contiguous():
                raise AmtError("hpsi: out must be a contiguous block shaped like psi (zero padding)")
        if nb == 0:
            return h[0] if single else h
        if autotune.fft_enabled() and self.info[12 + prec] and self.plane_ok[prec]:
            autotune.tune_fft_launch(self, prec, p3, v, k0)
        if coupling is None:
            # the whole operator in one library call (local part + projector GEMMs)
            self._attach(prec)
            nbytes = lib().amt_hpsi_workspace(self._handle, prec, nb, nkk)
            ws = ops.workspace(nbytes, f"hpsi{id(self)}")
            check(lib().amt_hpsi(self._handle, prec, p3.data_ptr(), h.data_ptr(), nb, k0, nkk, v.data_ptr(),
                                 ws.data_ptr(), nbytes, ops.stream_handle()), "hpsi")
            return h[0] if single else h
        st = self.sticks(prec, nkk * nb)
        check(lib().amt_hpsi_local(self._handle, prec, p3.data_ptr(), h.data_ptr(), nb, k0, nkk, v.data_ptr(),
                                   st.data_ptr(), ops.stream_handle()), "hpsi_local")
        kb = self.kb_tables(prec, coupling)
        if kb is not None:
            mat, dij, start, length = kb
            m = mat[k0 : k0 + nkk]
            d = dij[k0 : k0 + nkk]
            proj = ops.zgemm(p3, m, 1)  # [nkk][nb][nproj] = psi beta^H
            pd = torch.empty_like(proj)
            check(lib().amt_kb_dij(prec, proj.data_ptr(), d.data_ptr(), start.data_ptr(), length.data_ptr(),
                                   pd.data_ptr(), nb, self.nproj, nb * self.nproj, self.nproj * self.nproj,
                                   nb * self.nproj, nkk, ops.stream_handle()), "kb_dij")
            ops.zgemm(pd, m, 0, alpha=1.0, beta=1.0, out=h)  # += (P D) beta
        return h[0] if single else h

    def set_route(self, prec: int, full_grid: bool) -> None:
        """Send this precision's local H.psi and density through the library's full-grid route
        (``full_grid=True``) even where the plane kernels fit, or back (amt_plan_set_route)."""
        check(lib().amt_plan_set_route(self._handle, prec, 1 if full_grid else 0), "plan_set_route")
        info = (ctypes.c_int * 16)()
        check(lib().amt_plan_info(self._handle, info, 16), "plan_info")
        self.info = list(info)
        self.plane_ok = {FP32: self.info[10] > 0, FP64: self.info[11] > 0}

    # ---------------------------------------------------------------- density
    def accumulate_density(self, psi: torch.Tensor, rows: np.ndarray, weights: np.ndarray, rho: torch.Tensor,
                           *, k0: int = 0, beta: float = 1.0, scale: float = 1.0, chunk: int = 8) -> None:
        """rho = beta rho + scale sum_l weights[l] |u_{rows[l]}(r)|^2 for a [nkk][nb][npw_pad] block."""
        p3 = psi.unsqueeze(0) if psi.dim() == 2 else psi
        if not p3.is_contiguous():
            p3 = p3.contiguous()
        nkk, nb, _ = p3.shape
        prec = ops.precision_of(p3)
        if rho.dtype != torch.float64 or not rho.is_contiguous():
            raise AmtError("rho must be a contiguous float64 device tensor")
        nlist = int(len(rows))
        rows = np.asarray(rows, dtype=np.int32)
        if nlist and (rows.min() < 0 or rows.max() >= nkk * nb):
            raise AmtError("density: band row outside the block")
        if len(weights) != nlist:
            raise AmtError("density: one weight per listed band")
        if k0 < 0 or k0 + nkk > self.nk or p3.shape[2] != self.npw_pad:
            raise AmtError("density: block does not match the plan")
        if rho.numel() != self.shape[0] * self.shape[1] * self.shape[2] or not rho.is_cuda:
            raise AmtError("density: rho must be one device grid")
        lst = torch.as_tensor(rows, device=self.device)
        wgt = torch.as_tensor(np.asarray(weights, dtype=np.float64), device=self.device)
        st = self.sticks(prec, nkk * nb)
        part = ops.workspace(max(lib().amt_density_part_bytes(self._handle, max(nlist, 1), chunk), 8),
                             f"dens{id(self)}")
        check(lib().amt_density(self._handle, prec, p3.data_ptr(), nb, k0, nkk, lst.data_ptr(), wgt.data_ptr(),
                                nlist, chunk, st.data_ptr(), part.data_ptr(), rho.data_ptr(), float(beta),
                                float(scale), ops.stream_handle()), "density")

    # ---------------------------------------------------------------- maps
    def scatter(self, psi: torch.Tensor, k: int) -> torch.Tensor:
        p2 = psi.unsqueeze(0) if psi.dim() == 1 else psi
        p2 = p2.contiguous()
        nb = p2.shape[0]
        if p2.shape[1] != self.npw_pad or not 0 <= k < self.nk:
            raise AmtError("scatter: rows must have npw_pad columns and k must be in the plan")
        grid = torch.empty((nb, *self.shape), dtype=p2.dtype, device=self.device)
        check(lib().amt_scatter(self._handle, ops.precision_of(p2), p2.data_ptr(), grid.data_ptr(), nb, k,
                                ops.stream_handle()), "scatter")
        return grid[0] if psi.dim() == 1 else grid

    def gather(self, grid: torch.Tensor, k: int, scale: float = 1.0) -> torch.Tensor:
        g = grid.unsqueeze(0) if grid.dim() == 3 else grid
        g = g.contiguous()
        nb = g.shape[0]
        if tuple(g.shape[1:]) != self.shape or not 0 <= k < self.nk:
            raise AmtError("gather: grid shape or k-point does not match the plan")
        out = torch.empty((nb, self.npw_pad), dtype=g.dtype, device=self.device)
        check(lib().amt_gather(self._handle, ops.precision_of(g), g.data_ptr(), out.data_ptr(), nb, k,
                               float(scale), ops.stream_handle()), "gather")
        return out[0] if grid.dim() == 3 else out
