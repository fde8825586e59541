"""Torch-level wrappers around the C ABI: tensors in, tensors out, on the current stream.

Every function here requires CUDA (HIP) tensors of the right dtype and layout and
raises otherwise; nothing falls back to a CPU implementation.
"""

from __future__ import annotations

import ctypes

import torch

from . import autotune
from ._lib import FP32, FP64, AmtError, check, lib

_WORKSPACES: dict = {}


def stream_handle() -> int:
    return torch.cuda.current_stream().cuda_stream


def precision_of(t: torch.Tensor) -> int:
    if t.dtype in (torch.complex128, torch.float64):
        return FP64
    if t.dtype in (torch.complex64, torch.float32):
        return FP32
    raise AmtError(f"unsupported dtype {t.dtype}")


def complex_dtype(prec: int):
    return torch.complex128 if prec == FP64 else torch.complex64


def real_dtype(prec: int):
    return torch.float64 if prec == FP64 else torch.float32


def _require(t: torch.Tensor, name: str):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise AmtError(f"{name} must be a HIP device tensor")
    if not t.is_contiguous():
        raise AmtError(f"{name} must be contiguous")


def _require_rows(t: torch.Tensor, name: str):
    """A 2-D or 3-D device matrix (batch) whose rows are contiguous and do not overlap."""
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise AmtError(f"{name} must be a HIP device tensor")
    if t.dim() not in (2, 3):
        raise AmtError(f"{name} must be 2-D or 3-D")
    if t.shape[-1] > 1 and t.stride(-1) != 1:
        raise AmtError(f"{name} must have unit stride along its last axis")
    if t.shape[-2] > 1 and t.stride(-2) < t.shape[-1]:
        raise AmtError(f"{name} rows overlap")
    if t.dim() == 3 and t.shape[0] > 1 and t.stride(0) < t.shape[-2] * t.stride(-2) and t.shape[-2] > 0:
        raise AmtError(f"{name} batch entries overlap")


def workspace(nbytes: int, tag: str = "ws") -> torch.Tensor:
    """A cached device byte buffer of at least ``nbytes`` (grown, never shrunk)."""
    key = (tag, torch.cuda.current_device())
    buf = _WORKSPACES.get(key)
    if buf is None or buf.numel() < nbytes:
        buf = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device="cuda")
        _WORKSPACES[key] = buf
    return buf


def zgemm(a: torch.Tensor, b: torch.Tensor, op: int, *, alpha: float = 1.0, beta: float = 0.0,
          out: torch.Tensor | None = None) -> torch.Tensor:
    """``alpha * a @ op(b) + beta * out`` on the matrix cores.

    ``op = 0``: b is (k, n). ``op = 1``: b is (n, k) and ``conj(b)^T`` is used.
    2-D operands, or 3-D operands batched over the leading axis.
    Operands may be strided views: each matrix needs unit stride along its rows' elements,
    a row stride of at least its column count and any batch stride.
    """
    _require_rows(a, "a")
    _require_rows(b, "b")
    if a.dtype != b.dtype or a.dtype not in (torch.complex64, torch.complex128):
        raise AmtError("zgemm operands must share a complex dtype")
    batched = a.dim() == 3
    a3 = a if batched else a.unsqueeze(0)
    b3 = b if batched else b.unsqueeze(0)
    nb, m, k = a3.shape
    if op == 0:
        kb, n = b3.shape[1], b3.shape[2]
    else:
        n, kb = b3.shape[1], b3.shape[2]
    if kb != k or b3.shape[0] != nb:
        raise AmtError(f"zgemm shape mismatch: a {tuple(a.shape)}, b {tuple(b.shape)}, op {op}")
    if out is None:
        if beta != 0.0:
            raise AmtError("beta != 0 needs out")
        out = torch.empty((nb, m, n) if batched else (m, n), dtype=a.dtype, device=a.device)
    _require_rows(out, "out")
    c3 = out if batched else out.unsqueeze(0)
    if tuple(c3.shape) != (nb, m, n):
        raise AmtError(f"zgemm out has shape {tuple(out.shape)}, expected {(nb, m, n)}")
    prec = precision_of(a)

    def launch(dst, al, be):
        wsb = lib().amt_zgemm_workspace(prec, op, m, n, k, nb)
        ws = workspace(wsb, "gemm") if wsb else None
        check(lib().amt_zgemm(prec, op, m, n, k, float(al), a3.data_ptr(), a3.stride(1), a3.stride(0),
                              b3.data_ptr(), b3.stride(1), b3.stride(0), float(be), dst.data_ptr(), dst.stride(1),
                              dst.stride(0), nb, ws.data_ptr() if ws is not None else None, int(wsb),
                              stream_handle()), "zgemm")

    if autotune.enabled() and autotune.bucket(prec, op, m, n, k, nb) not in autotune._TUNED:
        scratch = torch.empty((nb, m, n), dtype=a.dtype, device=a.device)
        autotune.tune_zgemm(lambda: launch(scratch, 1.0, 0.0), prec, op, m, n, k, nb)
        del scratch
    launch(c3, alpha, beta)
    return out


class _FFTPlan:
    """amt_fft_plan_create / amt_fft_plan_destroy for one grid shape on one device."""

    def __init__(self, shape):
        handle = ctypes.c_void_p()
        check(lib().amt_fft_plan_create(ctypes.byref(handle), *(int(n) for n in shape)), "fft_plan_create")
        self.handle = handle

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                lib().amt_fft_plan_destroy(h)
            except Exception:
                pass
            self.handle = None


_FFT_PLANS: dict = {}


def fft_plan(shape) -> "_FFTPlan":
    """The (cached) transform plan of a grid shape on the current device: its twiddle tables are
    allocated once here, so the transforms themselves allocate nothing."""
    key = (tuple(int(n) for n in shape), torch.cuda.current_device())
    plan = _FFT_PLANS.get(key)
    if plan is None:
        plan = _FFTPlan(key[0])
        _FFT_PLANS[key] = plan
    return plan


def fft3d(x: torch.Tensor, sign: int, scale: float = 1.0, *, inplace: bool = False) -> torch.Tensor:
    """Batched 3D transform over the last three axes (sign +1: e^{+iGr}, -1: e^{-iGr})."""
    _require(x, "x")
    y = x if inplace else x.clone()
    n1, n2, n3 = y.shape[-3:]
    batch = int(y.numel() // (n1 * n2 * n3))
    plan = fft_plan((n1, n2, n3))
    check(lib().amt_fft3d(plan.handle, precision_of(y), batch, y.data_ptr(), int(sign), float(scale),
                          stream_handle()), "fft3d")
    return y


def residual(hv: torch.Tensor, v: torch.Tensor, eps: torch.Tensor):
    """R = HV - eps V and ||R|| per row (float64)."""
    _require(hv, "hv")
    _require(v, "v")
    if hv.shape != v.shape or hv.dtype != v.dtype or v.dim() != 2:
        raise AmtError(f"residual: hv {tuple(hv.shape)} and v {tuple(v.shape)} must be equal 2-D blocks")
    nrow, npw = v.shape
    if eps.numel() != nrow:
        raise AmtError("residual: one eigenvalue per row")
    res = torch.empty_like(v)
    norms = torch.empty(nrow, dtype=torch.float64, device=v.device)
    e = eps.to(device=v.device, dtype=torch.float64).contiguous()
    check(lib().amt_residual(precision_of(v), hv.data_ptr(), v.data_ptr(), e.data_ptr(), res.data_ptr(),
                             norms.data_ptr(), nrow, npw, stream_handle()), "residual")
    return res, norms


def precondition(res: torch.Tensor, v: torch.Tensor, kinetic: torch.Tensor, rows: torch.Tensor, *,
                 rows_checked: bool = False) -> torch.Tensor:
    """Teter-Payne-Allan preconditioned, normalised residuals of the listed rows.
    ``rows_checked=True`` skips the range check of ``rows`` (two device reads), for callers
    whose rows come from indices of ``v`` itself (``nonzero`` of a per-row mask)."""
    _require(res, "res")
    _require(v, "v")
    if res.shape != v.shape or res.dtype != v.dtype or v.dim() != 2:
        raise AmtError("precondition: res and v must be equal 2-D blocks")
    prec = precision_of(v)
    kin = kinetic.to(device=v.device, dtype=torch.float64 if prec == FP64 else torch.float32).contiguous()
    if kin.numel() != v.shape[1]:
        raise AmtError("precondition: kinetic must have one entry per plane wave")
    r = rows.to(device=v.device, dtype=torch.int32).contiguous()
    if r.numel() and not rows_checked:
        lo, hi = torch.stack(torch.aminmax(r)).tolist()  # one device read
        if lo < 0 or hi >= v.shape[0]:
            raise AmtError("precondition: row index out of range")
    nout = int(r.numel())
    out = torch.empty((nout, v.shape[1]), dtype=v.dtype, device=v.device)
    check(lib().amt_precondition(prec, res.data_ptr(), v.data_ptr(), kin.data_ptr(), r.data_ptr(),
                                 out.data_ptr(), nout, v.shape[1], stream_handle()), "precondition")
    return out


def heev_lowest(a: torch.Tensor, nev: int, *, workspace_tag: str = "heev"):
    """Lowest ``nev`` eigenpairs of a Hermitian complex128 matrix on the device: (w, zt) with
    w ascending (float64) and zt[k] the k-th eigenvector (row). See amt_heev_lowest. Calls
    that may run at the same time (on different streams) need different ``workspace_tag``s."""
    _require(a, "a")
    if a.dtype != torch.complex128 or a.dim() != 2 or a.shape[0] != a.shape[1]:
        raise AmtError("heev: a must be a square complex128 device matrix")
    n = a.shape[0]
    if not 1 <= nev <= n:
        raise AmtError(f"heev: nev={nev} outside 1..{n}")
    w = torch.empty(nev, dtype=torch.float64, device=a.device)
    zt = torch.empty((nev, n), dtype=torch.complex128, device=a.device)
    nbytes = int(lib().amt_heev_workspace(n))
    ws = workspace(nbytes, workspace_tag)
    check(lib().amt_heev_lowest(n, a.data_ptr(), nev, w.data_ptr(), zt.data_ptr(), ws.data_ptr(), nbytes,
                                stream_handle()), "heev")
    return w, zt

