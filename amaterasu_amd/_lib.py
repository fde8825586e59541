"""ctypes binding of libamaterasu_amd.so (the C ABI declared in include/amaterasu_amd.h).

The library is built in-tree (``python -m amaterasu_amd.build``). There is no fallback:
if the shared library is missing or fails to load, every compute entry point raises.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().with_name("libamaterasu_amd.so")

FP32, FP64 = 0, 1

_c_int, _c_dbl, _c_ll, _c_sz, _vp = ctypes.c_int, ctypes.c_double, ctypes.c_longlong, ctypes.c_size_t, ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_c_float = ctypes.c_float
_dp = ctypes.POINTER(ctypes.c_double)

#: name -> (restype, argtypes); every symbol the header declares.
SIGNATURES = {
    "amt_version": (ctypes.c_char_p, []),
    "amt_last_error": (ctypes.c_char_p, []),
    "amt_device_info": (_c_int, [ctypes.c_char_p, _c_int]),
    "amt_plan_create": (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _c_int, _c_int, _ip, _c_int, _ip, _dp]),
    "amt_plan_destroy": (_c_int, [_vp]),
    "amt_plan_info": (_c_int, [_vp, _ip, _c_int]),
    "amt_plan_nsticks": (_c_int, [_vp, _ip]),
    "amt_plan_stick_bytes": (_c_sz, [_vp, _c_int, _c_int]),
    "amt_plan_set_route": (_c_int, [_vp, _c_int, _c_int]),
    "amt_plan_set_launch": (_c_int, [_vp, _c_int, _c_int, _c_int]),
    "amt_hpsi_local": (_c_int, [_vp, _c_int, _vp, _vp, _c_int, _c_int, _c_int, _vp, _vp, _vp]),
    "amt_plan_set_projectors": (_c_int, [_vp, _c_int, _c_int, _vp, _vp, _vp, _vp]),
    "amt_hpsi_workspace": (_c_sz, [_vp, _c_int, _c_int, _c_int]),
    "amt_hpsi": (_c_int, [_vp, _c_int, _vp, _vp, _c_int, _c_int, _c_int, _vp, _vp, _c_sz, _vp]),
    "amt_hpsi_local_noalloc": (_c_int, [_vp, _c_int, _vp, _vp, _c_int, _c_int, _c_int, _vp, _vp]),
    "amt_hpsi_stage": (_c_int, [_vp, _c_int, _c_int, _vp, _vp, _c_int, _c_int, _c_int, _vp, _vp, _vp]),
    "amt_density_part_bytes": (_c_sz, [_vp, _c_int, _c_int]),
    "amt_density": (_c_int, [_vp, _c_int, _vp, _c_int, _c_int, _c_int, _vp, _vp, _c_int, _c_int, _vp, _vp, _vp,
                             _c_dbl, _c_dbl, _vp]),
    "amt_scatter": (_c_int, [_vp, _c_int, _vp, _vp, _c_int, _c_int, _vp]),
    "amt_gather": (_c_int, [_vp, _c_int, _vp, _vp, _c_int, _c_int, _c_dbl, _vp]),
    "amt_fft_plan_create": (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _c_int]),
    "amt_fft_plan_destroy": (_c_int, [_vp]),
    "amt_fft3d": (_c_int, [_vp, _c_int, _c_int, _vp, _c_int, _c_dbl, _vp]),
    "amt_zgemm_workspace": (_c_sz, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int]),
    "amt_zgemm_set_split": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int]),
    "amt_zgemm": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_dbl, _vp, _c_ll, _c_ll, _vp, _c_ll, _c_ll,
                           _c_dbl, _vp, _c_ll, _c_ll, _c_int, _vp, _c_sz, _vp]),
    "amt_residual": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _vp]),
    "amt_precondition": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _vp]),
    "amt_kb_dij": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_ll, _c_ll, _c_ll, _c_int, _vp]),
    "amt_heev_workspace": (_c_sz, [_c_int]),
    "amt_heev_lowest": (_c_int, [_c_int, _vp, _c_int, _vp, _vp, _vp, _c_sz, _vp]),
    "amt_mlp_param_count": (_c_int, []),
    "amt_mlp_features": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _ip, _c_dbl, _c_dbl, _vp, _vp]),
    "amt_mlp_forward": (_c_int, [_vp, _c_ll, _vp, _vp, _vp, _vp, _vp]),
    "amt_mlp_workspace": (_c_sz, [_c_ll]),
    "amt_mlp_loss_grad": (_c_int, [_vp, _vp, _c_ll, _c_dbl, _vp, _vp, _vp, _vp, _vp, _vp, _c_sz, _vp]),
    "amt_adam": (_c_int, [_vp, _vp, _vp, _vp, _c_int, _c_float, _c_float, _c_float, _c_float, _c_int, _vp]),
}


class AmtError(RuntimeError):
    """A failure reported by the HIP library (shape, launch or device error)."""


_LIB = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built or cannot be loaded."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise AmtError(
                f"{LIB_PATH.name} is not built; run `python -m amaterasu_amd.build` (hipcc, gfx950). "
                "There is no CPU fallback for the HIP path."
            )
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def check(status: int, what: str = "") -> None:
    if status != 0:
        msg = lib().amt_last_error().decode(errors="replace")
        raise AmtError(f"{what}: {msg}" if what else msg)


def last_error() -> str:
    return lib().amt_last_error().decode(errors="replace")
