"""The C ABI library loads and exports every symbol include/amaterasu_amd.h declares.

No compute call is made here (this runs without a GPU)."""

from __future__ import annotations

import re

import pytest

from tests.conftest import REPO


def header_symbols():
    text = (REPO / "include" / "amaterasu_amd.h").read_text()
    return sorted(set(re.findall(r"\b(amt_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    syms = header_symbols()
    for name in ("amt_plan_create", "amt_hpsi_local", "amt_density", "amt_zgemm", "amt_fft3d", "amt_precondition"):
        assert name in syms


def test_library_exports_every_declared_symbol():
    from amaterasu_amd import _lib

    lib = _lib.lib()
    for name in header_symbols():
        assert hasattr(lib, name), name
    assert set(header_symbols()) == set(_lib.SIGNATURES), "ctypes table and header disagree"


def test_version_and_error_string():
    from amaterasu_amd import _lib

    assert _lib.lib().amt_version().startswith(b"amaterasu_amd")
    assert isinstance(_lib.last_error(), str)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from amaterasu_amd import _lib

    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "libamaterasu_amd.so")
    with pytest.raises(_lib.AmtError, match="not built"):
        _lib.lib()


def test_split_hook_reports_large_tile_products():
    """amt_zgemm_set_split (a host-side table, no GPU needed) records a split for 64 x 64-tile
    products and answers 2 for the large-tile path, whose K chunks are fixed."""
    from amaterasu_amd._lib import FP32, FP64, lib

    assert lib().amt_zgemm_set_split(FP32, 1, 512, 512, 100000, 1, 4) == 2
    assert lib().amt_zgemm_set_split(FP32, 0, 512, 512, 100000, 1, 4) == 0
    assert lib().amt_zgemm_set_split(FP64, 1, 512, 512, 100000, 1, 4) == 0
    assert lib().amt_zgemm_set_split(FP32, 0, 512, 512, 100000, 1, 0) == 0
    assert lib().amt_zgemm_set_split(FP64, 1, 512, 512, 100000, 1, 0) == 0
    assert lib().amt_zgemm_set_split(FP32, 1, 100, 512, 100000, 1, 300) == 1
