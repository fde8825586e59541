"""Measured choice of the split along k for the complex GEMMs.

``amt_zgemm`` splits a long-k product over several workgroups per output tile when the
output alone has too few tiles to fill the GPU, with a deterministic second-pass
reduction; how many splits pay depends on the shape (tiles, k) and on the machine. The
first time a product of at least ``MIN_GFLOP`` falls in a new bucket -- (precision, op,
ceil(m/64), ceil(n/64), ceil(k/256), batch), the key the library uses -- this module times
the candidate splits on the caller's stream with HIP events (one warm-up and ``REPEATS``
timed runs each, into a scratch output so the caller's data is untouched), records the
fastest with ``amt_zgemm_set_split`` and remembers the bucket for the rest of the process.
The result of a product does not depend on the split beyond rounding.

``AMT_AUTOTUNE=0`` turns it off (the built-in heuristic is used throughout).
"""

from __future__ import annotations

import os

import torch

from ._lib import check, lib

MIN_GFLOP = 2.0
REPEATS = 3
CANDIDATES = (1, 2, 4, 8, 16, 32, 64, 128)

_TUNED: dict = {}


def enabled() -> bool:
    return os.environ.get("AMT_AUTOTUNE", "1") != "0"


def bucket(prec: int, op: int, m: int, n: int, k: int, batch: int) -> tuple:
    return (prec, op, (m + 63) // 64, (n + 63) // 64, (k + 255) // 256, batch)


def tuned_splits() -> dict:
    """The buckets tuned so far and their splits (for reports)."""
    return dict(_TUNED)


def tune_zgemm(run, prec: int, op: int, m: int, n: int, k: int, batch: int) -> int:
    """Time ``run()`` (one product into scratch output) for each candidate split and keep the
    fastest for the bucket. Returns the chosen split (0 when nothing was tuned)."""
    key = bucket(prec, op, m, n, k, batch)
    if key in _TUNED:
        return _TUNED[key]
    if 8.0 * m * n * k * batch < MIN_GFLOP * 1e9:
        _TUNED[key] = 0
        return 0
    kmax = max(1, (k + 255) // 256)
    best, best_ms = 0, float("inf")
    stream = torch.cuda.current_stream()
    for split in sorted({c for c in CANDIDATES if c <= kmax} | {0}):
        check(lib().amt_zgemm_set_split(prec, op, m, n, k, batch, split), "zgemm_set_split")
        run()  # warm-up (and workspace growth)
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(REPEATS):
            run()
        stop.record(stream)
        stop.synchronize()
        ms = start.elapsed_time(stop) / REPEATS
        if ms < best_ms:
            best, best_ms = split, ms
    check(lib().amt_zgemm_set_split(prec, op, m, n, k, batch, best), "zgemm_set_split")
    _TUNED[key] = best
    return best


__all__ = ["CANDIDATES", "MIN_GFLOP", "bucket", "enabled", "tune_zgemm", "tuned_splits"]
