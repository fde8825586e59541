"""Measured choices of launch parameters: the split along k for the complex GEMMs and the
launch shapes of the compile-time-length H.psi kernels.

``amt_zgemm`` splits a long-k product over several workgroups per output tile when the
output alone has too few tiles to fill the GPU, with a deterministic second-pass
reduction; how many splits pay depends on the shape (tiles, k) and on the machine. With
``AMT_AUTOTUNE=1``, the first time a product of at least ``MIN_GFLOP`` falls in a new bucket --
(precision, op, ceil(m/64), ceil(n/64), ceil(k/256), batch), the key the library uses -- this
module times the built-in heuristic's split and its neighbours within a factor of four on the
caller's stream with HIP events (one warm-up and ``REPEATS`` timed runs each, into a scratch
output so the caller's data is untouched), records the fastest with ``amt_zgemm_set_split``
and remembers the bucket for the rest of the process. The result of a product does not
depend on the split beyond rounding.

Products on the library's large-tile path (FP32 overlaps with both output sides >= 256)
have their split fixed by accuracy; the hook says so and they are left alone.

It is off by default: on the Si64 and Si8 shapes the heuristic was within a few per cent of
the best candidate, and tuning the many buckets of a short run (band counts change as bands
converge) costs more than it saves unless the process runs long enough to reuse them.
"""

from __future__ import annotations

import math
import os

import torch

from ._lib import check, lib

MIN_GFLOP = 2.0
REPEATS = 3
NEIGHBOURS = (0.25, 0.5, 1.0, 2.0, 4.0)  # multiples of the heuristic split that are timed

_TUNED: dict = {}


def enabled() -> bool:
    return os.environ.get("AMT_AUTOTUNE", "0") == "1"


def heuristic_split(m: int, n: int, k: int, batch: int, prec: int = 0, op: int = 1) -> int:
    """The library's built-in choice (``plan_gemm`` / ``choose_split`` in gemm.hip): the tile
    (80 rows where that pads less, for FP32 NN, FP32 NC with m < 256 and FP64 NN, unless the
    FP32 chunks would then need the two-level kernel; else 64 rows; 64 columns), then the
    smallest split giving about 1024 workgroups and, in FP32 (prec 0), K chunks of at most
    1024; then as many splits as fit in the same number of 1024-workgroup rounds; at least 256
    of k per split, at most 256 splits."""

    def split(rows):
        tiles = ((m + rows - 1) // rows) * ((n + 63) // 64) * batch
        if tiles >= 512:
            return 1
        single = (k + 1023) // 1024 if prec == 0 else 1
        want = max((1024 + tiles - 1) // tiles, single)
        rounds = max(1, math.ceil(tiles * want / 1024 - 0.05))
        want = max(rounds * 1024 // tiles, single)
        return int(max(1, min(want, (k + 255) // 256, 256)))

    tall = (m + 79) // 80 * 80 < (m + 63) // 64 * 64 and (op == 0 or (prec == 0 and m < 256))
    ns = split(80 if tall else 64)
    if tall and prec == 0 and (k + ns - 1) // ns > 1024:
        ns = split(64)
    return ns


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
    h = heuristic_split(m, n, k, batch, prec, op)
    if h == 1 and kmax == 1:
        _TUNED[key] = 0
        return 0
    best, best_ms = 0, float("inf")
    stream = torch.cuda.current_stream()
    candidates = {min(kmax, 256, max(1, int(round(h * f)))) for f in NEIGHBOURS} - {h}
    for split in [0] + sorted(candidates):
        rc = lib().amt_zgemm_set_split(prec, op, m, n, k, batch, split)
        if rc == 2:  # a large-tile product: its split is fixed, nothing to tune
            _TUNED[key] = 0
            return 0
        check(rc, "zgemm_set_split")
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


# ----------------------------------------------------------------------- FFT launch shapes
#: (grid length, precision) -> (plane workgroup, sticks per workgroup) chosen so far
_FFT: dict = {}


def fft_enabled() -> bool:
    """Launch shapes of the compile-time-length H.psi kernels are tuned unless AMT_AUTOTUNE_FFT=0."""
    return os.environ.get("AMT_AUTOTUNE_FFT", "1") != "0"


def tuned_fft_launches() -> dict:
    return dict(_FFT)


def _time_stage(plan, prec, stage, psi, out, v, st, k0, reps=REPEATS) -> float:
    from . import ops

    nkk, nb, _ = psi.shape

    def run():
        check(lib().amt_hpsi_stage(plan.handle, prec, stage, psi.data_ptr(), out.data_ptr(), nb, k0, nkk,
                                   v.data_ptr(), st.data_ptr(), ops.stream_handle()), "hpsi_stage")

    run()
    stream = torch.cuda.current_stream()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        run()
    stop.record(stream)
    stop.synchronize()
    return start.elapsed_time(stop) / reps


def tune_fft_launch(plan, prec: int, psi: torch.Tensor, v: torch.Tensor, k0: int = 0) -> tuple:
    """Pick the plane kernel's workgroup (256 or 512 threads) and the stick kernels' width (the
    default or half) for this grid length and precision by timing each candidate on the
    caller's block (read only; sticks and output go to scratch), with HIP events on the current
    stream, and set them on the plan (``amt_plan_set_launch``). The choice is remembered per
    (length, precision) and applied to later plans of that length without timing again. The
    result of H.psi does not depend on it (every stick and plane is transformed by the same
    arithmetic whatever the grouping)."""
    key = (plan.shape[0], prec)
    info = plan.info
    if key not in _FFT:
        info[12 + prec] = info[14 + prec] = -1  # the timing below leaves the plan on its last candidate
        nkk, nb, _ = psi.shape
        wdef = 32 if prec == 1 else 64
        st = torch.empty(lib().amt_plan_stick_bytes(plan.handle, prec, nkk * nb), dtype=torch.uint8,
                         device=psi.device)
        out = torch.zeros_like(psi)
        times = {}
        for w in (wdef, wdef // 2):
            check(lib().amt_plan_set_launch(plan.handle, prec, 0, w), "plan_set_launch")
            times[("w", w)] = _time_stage(plan, prec, 1, psi, out, v, st, k0) + \
                _time_stage(plan, prec, 3, psi, out, v, st, k0)
        for nth in (256, 512):
            check(lib().amt_plan_set_launch(plan.handle, prec, nth, 0), "plan_set_launch")
            _time_stage(plan, prec, 1, psi, out, v, st, k0, reps=1)  # fresh sticks for the in-place plane pass
            times[("p", nth)] = _time_stage(plan, prec, 2, psi, out, v, st, k0)
        best_w = min((w for (kind, w) in times if kind == "w"), key=lambda w: times[("w", w)])
        best_p = min((n for (kind, n) in times if kind == "p"), key=lambda n: times[("p", n)])
        _FFT[key] = (best_p, best_w)
        del st, out
    nth, w = _FFT[key]
    if (info[12 + prec], info[14 + prec]) != (nth, w):
        check(lib().amt_plan_set_launch(plan.handle, prec, nth, w), "plan_set_launch")
        info[12 + prec], info[14 + prec] = nth, w
    return nth, w


__all__ = ["MIN_GFLOP", "NEIGHBOURS", "bucket", "enabled", "fft_enabled", "heuristic_split", "tune_fft_launch",
           "tune_zgemm", "tuned_fft_launches", "tuned_splits"]
