"""Kernel-variant tuner for a CellBatch.

For every tunable kernel the tuner times each variant on the batch's own shapes
(its real cell sizes, every load case, on the batch's stream) in the roles the
solver gives it, and keeps the fastest:

* ``fine_operator``: the level-0 elastic operator, one CG product (FP64) plus one
  smoother sweep of the preconditioner's precision (FP32 in the mixed form);
* ``thermal_operator``: the level-0 thermal smoother sweep;
* ``coarse_operator``: the smoother sweeps of every coarse stencil level.

The choice depends on the cell mix (for a batch with cells whose nz is not 64 the
lane-rotation variants do not apply and the 2x1 node blocks usually win), so it is
kept per batch shape (cell sizes, precision, level count) for the process:
``autotune(cells)`` on a second batch of the same shape reuses it.

``streams > 1`` times every variant as that many concurrent copies on separate
streams (best of two interleaved trials), for batches that will share the device with
as many others (concurrent groups): a kernel that wins alone can lose there.  Such
timings scatter by a few per cent, so a candidate replaces the library's default only
when it is faster by more than CONCURRENT_MARGIN.  The FP32 MFMA operator (variant 9)
is the case in point: alone it is a little faster than the vector node-block form for
4-cell 64^3 groups; timed as four concurrent copies the two lie within 2 %, and in the
16-cell headline (four such groups beside each other's coarse-level and vector
kernels) variant 9 runs 92-94 cell it/s against 97.5-100.4 for variant 8.
"""
from __future__ import annotations

import time
from typing import Dict, Optional, Tuple

import torch

from . import _lib as L
from .cells import CellBatch

KERNELS = {
    "fine_operator": (L.MAPLE_KERNEL_FINE_OPERATOR, (0, 1, 2, 3, 4, 5, 6, 7, 8, 9)),
    "thermal_operator": (L.MAPLE_KERNEL_THERMAL_OPERATOR, (0, 1, 2, 3)),
    "coarse_operator": (L.MAPLE_KERNEL_COARSE_OPERATOR, (0, 1, 2)),
}

# the library's default variant of each kernel (the fine operator's 8 becomes 5 in a
# batch with a cell whose nz is not 64)
DEFAULTS = {"fine_operator": 8, "thermal_operator": 3, "coarse_operator": 2}
CONCURRENT_MARGIN = 0.05

_CACHE: Dict[Tuple, Dict[str, int]] = {}


def _time(fn, reps: int, streams=None) -> float:
    """ms per repetition; with streams, every repetition issues fn once on each stream."""
    if not streams:
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def once():
        for st in streams:
            with torch.cuda.stream(st):
                fn()
    torch.cuda.synchronize()   # the inputs were made on the caller's stream
    once()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        once()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps


def choose(times: Dict[int, float], default: Optional[int], concurrent: bool) -> int:
    """The variant to keep: the fastest; under concurrent timing the default unless the
    fastest beats it by more than CONCURRENT_MARGIN."""
    best = min(times, key=times.get)
    if concurrent and default in times and times[best] > (1.0 - CONCURRENT_MARGIN) * times[default]:
        return default
    return best


def shape_key(cells: CellBatch, streams: int = 1) -> Tuple:
    return (tuple(cells.dims), cells.precision, cells.nlevels, int(streams))


def autotune(cells: CellBatch, reps: int = 5, seed: int = 0, force: bool = False, streams: int = 1) -> dict:
    """Pick the fastest variant of every tunable kernel for this batch shape and set it;
    returns {kernel: (variant, {variant: ms})} ({kernel: (variant, None)} from the cache).
    streams > 1: time each variant as that many concurrent copies (one per stream)."""
    key = shape_key(cells, streams)
    if not force and key in _CACHE:
        for name, v in _CACHE[key].items():
            cells.set_variant(KERNELS[name][0], v)
        return {name: (v, None) for name, v in _CACHE[key].items()}
    # streams of its own: timed on the process's group streams (maple_amd/streams.py) the
    # choice varied from run to run and was up to 13 % slower on the 4-group headline
    sts = [torch.cuda.Stream(device=cells.device) for _ in range(streams)] if streams > 1 else None
    g = torch.Generator(device=cells.device).manual_seed(seed)
    dev, n = cells.device, cells.nelem
    x = torch.rand(n, generator=g, device=dev, dtype=torch.float64)
    rho = 1e-9 + (1 - 1e-9) * (0.05 + 0.95 * x) ** 3
    pre = cells.precision
    keep = []
    for phys in ("elastic", "thermal"):
        keep.append(cells.mg_setup(phys, rho, "fp64"))
        if pre != "fp64":
            cells.mg_setup(phys, rho, pre)
    u = torch.randn(6, 3, n, generator=g, device=dev, dtype=torch.float64)
    ut = torch.randn(3, 1, n, generator=g, device=dev, dtype=torch.float64)
    uc = {l: torch.randn(6, 3, cells.num_nodes(l), generator=g, device=dev, dtype=torch.float64)
          for l in range(1, cells.nlevels - 1)}
    roles = {
        "fine_operator": lambda: (cells.apply_operator("elastic", rho, u),
                                  cells.mg_apply("elastic", 0, u, u, mode=3, precision=pre)),
        "thermal_operator": lambda: cells.mg_apply("thermal", 0, ut, ut, mode=3, precision=pre),
        "coarse_operator": lambda: [cells.mg_apply("elastic", l, uc[l], uc[l], mode=3, precision=pre) for l in uc],
    }
    out, chosen = {}, {}
    for name, (kid, variants) in KERNELS.items():
        if name == "coarse_operator" and not uc:
            continue
        times = {}
        for trial in range(2 if sts else 1):   # concurrent timings are noisier: best of two, interleaved
            for v in variants:
                cells.set_variant(kid, v)
                t = _time(roles[name], reps, sts)
                times[v] = min(times.get(v, t), t)
        best = choose(times, DEFAULTS.get(name), sts is not None)
        cells.set_variant(kid, best)
        out[name] = (best, times)
        chosen[name] = best
    _CACHE[key] = chosen
    del keep
    return out
