"""Kernel-variant tuner for a CellBatch.

For every tunable kernel the tuner times each variant on the batch's own
shapes (the real cell sizes, all load cases, HIP events on the batch's
stream) and keeps the fastest.  The choice depends on the cell mix: the
register-staged level-0 operator wins when occupancy is the limit, the
all-in-registers one when L1 re-loads are; the node-block forms (six or three
load cases per thread) when the arithmetic is.
"""
from __future__ import annotations

import torch

from . import _lib as L
from .cells import CellBatch

KERNELS = {"fine_operator": (L.MAPLE_KERNEL_FINE_OPERATOR, (0, 1, 2, 3, 4, 5, 6, 7, 8))}


def _time(fn, reps: int) -> float:
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


def autotune(cells: CellBatch, reps: int = 10, seed: int = 0) -> dict:
    """Pick the fastest variant of every tunable kernel for this batch; returns {kernel: (variant, ms)}."""
    g = torch.Generator(device=cells.device).manual_seed(seed)
    rho = torch.rand(cells.nelem, generator=g, device=cells.device, dtype=torch.float64) + 1e-3
    u = torch.randn(6, 3, cells.nelem, generator=g, device=cells.device, dtype=torch.float64)
    out = {}
    for name, (kid, variants) in KERNELS.items():
        times = {}
        for v in variants:
            cells.set_variant(kid, v)
            times[v] = _time(lambda: cells.apply_operator("elastic", rho, u), reps)
        best = min(times, key=times.get)
        cells.set_variant(kid, best)
        out[name] = (best, times)
    return out
