"""Streams of this process for concurrent work on one device (groups of cells, the tuner).

They are made once per device and reused by every later caller, and the first
``CellBatch`` of a device makes them before its context creates the library's own
streams.  Fresh streams per run cost twice: torch's caching allocator keeps separate
blocks per stream (each run's tensors become new device allocations, reserved memory
grew by 3.7 GB per 16-cell run), and HIP binds each new stream to one of its few
hardware queues (four by default), so a run's groups could end up sharing one.
Measured on the 16-cell headline in 4 groups, runs one after the other in one process:
with fresh streams one run in about six was 10-15 % slower (81 and 88 against 93-98 cell
it/s) with unchanged iteration counts; with these streams no run fell behind and
reserved memory stays flat.
"""
from __future__ import annotations

import threading
from typing import Dict, List

import torch

RESERVED = 8   # streams made before the first context of a device

_STREAMS: Dict[int, List[torch.cuda.Stream]] = {}
_LOCK = threading.Lock()


def _index(device) -> int:
    device = torch.device(device)
    return device.index if device.index is not None else torch.cuda.current_device()


def group_streams(device, n: int) -> List[torch.cuda.Stream]:
    """The first n streams of this process on `device` (made on first use, then reused)."""
    idx = _index(device)
    with _LOCK:
        have = _STREAMS.setdefault(idx, [])
        while len(have) < n:
            have.append(torch.cuda.Stream(device=torch.device("cuda", idx)))
        return have[:n]


def reserve(device) -> None:
    """Make the device's first RESERVED streams now (before any context of the library)."""
    group_streams(device, RESERVED)
