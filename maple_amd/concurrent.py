"""Independent OC trajectories advanced by concurrent groups of cells.

The cells of a batched workload follow independent OC trajectories (own seed,
objective weights, volume fraction).  Advancing them in lockstep -- one
``OCDriver.step`` = one OC iteration of every cell -- makes every step last as
long as its slowest cell's solves: cells reach their hard phases (islands and
thin members forming, hundreds of CG iterations) at different iterations, and
meanwhile the others wait.  ``ConcurrentOC`` splits the cells into groups, each
with its own ``CellBatch`` (one launch set per group), ``OCDriver``, HIP stream
and host thread: a group in an easy phase keeps stepping while another grinds
through a hard solve, and the GPU interleaves the groups' kernels.  Each group
is still a batch (its cells share every launch), and every cell's trajectory is
exactly the one it has alone: groups share nothing but the device.

The reference runs its densification workers as independent processes, one
trajectory each (densify:1055-1255, 1299-1332); here the independent
trajectories share one device and one process.
"""
from __future__ import annotations

import threading
import time
from typing import Callable, List, Sequence, Tuple

import torch

from .cells import CellBatch
from .optim import OCDriver
from .streams import group_streams


class ConcurrentOC:
    """groups: (CellBatch, OCDriver, x0) per group, built by the caller on any thread.

    ``run(steps)`` advances every group by `steps` OC iterations, each group on its own
    stream and host thread, and returns per group (x, info of the last step, per-step
    wall times in s).  The solver calls release the GIL, so the host threads overlap."""

    def __init__(self, groups: Sequence[Tuple[CellBatch, OCDriver, torch.Tensor]]):
        self.groups = list(groups)
        if len({g[0].device for g in self.groups}) > 1:
            raise ValueError("the groups of a ConcurrentOC must share one device")
        self.streams = group_streams(self.groups[0][0].device, len(self.groups)) if self.groups else []

    def run(self, steps: int, on_step: Callable[[int, int, dict], None] = None):
        n = len(self.groups)
        results: List = [None] * n
        errors: List = [None] * n
        # the groups' inputs were produced on the caller's stream
        for s in self.streams:
            s.wait_stream(torch.cuda.current_stream(s.device))

        def work(g: int):
            cells, drv, x = self.groups[g]
            try:
                with torch.cuda.stream(self.streams[g]):
                    times, info = [], None
                    for k in range(steps):
                        t0 = time.perf_counter()
                        x, info = drv.step(x)
                        times.append(time.perf_counter() - t0)
                        if on_step is not None:
                            on_step(g, k, info)
                    self.streams[g].synchronize()
                results[g] = (x, info, times)
            except BaseException as e:  # re-raised on the caller's thread
                errors[g] = e

        threads = [threading.Thread(target=work, args=(g,), daemon=True) for g in range(n)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for e in errors:
            if e is not None:
                raise e
        cur = torch.cuda.current_stream(self.streams[0].device) if self.streams else None
        for s in self.streams:
            cur.wait_stream(s)
        for g in range(n):
            self.groups[g] = (self.groups[g][0], self.groups[g][1], results[g][0])
        return results
