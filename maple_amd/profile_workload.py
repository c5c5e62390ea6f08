"""Small fixed workloads to run under rocprofv3 (kernel timing and HBM counters).

    python -m maple_amd.profile_workload --case operator
    python -m maple_amd.profile_workload --case copy

``operator``: one 64^3 cell at the x0 of the 64^3 workload (0.25 + U(-0.05, 0.05),
seed 0), SIMP p = 3; 20 launches each of the FP64 level-0 elastic operator
(6 load cases) and, through two mixed-form homogenizations, its FP32
preconditioner instances.
``copy``: 8 device-to-device copies of a 512 MiB buffer (a kernel of known HBM
traffic, to calibrate FETCH_SIZE / WRITE_SIZE: 512 MiB read and 512 MiB written
per copy).
"""
from __future__ import annotations

import argparse

import numpy as np
import torch


def run_operator(reps: int = 20):
    import maple_amd
    dev = torch.device("cuda", 0)
    r = 64
    x = maple_amd.x0_uniform_noise(r, seed=0)
    rho = torch.from_numpy(1e-9 + (1 - 1e-9) * x ** 3).to(dev)
    cb = maple_amd.CellBatch([r], nlevels=4, device=dev, precision="mixed")
    u = torch.randn(6, 3, r ** 3, dtype=torch.float64, device=dev)
    for _ in range(reps):
        cb.apply_operator("elastic", rho, u)
    for _ in range(2):
        cb.homogenize(rho, ("elastic",), rtol=1e-6)
    torch.cuda.synchronize()
    print(f"operator: {reps} FP64 launches of 6 x {r}^3, 2 mixed homogenizations "
          f"(CG iterations {cb.last_iters[0, :6].tolist()})", flush=True)


def run_copy(reps: int = 8, mib: int = 512):
    dev = torch.device("cuda", 0)
    a = torch.ones(mib * 2 ** 20 // 4, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copy: {reps} x {mib} MiB", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", choices=("operator", "copy"), required=True)
    args = ap.parse_args()
    if args.case == "operator":
        run_operator()
    else:
        run_copy()


if __name__ == "__main__":
    main()
