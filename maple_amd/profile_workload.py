"""Small fixed workloads to run under rocprofv3 (kernel timing and HBM counters).

    python -m maple_amd.profile_workload --case operator
    python -m maple_amd.profile_workload --case elastic_sweep
    python -m maple_amd.profile_workload --case copy

``operator``: one 64^3 cell at the x0 of the 64^3 workload (0.25 + U(-0.05, 0.05),
seed 0), SIMP p = 3; 20 launches each of the FP64 level-0 elastic operator
(6 load cases) and, through two mixed-form homogenizations, its FP32
preconditioner instances.
``thermal``: the level-0 thermal Jacobi sweep (node form, three load cases): 20 FP64
sweeps of one 64^3 cell (the plain form's) and 20 FP32 sweeps of a 16-cell batch (the
mixed form's preconditioner), at the same density in every cell; --variant picks the kernel (3: one node per
thread, 0: element rows).
``elastic_sweep``: 20 FP32 Jacobi sweeps of the level-0 elastic operator over a
16-cell 64^3 batch (the mixed form's preconditioner at the 16-cell batch's shape:
6 load cases x 16 x 262144 = 25 165 824 element-cases per launch), same density.
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


def run_thermal(reps: int = 20, variant: int = 2):
    import maple_amd
    from maple_amd import _lib as L
    dev = torch.device("cuda", 0)
    r = 64
    x = maple_amd.x0_uniform_noise(r, seed=0)
    rho1 = torch.from_numpy(1e-9 + (1 - 1e-9) * x ** 3).to(dev)
    for ncells, prec in ((1, "fp64"), (16, "mixed")):
        cb = maple_amd.CellBatch([r] * ncells, nlevels=4, device=dev, precision="mixed")
        cb.set_variant(L.MAPLE_KERNEL_THERMAL_OPERATOR, variant)
        rho = rho1.repeat(ncells)
        keep = cb.mg_setup("thermal", rho, prec)
        u = torch.randn(3, 1, ncells * r ** 3, dtype=torch.float64, device=dev)
        f = torch.randn_like(u)
        for _ in range(reps):
            cb.mg_apply("thermal", 0, u, f, mode=3, precision=prec)
        torch.cuda.synchronize()
        print(f"thermal: {reps} {prec} Jacobi sweeps of 3 cases x {ncells} x {r}^3 (variant {variant})", flush=True)
        del keep
        cb.close()


def run_elastic_sweep(reps: int = 20, variant: int = -1):
    import maple_amd
    from maple_amd import _lib as L
    dev = torch.device("cuda", 0)
    r, ncells = 64, 16
    x = maple_amd.x0_uniform_noise(r, seed=0)
    rho = torch.from_numpy(1e-9 + (1 - 1e-9) * x ** 3).to(dev).repeat(ncells)
    cb = maple_amd.CellBatch([r] * ncells, nlevels=4, device=dev, precision="mixed")
    if variant >= 0:
        cb.set_variant(L.MAPLE_KERNEL_FINE_OPERATOR, variant)
    keep = cb.mg_setup("elastic", rho, "fp32")
    u = torch.randn(6, 3, ncells * r ** 3, dtype=torch.float64, device=dev)
    f = torch.randn_like(u)
    for _ in range(reps):
        cb.mg_apply("elastic", 0, u, f, mode=3, precision="fp32")
    torch.cuda.synchronize()
    print(f"elastic_sweep: {reps} fp32 Jacobi sweeps of 6 cases x {ncells} x {r}^3 "
          f"({6 * ncells * r ** 3} element-cases per launch, variant {variant})", flush=True)
    del keep
    cb.close()


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
    ap.add_argument("--case", choices=("operator", "thermal", "elastic_sweep", "copy"), required=True)
    ap.add_argument("--variant", type=int, default=None,
                    help="thermal: operator variant (0/1 element rows, 2 node form with two nodes per thread "
                         "(default), 3 one node per thread); elastic_sweep: level-0 operator variant "
                         "(default: the library's choice)")
    args = ap.parse_args()
    if args.case == "operator":
        run_operator()
    elif args.case == "thermal":
        run_thermal(variant=2 if args.variant is None else args.variant)
    elif args.case == "elastic_sweep":
        run_elastic_sweep(variant=-1 if args.variant is None else args.variant)
    else:
        run_copy()


if __name__ == "__main__":
    main()
