"""Benchmark: OC iterations per second of the periodic inverse-homogenization loop at 64^3.

Workload: 64^3 periodic cells, E0=1, nu=0.3, Emin=kmin=1e-9, SIMP p=3, box
filter k=3 on dJ/dx, 6 elastic unit-strain + 3 thermal unit-gradient cell
problems, 4-level multigrid-preconditioned CG to rtol 1e-6, J = 0.5 stiffness +
0.5 kappa, volume fraction 0.25 on mean(x), move 0.2, x0 = 0.25 + U(-0.05, 0.05)
(numpy default_rng(seed), cell k has seed k).  One OC iteration of a cell = both
homogenizations, sensitivities, filter and OC update.  Every timed run is a
whole 100-iteration trajectory from x0, including the hard middle where small
disconnected islands slow the cell solves down.

Forms (one JSON line; ``value`` is the headline form):
* ``plain``    one cell, FP64, zero start every solve: the reference's method on the GPU.
* ``headline`` 16 cells in 4 concurrent groups of 4 (each group one set of
               launches per OC step, its own stream and host thread),
               FP64 CG around an FP32 multigrid preconditioner with the FP64
               island correction (small solid islands solved exactly beside the
               V-cycle) and its coarsest dense inverse rebuilt every third
               density (an earlier density's SPD inverse in between), each
               solve starting from the best of zero, the previous
               iterate's fields and those fields plus the change the learned
               network predicts between the two densities.
               value = cells x iterations / s.
* optional (--forms): single-cell and batched ablations (``fp64-warm``,
  ``fp64-islands``, ``fp64-block``, ``mixed-warm``, ``mixed-warm-block``,
  ``mixed-warm-learned``, ``batched-mixed-warm``, ``batched-mixed-warm-islands``,
  ``batched-mixed-warm-block``, ``headline-block`` with block CG over the load
  cases, ``headline-lockstep`` with all 16 cells in one group); --config4 adds
  the 64-cell mixed-resolution batch (64^3-equivalent iterations / s).

Accuracy: the last iterate's C_H and kappa of the headline form (every cell)
against a tight (rtol 1e-11) FP64 solve of the same densities.  ``roofline`` is
measured live for the dominant kernel (the level-0 elastic operator of the
headline form's preconditioner) with HIP events on the stream it runs on.
``cpu_baseline`` times the FP64 CPU oracle on one OC iteration of the same cell.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "OC iters/sec at 64^3 (mixed+batched+learned vs plain FP64 GPU) at C_H rel err ≤1e-5"
R = 64
NLEVELS = 4
OBJ = "stiffness:0.5,thermal:0.5"
PEAK = {"fp64": 78.6, "mixed": 157.3}
FLOP_PER_ELEM_ELASTIC = 1152      # 24 rows x 24 cols x 2 per element and load case
# HBM bytes per 64^3 cell and launch of the level-0 elastic operator, from rocprofv3
# FETCH_SIZE + WRITE_SIZE runs calibrated against a copy kernel
# (profiles/fine_operator_64cube_hbm_counters.md): FP64 product 93.5 MB, FP32 Jacobi sweep 81.5 MB
TRAFFIC_PER_CELL = {"fp64": 93.5e6, "mixed": 81.5e6}


def _x0(ncells):
    import numpy as np

    import maple_amd
    return np.concatenate([maple_amd.x0_uniform_noise(R, seed=k) for k in range(ncells)])


def run_form(steps: int, warmup: int, ncells: int = 1, precision: str = "fp64", warm: bool = False,
             learned: bool = False, block: bool = False, islands: bool = False, groups: int = 1, refresh: int = 1,
             maxit: int = 5000, profile_kernel: bool = True):
    """groups > 1: the cells advance in that many concurrent groups (maple_amd.concurrent),
    each its own batch, stream and host thread; every cell's trajectory is unchanged."""
    import numpy as np
    import torch

    import maple_amd
    from maple_amd.concurrent import ConcurrentOC

    dev = torch.device("cuda", 0)
    per = ncells // groups

    def make():
        out = []
        for g in range(groups):
            cells = maple_amd.CellBatch([R] * per, nlevels=NLEVELS, device=dev, warm_start=warm,
                                        precision=precision, block_cg=block, islands=islands)
            cells.set_coarse_refresh(refresh)
            if groups > 1:
                # the groups already overlap on the device: each solves its two physics in turn
                # on its own stream (4 hardware queues per process)
                cells.set_concurrency(False)
            start = None
            if learned:
                from maple_amd import learned as ml
                start = ml.LearnedStart(cells, ml.UNet.load(max_batch=min(16, per), device=dev), mode="delta")
            drv = maple_amd.OCDriver(cells, OBJ, 0.25, move=0.2, xmin=1e-3, filter_k=3, rtol=1e-6, maxit=maxit,
                                     start=start)
            out.append((cells, drv))
        return out

    x0 = torch.from_numpy(_x0(ncells)).to(dev)
    xs = [x0[g * per * R ** 3:(g + 1) * per * R ** 3].clone() for g in range(groups)]
    if warmup:
        sets = make()
        ConcurrentOC([(c, d, x.clone()) for (c, d), x in zip(sets, xs)]).run(warmup)
        torch.cuda.synchronize()
        for c, _ in sets:
            c.close()
    sets = make()
    kind = 0 if precision == "fp64" else 2
    if profile_kernel:
        for c, _ in sets:
            c.timing(True)
            c.timing_read(kind)
    iters = np.zeros((groups, steps, per, 9), dtype=np.int64)

    def on_step(g, k, info):
        iters[g, k] = info["iters"]

    co = ConcurrentOC([(c, d, x.clone()) for (c, d), x in zip(sets, xs)])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = co.run(steps, on_step)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    count, kms, work = 0, 0.0, 0
    if profile_kernel:
        for c, _ in sets:
            n_, ms_, w_ = c.timing_read(kind, work=True)
            count, kms, work = count + n_, kms + ms_, work + w_
            c.timing(False)
    # accuracy: the last step's C_H and kappa (solved at info["rho"]) against a tight FP64 solve
    rho = torch.cat([r[1]["rho"] for r in res])
    CHl = np.concatenate([r[1]["CH"] for r in res])
    kl = np.concatenate([r[1]["kappa"] for r in res])
    ref = maple_amd.CellBatch([R] * ncells, nlevels=NLEVELS, device=dev)
    CHt, kt, _ = ref.homogenize(rho, ("elastic", "thermal"), rtol=1e-11, maxit=4 * maxit, strict=False)
    CH6 = torch.from_numpy(CHl).to(dev)
    k6 = torch.from_numpy(kl).to(dev)
    err = 0.0
    for c in range(ncells):
        err = max(err, float(torch.linalg.norm(CH6[c] - CHt[c]) / torch.linalg.norm(CHt[c])),
                  float(torch.linalg.norm(k6[c] - kt[c]) / torch.linalg.norm(kt[c])))
    ref.close()
    src = None
    if learned or warm:
        src = {p: np.sum([np.bincount(c.start_sources(p).ravel(), minlength=3) for c, _ in sets], axis=0).tolist()
               for p in ("elastic", "thermal")}
    J_last = float(res[0][1]["J"][0])
    for c, _ in sets:
        c.close()
    dt = t1 - t0
    # per step: the largest count over the cells of a group, averaged over the groups
    it_el = iters[..., :6].max(axis=(2, 3))
    it_th = iters[..., 6:].max(axis=(2, 3))
    out = {"value": ncells * steps / dt, "ms_per_step": dt * 1e3 / steps, "cells": ncells, "groups": groups,
           "iters_el_max": int(it_el.max()), "iters_el_mean": float(it_el.mean()),
           "iters_th_mean": float(it_th.mean()), "ch_rel_err": err, "J_last": J_last,
           "last_starts": src}
    if count:
        # launches skip converged cells and load cases: the work of each launch is
        # counted on the device (elements x load cases it processed)
        avg_s = kms / count * 1e-3
        flops = work / count * FLOP_PER_ELEM_ELASTIC
        tname = "double" if precision == "fp64" else "float"
        out["kernel"] = {"name": f"k_fine_nodeb64<{tname},*,3> (level-0 elastic operator, node blocks, "
                                 f"up to 6 load cases x {per} cells per launch)",
                         "launches": count, "avg_us": avg_s * 1e6, "flop_per_launch": flops,
                         "elem_cases_per_launch": work / count,
                         "achieved_tflops": flops / avg_s / 1e12, "peak": PEAK[precision],
                         "traffic": work / count / (6 * R ** 3) * TRAFFIC_PER_CELL[precision]}
    return out


def run_config4(steps: int, precision: str = "mixed", maxit: int = 5000):
    """The 64-cell mixed-resolution batch (16 x 32^3, 32 x 48^3, 16 x 64^3)."""
    import numpy as np
    import torch

    import maple_amd

    dev = torch.device("cuda", 0)
    spec = maple_amd.batched_workload()
    cells = maple_amd.CellBatch([c["r"] for c in spec], nlevels=NLEVELS, device=dev, precision=precision,
                                warm_start=True)
    cells.set_islands(True)
    cells.set_coarse_refresh(3)
    drv = maple_amd.OCDriver(cells, [c["weights"] for c in spec], [c["volfrac"] for c in spec], move=0.2, xmin=1e-3,
                             filter_k=3, rtol=1e-6, maxit=maxit)
    x = torch.from_numpy(np.concatenate([c["x0"] for c in spec])).to(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        x, info = drv.step(x)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    elems = sum(c["r"] ** 3 for c in spec)
    cells.close()
    return {"steps": steps, "ms_per_step": dt * 1e3 / steps, "cells": len(spec), "elements": elems,
            "cell_oc_iters_per_s": len(spec) * steps / dt, "equiv_64cube_oc_iters_per_s": elems / R ** 3 * steps / dt}


def cpu_baseline(n_iter: int):
    from oracle import driver

    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    x0 = driver.x0_uniform_noise(R, seed=0)
    t0 = time.perf_counter()
    x, hist, times = driver.run(R, x0, {"stiffness": 0.5, "thermal": 0.5}, n_iter, volfrac=0.25, nlevels=NLEVELS,
                                rtol=1e-6, move=0.2, xmin=1e-3)
    total = time.perf_counter() - t0
    per = sum(times) / len(times)
    return {"value": 1.0 / per, "unit": "OC iters/sec", "cores": cores, "kind": "port",
            "sample": f"{n_iter} OC iteration(s) of the same 64^3 workload from x0, FP64 numpy/scipy oracle "
                      f"(multithreaded BLAS/sparse kernels), {total:.1f} s wall"}


FORMS = {
    "plain": dict(ncells=1, precision="fp64", warm=False, learned=False, block=False, islands=False),
    "headline": dict(ncells=16, precision="mixed", warm=True, learned=True, block=False, islands=True, groups=4,
                     refresh=3),
    "headline-lockstep": dict(ncells=16, precision="mixed", warm=True, learned=True, block=False, islands=True,
                              refresh=3),
    "headline-block": dict(ncells=16, precision="mixed", warm=True, learned=True, block=True, islands=True,
                           groups=4, refresh=3),
    "fp64-warm": dict(ncells=1, precision="fp64", warm=True, learned=False, block=False, islands=False),
    "fp64-islands": dict(ncells=1, precision="fp64", warm=False, learned=False, block=False, islands=True),
    "fp64-block": dict(ncells=1, precision="fp64", warm=False, learned=False, block=True, islands=False),
    "mixed-warm": dict(ncells=1, precision="mixed", warm=True, learned=False, block=False, islands=False),
    "mixed-warm-block": dict(ncells=1, precision="mixed", warm=True, learned=False, block=True, islands=False),
    "mixed-warm-learned": dict(ncells=1, precision="mixed", warm=True, learned=True, block=False, islands=False),
    "batched-mixed-warm": dict(ncells=16, precision="mixed", warm=True, learned=False, block=False, islands=False),
    "batched-mixed-warm-islands": dict(ncells=16, precision="mixed", warm=True, learned=False, block=False,
                                       islands=True),
    "batched-mixed-warm-block": dict(ncells=16, precision="mixed", warm=True, learned=False, block=True,
                                     islands=False),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--forms", default="plain,headline,headline-lockstep", help="comma list of " + ", ".join(FORMS))
    ap.add_argument("--config4", action="store_true", help="also run the 64-cell mixed-resolution batch")
    ap.add_argument("--cpu-iters", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the solver runs on a single MI355X (cells shard by process "
              "per GPU with no collectives); --gpus above 1 is not supported.", file=sys.stderr)
        sys.exit(2)
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    for f in forms:
        if f not in FORMS:
            raise SystemExit(f"unknown form {f}")
    res = {}
    for f in forms:
        res[f] = run_form(args.steps, args.warmup, **FORMS[f])
        print(f"# {f}: {res[f]['value']:.3f} 64^3 OC it/s, C_H/kappa rel err {res[f]['ch_rel_err']:.2e}, "
              f"elastic CG mean {res[f]['iters_el_mean']:.1f}", file=sys.stderr, flush=True)
    ok = [f for f in forms if res[f]["ch_rel_err"] <= 1e-5]
    head = "headline" if "headline" in ok else (max(ok, key=lambda f: res[f]["value"]) if ok else forms[0])
    g = res[head]
    spec = FORMS[head]
    line = {
        "metric": METRIC,
        "value": g["value"],
        "unit": "OC iters/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": g["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if spec["precision"] == "fp64" else "f64+f32 (+bf16 network)" if spec["learned"] else "f64+f32",
        "data": "synthetic",
        "config": {"workload": f"{spec['ncells']} x 64^3 periodic cell(s), 100 OC iterations each from "
                               "x0=0.25+U(-0.05,0.05) (seed = cell index), J=0.5E+0.5kappa, vf 0.25, move 0.2, "
                               "filter k=3, 4-level MG-PCG rtol 1e-6",
                   "form": head, "cells": spec["ncells"], "precision": spec["precision"],
                   "warm_start": spec["warm"], "learned_start": spec["learned"], "block_cg": spec["block"],
                   "island_correction": spec["islands"], "groups": spec.get("groups", 1),
                   "coarsest_inverse_every": spec.get("refresh", 1),
                   "resolution": R,
                   "levels": NLEVELS, "rtol": 1e-6},
        "ch_rel_err_vs_fp64_tight_solve": g["ch_rel_err"],
        "forms": {f: {k: res[f][k] for k in ("value", "ms_per_step", "cells", "groups", "ch_rel_err", "iters_el_mean",
                                             "iters_th_mean", "iters_el_max", "last_starts")}
                  | {"kernel_avg_us": res[f].get("kernel", {}).get("avg_us")} for f in forms},
    }
    if "plain" in res and head != "plain":
        line["speedup_vs_plain_fp64"] = g["value"] / res["plain"]["value"]
    # roofline of the dominant kernel: from the lockstep run of the same workload when it
    # was run (its launches do not overlap another group's, so a launch's HIP-event span
    # is its own duration), else from the headline run itself
    rl_form = "headline-lockstep" if (head == "headline" and "headline-lockstep" in res) else head
    k = res[rl_form].get("kernel")
    if k:
        line["roofline"] = {"bound": "mfma", "achieved": k["achieved_tflops"], "peak": k["peak"],
                            "unit": "TFLOP/s", "frac": k["achieved_tflops"] / k["peak"], "traffic": k["traffic"],
                            "traffic_unit": "bytes per launch (rocprofv3 PMC, copy-calibrated)",
                            "kernel": k["name"], "avg_us": k["avg_us"], "launches": k["launches"],
                            "algorithmic_flop_per_launch": k["flop_per_launch"],
                            "measured_on_form": rl_form,
                            "roof": "vector FMA and MFMA share the MI355X peak for this precision "
                                    "(FP64 78.6, FP32 157.3 TFLOP/s)"}
    if args.config4:
        line["config4"] = run_config4(args.steps)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.cpu_iters)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
