"""Train the warm-start network on pairs the solver generates, then measure its gain.

    python -m maple_amd.train_warmstart --out maple_amd/data/warmstart_unet32.npz

1. Pairs: batched OC runs at 32^3 (``learned.generate_pairs``; default 400
   trajectories x 50 iterates = 20 000 pairs, random seeds, vf ~ U[0.01, 0.40]).
2. Training: Adam with a cosine schedule on the GPU kernels of csrc/unet.hip.
3. Gain: on a held-out 64^3 OC trajectory (seed and objective unlike the training
   runs), CG iterations of every solve from a zero start against the same solve
   from the network's start, and the starts the solver picked.
Writes the weights (with the training history and the gain figures) to --out and
prints one JSON line with the summary.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np
import torch

from . import learned
from .cells import CellBatch
from .optim import OCDriver, x0_uniform_noise


def measure_gain(net: learned.UNet, dev, res: int = 64, iters: int = 60, probe_every: int = 5, seed: int = 7,
                 objective: str = "stiffness:0.4,shear:0.3,thermal:0.3", volfrac: float = 0.2, log=print):
    """Iterations of cold solves (zero start) vs learned-start solves on the designs of a
    held-out OC trajectory, probed every `probe_every` iterations."""
    cells = CellBatch([res], nlevels=4, device=dev)
    drv = OCDriver(cells, objective, volfrac, move=0.2, rtol=1e-6, maxit=5000, physics=("elastic", "thermal"))
    probe = CellBatch([res], nlevels=4, device=dev)
    start = learned.LearnedStart(probe, net)
    x = torch.from_numpy(x0_uniform_noise(res, seed=seed, vf=volfrac)).to(dev)
    rows = []
    for it in range(iters):
        x_new, info = drv.step(x)
        if it % probe_every == 0:
            rho = info["rho"]
            _, _, cold = probe.homogenize(rho, rtol=1e-6, maxit=5000)
            cold = cold[0].copy()
            # residual of the predicted start relative to the zero start's, per case
            ratio = []
            for phys, u in zip(("elastic", "thermal"), start.predict(rho)):
                f = probe.rhs(phys, rho)
                r = f - probe.apply_operator(phys, rho, u)
                ratio += (torch.linalg.norm(r.reshape(r.shape[0], -1), dim=1)
                          / torch.linalg.norm(f.reshape(f.shape[0], -1), dim=1)).tolist()
            start(rho)
            _, _, warm = probe.homogenize(rho, rtol=1e-6, maxit=5000)
            warm = warm[0].copy()
            src = np.concatenate([probe.start_sources("elastic")[0], probe.start_sources("thermal")[0]])
            rows.append({"oc_iter": it, "cold": cold.tolist(), "learned": warm.tolist(), "src": src.tolist(),
                         "start_residual_ratio": ratio})
            log(f"OC {it}: cold max el {cold[:6].max()} th {cold[6:].max()}  learned max el {warm[:6].max()} "
                f"th {warm[6:].max()}  starts {src.tolist()}  start residual / zero-start residual "
                f"{np.round(ratio, 3).tolist()}", flush=True)
        x = x_new
    cold_tot = sum(sum(r["cold"]) for r in rows)
    warm_tot = sum(sum(r["learned"]) for r in rows)
    cells.close()
    probe.close()
    ratios = np.array([r["start_residual_ratio"] for r in rows])
    return {"probes": rows, "cold_total_iters": cold_tot, "learned_total_iters": warm_tot,
            "iteration_ratio": warm_tot / max(1, cold_tot),
            "median_start_residual_ratio": float(np.median(ratios))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--cells", type=int, default=400)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--gain-iters", type=int, default=60)
    ap.add_argument("--unweighted", action="store_true", help="plain MSE instead of the density-weighted loss")
    args = ap.parse_args()
    dev = torch.device("cuda", 0)
    t0 = time.time()
    feat, tgt = learned.generate_pairs(args.cells, args.iters, res=32, seed=args.seed, device=dev)
    t_data = time.time() - t0
    scales = learned.channel_scales(tgt)
    print(f"{feat.shape[0]} pairs in {t_data:.1f} s; channel scales {np.round(scales, 5).tolist()}", flush=True)
    net = learned.UNet(n=32, max_batch=args.batch, device=dev, seed=0, scales=scales)
    net.set_loss(not args.unweighted)
    t1 = time.time()
    hist = learned.train(net, feat, tgt, epochs=args.epochs, batch=args.batch, lr=args.lr)
    t_train = time.time() - t1
    del feat, tgt
    torch.cuda.empty_cache()
    gain = measure_gain(net, dev, iters=args.gain_iters)
    net.save(args.out, history=np.array(hist), data_pairs=args.cells * args.iters, data_seconds=t_data,
             train_seconds=t_train, cold_total_iters=gain["cold_total_iters"],
             learned_total_iters=gain["learned_total_iters"])
    print(json.dumps({"pairs": args.cells * args.iters, "data_s": t_data, "train_s": t_train,
                      "final_train_loss": hist[-1][1], "final_val_loss": hist[-1][2],
                      "cold_total_iters": gain["cold_total_iters"], "learned_total_iters": gain["learned_total_iters"],
                      "iteration_ratio": gain["iteration_ratio"],
                      "median_start_residual_ratio": gain["median_start_residual_ratio"], "out": args.out}), flush=True)


if __name__ == "__main__":
    main()
