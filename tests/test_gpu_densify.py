"""The batched densification workload (maple_amd.densify) on the GPU.

Each cell of a DensifyBatch follows the trajectory it has alone in the single-cell
drop-in optimize_inverse_homogenization with the worker's settings (voxel
initialisation, iso symmetry, overhang and trapped-void penalties at 0.9 with the
2-tile array check and channel radius 3, move 0.05, rho_min = xmin = 1e-5,
densify_pareto_property_network_32.py:184-201): J_total, J_raw, P_print and P_cavity
of every iteration within 1e-8, and the sample files and dataset_log.csv rows are
written.
"""
import csv
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0) if torch.cuda.is_available() else None


def _warm(r, seed, vf):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.0, 1.0, (r, r, r))
    for _ in range(3):   # a smooth field, mean vf
        rho = sum(np.roll(rho, s, a) for a in range(3) for s in (-1, 0, 1)) / 9.0
    rho = np.clip(rho - rho.mean() + vf, 1e-5, 1.0)
    return rho.astype(np.float32)


def test_densify_batch_follows_single_cell_runs(tmp_path):
    import maple_amd
    from maple_amd import densify
    from maple_amd.driver import optimize_inverse_homogenization
    r, iters = 16, 4
    specs = [("stiffness:0.5,thermal:0.5", 0.25, 1), ("stiffness:0.4,shear:0.3,thermal:0.3", 0.3, 2),
             ("shear:1.0", 0.2, 3)]
    tasks = [densify.Task(rho_warm=_warm(r, s, vf), objective=o, vf_target=vf, seed=s,
                          sample_dir=str(tmp_path / f"s{s}"), meta={"weight_tag": f"w{s}"})
             for o, vf, s in specs]
    print(f"[densify] {len(tasks)} cells at {r}^3, {iters} iterations", flush=True)
    db = densify.DensifyBatch(tasks, device=DEV, precision="fp64", rtol=1e-10, maxit=4000, islands=False,
                              cavity_iters=24)
    rows = db.run(iters=iters)
    db.close()
    s = densify.DENSIFY_DEFAULTS
    for t, row in zip(tasks, rows):
        ihd = maple_amd.IHD(maple_amd.make_cfg(r), DEV, rtol=1e-10, maxit=4000, nlevels=3)
        batch = maple_amd.build_unitcell_batch(DEV, r, torch.from_numpy(t.rho_warm.astype(np.float64)).to(DEV))
        out = str(tmp_path / f"single{t.seed}")
        res = optimize_inverse_homogenization(
            ihd, batch, res=r, objective=t.objective, init_mode="voxel", volfrac=t.vf_target, penal=s["penal"],
            rho_min=s["rho_min"], move=s["move"], xmin=s["xmin"], iters=iters, seed=t.seed, symmetry="iso",
            lambda_print=0.9, overhang_deg=45.0, lambda_cavity=0.9, min_channel_rad=3, cavity_iters=24,
            cavity_tiles=2, filter_k=3, output_dir=out, verbose=False)
        want = list(csv.reader(open(res.log_csv)))[1:]
        got = list(csv.reader(open(os.path.join(t.sample_dir, "iter_log.csv"))))
        assert got[0] == densify.ITER_LOG_HEADER and len(got) - 1 == len(want) == iters
        for g, w in zip(got[1:], want):
            for col in (1, 2, 3, 4):   # J_total, J_raw, P_print, P_cavity
                assert abs(float(g[col]) - float(w[col])) <= 1e-8 * max(abs(float(w[col])), 1e-12), (t.seed, col)
        rf = np.load(row["rho_final_npy"])
        assert rf.dtype == np.float32 and rf.shape == (r, r, r)
        assert np.abs(rf - res.rho_final).max() <= 1e-6
        assert row["vf_final"] == pytest.approx(float(rf.mean()))
        assert os.path.exists(row["condition_npy"]) and np.load(row["condition_npy"]).shape == (7,)
        assert row["weight_tag"] == f"w{t.seed}"
    log = str(tmp_path / "dataset_log.csv")
    densify.write_dataset_log(log, rows)
    assert list(csv.reader(open(log)))[0] == densify.FIELDNAMES
