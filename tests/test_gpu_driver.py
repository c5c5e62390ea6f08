"""GPU checks of the OC-driver drop-in (optimize_inverse_homogenization,
main_IHD_v3.py:492-891): the symmetry projection against the oracle's
restatement (1e-14), the driver's trajectory against the oracle's loop
(J 1e-8 relative, x 1e-7), early stop, and the output files."""
import csv
import json
import os

import numpy as np
import pytest
import torch

from oracle import driver as odrv, symmetry as osym

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0) if torch.cuda.is_available() else None


@pytest.mark.parametrize("symmetry", ["x", "xz", "octant", "cubic"])
def test_symmetrize_matches_oracle(symmetry):
    import maple_amd
    dims = [16, 8]
    cb = maple_amd.CellBatch(dims, nlevels=2, device=DEV)
    rng = np.random.default_rng(0)
    v = [rng.standard_normal(d ** 3) for d in dims]
    kind, axes = osym.parse_symmetry(symmetry)
    out = cb.symmetrize(torch.from_numpy(np.concatenate(v)).to(DEV), kind, axes).cpu().numpy()
    ref = np.concatenate([osym.symmetrize(vi, d, kind, axes) for vi, d in zip(v, dims)])
    assert np.abs(out - ref).max() <= 1e-14


@pytest.mark.parametrize("symmetry,objective", [("cubic", "bulk"), ("xy", "stiffness:0.5,thermal:0.5")])
def test_driver_matches_oracle(tmp_path, symmetry, objective):
    import maple_amd
    from maple_amd.driver import optimize_inverse_homogenization
    r, iters = 16, 6
    print(f"[driver] r={r} symmetry={symmetry} objective={objective}", flush=True)
    x0 = maple_amd.x0_sphere_void(r, seed=2)
    rho0 = 1e-9 + (1 - 1e-9) * x0 ** 3
    ihd = maple_amd.IHD(maple_amd.make_cfg(r), DEV, rtol=1e-10, maxit=4000, nlevels=3)
    batch = maple_amd.build_unitcell_batch(DEV, r, torch.from_numpy(rho0.reshape(r, r, r)).to(DEV))
    res = optimize_inverse_homogenization(
        ihd, batch, res=r, objective=objective, init_mode="rho", volfrac=0.25, penal=3.0, rho_min=1e-9, move=0.2,
        xmin=1e-3, iters=iters, seed=0, symmetry=symmetry, filter_k=3, output_dir=str(tmp_path), verbose=False)
    w = maple_amd.parse_objective_spec(objective)
    x_init = np.clip(((rho0 - 1e-9) / (1 - 1e-9)).clip(0, 1) ** (1 / 3), 1e-3, 1)
    xo, rows = odrv.run_driver(r, x_init, w, iters, 0.25, symmetry=symmetry, rtol=1e-11)
    with open(res.log_csv) as f:
        log = list(csv.reader(f))
    assert log[0] == ["iter", "J_total", "J_raw", "P_print", "P_cavity", "mean_x", "mean_rho", "dx_max", "dx_mean",
                      "time_sec"]
    assert len(log) == 1 + len(rows) == 1 + res.iters_run
    for got, want in zip(log[1:], rows):
        assert int(got[0]) == want[0]
        assert abs(float(got[1]) - want[1]) <= 1e-8 * abs(want[1])
        assert abs(float(got[5]) - want[2]) <= 1e-9 and abs(float(got[7]) - want[4]) <= 1e-7
    x_final = osym.symmetrize(xo, r, *osym.parse_symmetry(symmetry))
    rho_ref = (1e-9 + (1 - 1e-9) * np.clip(x_final, 1e-3, 1) ** 3).reshape(r, r, r)
    rf = np.load(res.rho_final_npy)
    assert rf.dtype == np.float32 and rf.shape == (r, r, r)
    assert np.abs(rf - rho_ref).max() <= 1e-6
    summary = json.load(open(res.summary_json))
    for k in ("objective", "symmetry", "resolution", "iters_run", "converged", "J_raw_init", "J_raw_final",
              "J_total_final", "volfrac_target", "rho_final_npy", "time_sec"):
        assert k in summary


def test_driver_early_stop(tmp_path):
    """A uniform design at the target volume fraction is a fixed point of OC: the loop
    stalls from iteration min_iters on and stops after `patience` stalls."""
    import maple_amd
    from maple_amd.driver import optimize_inverse_homogenization
    r = 16
    ihd = maple_amd.IHD(maple_amd.make_cfg(r), DEV, rtol=1e-8, nlevels=3)
    rho0 = torch.full((r, r, r), 1e-9 + (1 - 1e-9) * 0.3 ** 3, dtype=torch.float64, device=DEV)
    batch = maple_amd.build_unitcell_batch(DEV, r, rho0)
    res = optimize_inverse_homogenization(
        ihd, batch, res=r, objective="bulk", init_mode="rho", volfrac=0.3, penal=3.0, rho_min=1e-9, move=0.2,
        xmin=1e-3, iters=50, seed=0, symmetry="cubic", filter_k=3, output_dir=str(tmp_path), verbose=False)
    assert res.converged and res.iters_run == 20


@pytest.mark.parametrize("tiles", [1, 2])
def test_driver_with_penalties_matches_oracle(tmp_path, tiles):
    """Printability and cavity penalties at lambda = 0.9 (the experiment scripts'
    weights, experiments/biobjective/run_weight_sweep.sh:16-20) composed into the
    trajectory: J_total, J_raw, P_print and P_cavity of every iteration against the
    oracle's loop with its own penalty restatement."""
    import maple_amd
    from maple_amd.driver import optimize_inverse_homogenization
    r, iters = 16, 5
    print(f"[driver penalties] tiles={tiles}", flush=True)
    x0 = maple_amd.x0_sphere_void(r, seed=4)
    rho0 = 1e-9 + (1 - 1e-9) * x0 ** 3
    ihd = maple_amd.IHD(maple_amd.make_cfg(r), DEV, rtol=1e-10, maxit=4000, nlevels=3)
    batch = maple_amd.build_unitcell_batch(DEV, r, torch.from_numpy(rho0.reshape(r, r, r)).to(DEV))
    res = optimize_inverse_homogenization(
        ihd, batch, res=r, objective="stiffness:0.5,thermal:0.5", init_mode="rho", volfrac=0.25, penal=3.0,
        rho_min=1e-9, move=0.2, xmin=1e-3, iters=iters, seed=0, symmetry="none", lambda_print=0.9,
        overhang_deg=45.0, lambda_cavity=0.9, min_channel_rad=1, cavity_iters=24, cavity_tiles=tiles,
        filter_k=3, output_dir=str(tmp_path), verbose=False)
    x_init = np.clip(((rho0 - 1e-9) / (1 - 1e-9)).clip(0, 1) ** (1 / 3), 1e-3, 1)
    cav = dict(min_channel_rad=1, iters=24)
    if tiles > 1:
        cav["tiles_xyz"] = (tiles,) * 3
    xo, rows = odrv.run_driver(r, x_init, {"stiffness": 0.5, "thermal": 0.5}, iters, 0.25, rtol=1e-11,
                               lambda_print=0.9, lambda_cavity=0.9, cavity=cav)
    with open(res.log_csv) as f:
        log = list(csv.reader(f))[1:]
    assert len(log) == len(rows) == iters
    for got, want in zip(log, rows):
        assert abs(float(got[1]) - want[1]) <= 1e-8 * abs(want[1])        # J_total
        assert abs(float(got[2]) - want[7]) <= 1e-8 * abs(want[7])        # J_raw
        assert abs(float(got[3]) - want[8]) <= 1e-9 * max(want[8], 1e-12)  # P_print
        assert abs(float(got[4]) - want[9]) <= 1e-9 * max(want[9], 1e-12)  # P_cavity
        assert abs(float(got[7]) - want[4]) <= 1e-7                        # dx_max
    assert float(log[0][3]) > 0
