"""GPU parity of the mixed-precision form (FP32 multigrid preconditioner inside
FP64 CG) against the FP64 oracle.  Bar: C_H and kappa_H within
1e-5 Frobenius-relative at rtol 1e-6; sensitivities within 1e-5 of the largest."""
import os

import numpy as np
import pytest
import torch

from oracle import homog as ohomog, objective as oobj, oc as ooc

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0) if torch.cuda.is_available() else None


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(DEV)


@pytest.mark.parametrize("r,nlev,seed", [(8, 3, 0), (16, 3, 1), (16, 4, 2)])
def test_mixed_matches_oracle(r, nlev, seed):
    import maple_amd
    print(f"[mixed] r={r} nlev={nlev}", flush=True)
    rng = np.random.default_rng(seed)
    rho = 1e-9 + (1 - 1e-9) * rng.uniform(0, 1, r ** 3) ** 3
    cb = maple_amd.CellBatch([r], nlevels=nlev, device=DEV, precision="mixed")
    CH, kap, it = cb.homogenize(_t(rho), rtol=1e-6, maxit=2000)
    cs = ohomog.CellSolver((r, r, r), nlevels=3, rtol=1e-12)
    CHo, S6, _ = cs.elastic(rho)
    ko, S3, _ = cs.thermal(rho)
    assert np.linalg.norm(CH[0].cpu().numpy() - CHo) <= 1e-5 * np.linalg.norm(CHo)
    assert np.linalg.norm(kap[0].cpu().numpy() - ko) <= 1e-5 * np.linalg.norm(ko)
    J, G6, G3 = oobj.evaluate({"stiffness": 0.4, "shear": 0.3, "thermal": 0.3}, CHo, ko)
    d = cb.sensitivity(_t(G6).reshape(1, 6, 6), _t(G3).reshape(1, 3, 3)).cpu().numpy()
    do = np.einsum("ij,eij->e", G6, S6) + np.einsum("ij,eij->e", G3, S3)
    assert np.abs(d - do).max() <= 1e-5 * np.abs(do).max()
    # the FP64 outer CG still reaches a tight tolerance through the FP32 preconditioner
    CHt, kt, _ = cb.homogenize(_t(rho), rtol=1e-10, maxit=4000)
    assert np.linalg.norm(CHt[0].cpu().numpy() - CHo) <= 1e-9 * np.linalg.norm(CHo)


def test_mixed_golden_32(golden_dir):
    import maple_amd
    g = np.load(os.path.join(golden_dir, "sphere_void_32.npz"))
    cb = maple_amd.CellBatch([32], nlevels=3, device=DEV, precision="mixed")
    CH, kap, it = cb.homogenize(_t(ooc.simp(g["x0"])), rtol=1e-6)
    assert np.linalg.norm(CH[0].cpu().numpy() - g["CH"]) <= 1e-5 * np.linalg.norm(g["CH"])
    assert np.linalg.norm(kap[0].cpu().numpy() - g["kappa"]) <= 1e-5 * np.linalg.norm(g["kappa"])
    d = cb.sensitivity(_t(g["G6"]).reshape(1, 6, 6), _t(g["G3"]).reshape(1, 3, 3)).cpu().numpy()
    assert np.abs(d - g["dJdrho"]).max() <= 1e-5 * np.abs(g["dJdrho"]).max()


def test_mixed_vs_fp64_along_oc_trajectory():
    """Config-3 style: the mixed run's own rho at several OC iterates, checked against
    the FP64 path (itself pinned to the oracle) solved tightly on the same rho."""
    import maple_amd
    r = 32
    cm = maple_amd.CellBatch([r], nlevels=3, device=DEV, precision="mixed")
    cf = maple_amd.CellBatch([r], nlevels=3, device=DEV)
    drv = maple_amd.OCDriver(cm, "stiffness:0.4,shear:0.3,thermal:0.3", 0.25, move=0.2, rtol=1e-6, maxit=3000)
    x = _t(maple_amd.x0_uniform_noise(r, seed=3))
    worst = 0.0
    for it in range(30):
        x, info = drv.step(x)
        if it % 5 == 4:
            rho = info["rho"]
            CHf, kf, _ = cf.homogenize(rho, rtol=1e-10, maxit=6000)
            e = np.linalg.norm(info["CH"][0] - CHf[0].cpu().numpy()) / np.linalg.norm(CHf[0].cpu().numpy())
            ek = np.linalg.norm(info["kappa"][0] - kf[0].cpu().numpy()) / np.linalg.norm(kf[0].cpu().numpy())
            worst = max(worst, e, ek)
            print(f"OC it {it}: C_H rel err {e:.2e}  kappa rel err {ek:.2e}", flush=True)
    assert worst <= 1e-5


@pytest.mark.parametrize("fallback", [10, 100])
def test_mixed_islands_fall_back_to_fp64(fallback):
    """Solid islands floating in 1e-9 void have soft modes that an FP32 V-cycle cannot
    resolve; the mixed form finishes such solves on the FP64 hierarchy and still
    matches the FP64 path (itself pinned to the oracle) solved tightly."""
    import maple_amd
    r = 16
    print(f"[mixed islands] r={r} fallback={fallback}", flush=True)
    x = np.full((r, r, r), 1e-3)
    x[0, :, :] = 1.0          # a periodic wall keeps the solid connected
    x[:, 0, :] = 1.0
    rng = np.random.default_rng(11)
    for _ in range(12):       # isolated voxels and small blobs away from the walls
        i, j, k = rng.integers(3, r - 3, size=3)
        x[i:i + rng.integers(1, 3), j:j + rng.integers(1, 3), k] = 1.0
    rho = 1e-9 + (1 - 1e-9) * x.ravel() ** 3
    cm = maple_amd.CellBatch([r], nlevels=3, device=DEV, precision="mixed")
    cm.set_mixed_fallback(fallback)
    cf = maple_amd.CellBatch([r], nlevels=3, device=DEV)
    CH, kap, it = cm.homogenize(_t(rho), rtol=1e-6, maxit=3000)
    CHf, kf, itf = cf.homogenize(_t(rho), rtol=1e-10, maxit=6000)
    print(f"mixed iters {it.tolist()}  fp64 iters {itf.tolist()}", flush=True)
    assert np.linalg.norm(CH[0].cpu().numpy() - CHf[0].cpu().numpy()) <= 1e-5 * np.linalg.norm(CHf[0].cpu().numpy())
    assert np.linalg.norm(kap[0].cpu().numpy() - kf[0].cpu().numpy()) <= 1e-5 * np.linalg.norm(kf[0].cpu().numpy())
    with pytest.raises(ValueError):
        cm.set_mixed_fallback(-1)
