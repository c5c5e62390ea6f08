"""GPU parity of block CG (the load cases of a cell share one Krylov space)
against the FP64 oracle, at the FP64 form's bar: C_H / kappa_H Frobenius-relative
1e-9 with both sides at rtol 1e-10; closed forms of a homogeneous cell and a
1:1e9 laminate (whose in-plane shear right-hand sides vanish, so block
directions are dropped); batches of uneven cells equal single cells; the mixed
form within 1e-5."""
import numpy as np
import pytest
import torch

from oracle import fe as ofe, homog as ohomog

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0) if torch.cuda.is_available() else None


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(DEV)


def _rho(r, seed, contrast=1e-9):
    x = np.random.default_rng(seed).uniform(0.0, 1.0, r ** 3)
    return contrast + (1 - contrast) * x ** 3


def _cells(dims, nlev, precision="fp64"):
    import maple_amd
    return maple_amd.CellBatch(dims, nlevels=nlev, device=DEV, precision=precision, block_cg=True)


@pytest.mark.parametrize("r,nlev,seed", [(8, 3, 1), (16, 3, 2), (32, 3, 5)])
def test_block_cg_matches_oracle(r, nlev, seed):
    """Against the oracle at 8^3 and 16^3; at 32^3 (where the oracle's plain-Jacobi
    multigrid needs thousands of iterations on per-voxel 1:1e9 noise) against the
    per-case GPU solve, which the other parity tests pin to the oracle."""
    print(f"[block cg] r={r} nlev={nlev}", flush=True)
    import maple_amd
    rho = _rho(r, seed)
    cb = _cells([r], nlev)
    CH, kap, it = cb.homogenize(_t(rho), rtol=1e-10, maxit=3000)
    ref = maple_amd.CellBatch([r], nlevels=nlev, device=DEV)
    CH1, k1, it1 = ref.homogenize(_t(rho), rtol=1e-10, maxit=3000)
    if r <= 16:
        cs = ohomog.CellSolver((r, r, r), nlevels=3, rtol=1e-12)
        CHo, _, _ = cs.elastic(rho)
        ko, _, _ = cs.thermal(rho)
    else:
        CHo, ko = CH1[0].cpu().numpy(), k1[0].cpu().numpy()
    e1 = np.linalg.norm(CH[0].cpu().numpy() - CHo) / np.linalg.norm(CHo)
    e2 = np.linalg.norm(kap[0].cpu().numpy() - ko) / np.linalg.norm(ko)
    print(f"C_H rel {e1:.2e} kappa rel {e2:.2e}; block iterations {it[0].tolist()} vs per-case {it1[0].tolist()}",
          flush=True)
    assert e1 <= 1e-9 and e2 <= 1e-9
    # every case of a cell reports the block's count, and sharing the space pays
    assert len(set(it[0, :6].tolist())) == 1 and len(set(it[0, 6:].tolist())) == 1
    assert it[0, :6].max() <= it1[0, :6].max()


def test_block_cg_closed_forms():
    r = 16
    cb = _cells([r], 3)
    C0 = ofe.isotropic_elastic_tensor(1, 0.3)
    CH, kap, it = cb.homogenize(_t(np.full(r ** 3, 0.7)), rtol=1e-10)
    assert np.abs(CH[0].cpu().numpy() - 0.7 * C0).max() < 1e-14
    assert np.abs(kap[0].cpu().numpy() - 0.7 * np.eye(3)).max() < 1e-14
    prof = np.random.default_rng(3).uniform(0.2, 1.0, r)
    prof[::4] = 1e-9
    CH, kap, it = cb.homogenize(_t(np.tile(prof, r * r)), rtol=1e-12, maxit=3000)
    lam, mu = C0[0, 1], C0[3, 3]
    a = lam + 2 * mu
    h, m = 1 / np.mean(1 / prof), np.mean(prof)
    ref = np.zeros((6, 6))
    ref[0, 0] = ref[1, 1] = 4 * mu * (lam + mu) / a * m + lam ** 2 / a * h
    ref[0, 1] = ref[1, 0] = 2 * mu * lam / a * m + lam ** 2 / a * h
    ref[2, 2] = a * h
    ref[0, 2] = ref[2, 0] = ref[1, 2] = ref[2, 1] = lam * h
    ref[3, 3] = mu * m
    ref[4, 4] = ref[5, 5] = mu * h
    assert np.linalg.norm(CH[0].cpu().numpy() - ref) <= 1e-11 * np.linalg.norm(ref)
    kref = np.diag([m, m, h])
    assert np.linalg.norm(kap[0].cpu().numpy() - kref) <= 1e-11 * np.linalg.norm(kref)


def test_block_cg_batch_equals_single():
    dims = [8, 16, 8]
    rhos = [_rho(d, 20 + i) for i, d in enumerate(dims)]
    cb = _cells(dims, 3)
    CH, kap, it = cb.homogenize(_t(np.concatenate(rhos)), rtol=1e-10, maxit=3000)
    for i, d in enumerate(dims):
        c1, k1, it1 = _cells([d], 3).homogenize(_t(rhos[i]), rtol=1e-10, maxit=3000)
        np.testing.assert_array_equal(it[i], it1[0])
        assert torch.linalg.norm(CH[i] - c1[0]) <= 1e-12 * torch.linalg.norm(c1[0])
        assert torch.linalg.norm(kap[i] - k1[0]) <= 1e-12 * torch.linalg.norm(k1[0])


def test_block_cg_mixed_and_warm_start():
    import maple_amd
    r = 32
    print(f"[block cg mixed] r={r}", flush=True)
    cm = _cells([r], 3, precision="mixed")
    cf = maple_amd.CellBatch([r], nlevels=3, device=DEV)
    drv = maple_amd.OCDriver(cm, "stiffness:0.4,shear:0.3,thermal:0.3", 0.25, move=0.2, rtol=1e-6, maxit=3000)
    cm.set_solver(warm_start=True)
    x = _t(maple_amd.x0_uniform_noise(r, seed=3))
    worst = 0.0
    for it in range(20):
        x, info = drv.step(x)
        if it % 5 == 4:
            CHf, kf, _ = cf.homogenize(info["rho"], rtol=1e-10, maxit=6000)
            e = np.linalg.norm(info["CH"][0] - CHf[0].cpu().numpy()) / np.linalg.norm(CHf[0].cpu().numpy())
            ek = np.linalg.norm(info["kappa"][0] - kf[0].cpu().numpy()) / np.linalg.norm(kf[0].cpu().numpy())
            worst = max(worst, e, ek)
            print(f"OC it {it}: iters {info['iters'][0].tolist()} C_H rel err {e:.2e} kappa rel err {ek:.2e}",
                  flush=True)
    assert worst <= 1e-5
