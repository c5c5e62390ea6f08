"""Solves started on the projection over earlier solutions (maple_set_start_history).

The projected start is the combination of the basis (the last m solutions, newest
first, and a supplied start) that minimises the energy norm of the error.  When the
exact solution u* lies in the span of the basis, the start is u* itself and the solve
needs no iteration, whatever the individual vectors are: each check builds a basis
from u* and random fields w whose combination is u* (the coefficients differ per
check, so a transposed Gram entry, a wrong right-hand side or a misplaced
coefficient leaves a residual), including a duplicated vector that the factorisation
must drop instead of dividing by a zero pivot.  u* itself comes from an independent
tight solve, and is checked against the oracle's direct solve.  A trajectory with the
history on must agree with one without (same designs and C_H to the solver tolerance)
and take fewer CG iterations in total.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0) if torch.cuda.is_available() else None
NC = {"elastic": 6, "thermal": 3}
ND = {"elastic": 3, "thermal": 1}


def _rho(r, seed):
    rng = np.random.default_rng(seed)
    x = np.clip(0.45 + 0.5 * rng.standard_normal(r ** 3), 0.05, 1.0)
    return x ** 3


def _solved(r, physics, rho, m):
    import maple_amd
    cells = maple_amd.CellBatch([r], nlevels=3, device=DEV)
    cells.set_start_history(m)
    rt = torch.from_numpy(rho).to(DEV)
    cells.homogenize(rt, (physics,), rtol=1e-12, maxit=4000)
    return cells, rt, cells.fields(physics)


def test_tight_solution_matches_oracle():
    from oracle import fe, grid, mg
    r = 8
    rho = _rho(r, 3)
    cells, rt, u = _solved(r, "elastic", rho, 2)
    Ke, Fe, _ = fe.hexahedron(r, fe.isotropic_elastic_tensor())
    F = grid.assemble_rhs(rho, Fe, (r, r, r))
    ref = mg.direct_solve(rho, Ke, (r, r, r), F)                      # (N, 3, 6)
    got = u.cpu().numpy().transpose(2, 1, 0)                          # (6, 3, N) -> (N, 3, 6)
    got = got - got.mean(axis=0, keepdims=True)                       # both sides up to translations
    assert np.linalg.norm(got - ref) <= 1e-9 * np.linalg.norm(ref)
    cells.close()


@pytest.mark.parametrize("physics", ["elastic", "thermal"])
@pytest.mark.parametrize("r", [8, 16])
def test_projection_recovers_solution_in_span(physics, r):
    print(f"[projection] {physics} r={r}", flush=True)
    cells, rt, ustar = _solved(r, physics, _rho(r, 7 + r), 2)
    g = torch.Generator(device=DEV).manual_seed(11)
    scale = torch.linalg.norm(ustar) / ustar.numel() ** 0.5
    w1 = torch.randn(ustar.shape, generator=g, device=DEV, dtype=torch.float64) * scale
    w2 = torch.randn(ustar.shape, generator=g, device=DEV, dtype=torch.float64) * scale
    # u* = 0.2 v1 + 0.3 v2 + 0.5 s: the ring holds v1, v2; s is the supplied start
    v1 = ustar + w1
    v2 = ustar + 2.0 * w2
    s = ustar - 0.4 * w1 - 1.2 * w2
    cells.set_fields(physics, v1)
    cells.homogenize(rt, (physics,), rtol=1e-9, maxit=1, strict=False)   # pushes v1
    cells.set_fields(physics, v2)
    cells.set_start(physics, s)
    _, _, it = cells.homogenize(rt, (physics,), rtol=1e-9, maxit=1, strict=False)
    n = NC[physics]
    cols = slice(0, 6) if physics == "elastic" else slice(6, 9)
    assert (it[0, cols] == 0).all(), it
    assert (cells.start_sources(physics) == 3).all()
    u = cells.fields(physics)
    assert float(torch.linalg.norm(u - ustar) / torch.linalg.norm(ustar)) <= 1e-7
    # a dependent basis: ring (v2', v1') with the supplied start equal to v1'
    a1 = ustar + w1
    a2 = ustar - w1                                               # u* = (a1 + a2) / 2
    cells.set_fields(physics, a1)
    cells.homogenize(rt, (physics,), rtol=1e-9, maxit=1, strict=False)
    cells.set_fields(physics, a2)
    cells.set_start(physics, a1.clone())
    _, _, it = cells.homogenize(rt, (physics,), rtol=1e-9, maxit=1, strict=False)
    assert (it[0, cols] == 0).all(), it
    u = cells.fields(physics)
    assert torch.isfinite(u).all()
    assert float(torch.linalg.norm(u - ustar) / torch.linalg.norm(ustar)) <= 1e-7
    assert u.shape == (n, ND[physics], r ** 3)
    # a nearly dependent basis: the supplied start differs from a ring vector by 1e-13 of
    # its size, far below what the Gram matrix resolves; it must be left out, not amplified
    b1 = ustar + w2
    b2 = ustar - w2
    cells.set_fields(physics, b1)
    cells.homogenize(rt, (physics,), rtol=1e-9, maxit=1, strict=False)
    cells.set_fields(physics, b2)
    cells.set_start(physics, b1 * (1.0 + 1e-13 * torch.randn(b1.shape, generator=g, device=DEV, dtype=torch.float64)))
    _, _, it = cells.homogenize(rt, (physics,), rtol=1e-9, maxit=1, strict=False)
    assert (it[0, cols] == 0).all(), it
    u = cells.fields(physics)
    assert torch.isfinite(u).all()
    assert float(torch.linalg.norm(u - ustar) / torch.linalg.norm(ustar)) <= 1e-7
    cells.close()


def test_history_cleared_and_bounds():
    import maple_amd
    cells = maple_amd.CellBatch([8], nlevels=2, device=DEV)
    with pytest.raises(Exception):
        cells.set_start_history(6)
    with pytest.raises(Exception):
        cells.set_start_history(-1)
    rho = torch.from_numpy(_rho(8, 1)).to(DEV)
    cells.set_start_history(3)
    cells.homogenize(rho, ("elastic", "thermal"), rtol=1e-8)
    assert (cells.start_sources("elastic") == 0).all()              # nothing to project on yet
    cells.homogenize(rho, ("elastic", "thermal"), rtol=1e-8)
    assert (cells.start_sources("elastic") == 3).all()
    assert (cells.last_iters == 0).all()                            # the same density: the previous solution
    cells.set_start_history(0)
    cells.homogenize(rho, ("elastic", "thermal"), rtol=1e-8)
    assert (cells.start_sources("elastic") == 0).all()              # off, and no warm start
    cells.close()


def _trajectory(m, dims, precision, steps):
    import maple_amd
    cells = maple_amd.CellBatch(dims, nlevels=3, device=DEV, precision=precision, warm_start=True,
                                islands=precision == "mixed")
    cells.set_start_history(m)
    rng = np.random.default_rng(5)
    x = torch.from_numpy(np.concatenate([np.clip(0.3 + rng.uniform(-0.2, 0.2, r ** 3), 1e-3, 1.0)
                                         for r in dims])).to(DEV)
    drv = maple_amd.OCDriver(cells, "stiffness:0.5,thermal:0.5", 0.3, move=0.2, xmin=1e-3, filter_k=3, rtol=1e-9,
                             maxit=3000)
    out, total = [], 0
    for _ in range(steps):
        x, info = drv.step(x)
        out.append((info["CH"].copy(), info["kappa"].copy(), x.clone()))
        total += int(np.asarray(info["iters"]).sum())
    cells.close()
    return out, total


@pytest.mark.parametrize("case", [dict(dims=[16, 32], precision="fp64"), dict(dims=[32, 16], precision="mixed")])
def test_trajectory_with_history(case):
    print(f"[projection trajectory] {case}", flush=True)
    base, it0 = _trajectory(0, steps=12, **case)
    proj, it3 = _trajectory(3, steps=12, **case)
    for k, (a, b) in enumerate(zip(base, proj)):
        for c in range(len(case["dims"])):
            assert np.linalg.norm(a[0][c] - b[0][c]) <= 1e-6 * np.linalg.norm(a[0][c]), f"step {k} cell {c}"
            assert np.linalg.norm(a[1][c] - b[1][c]) <= 1e-6 * np.linalg.norm(a[1][c]), f"step {k} cell {c}"
        assert float((a[2] - b[2]).abs().max()) <= 1e-5, f"step {k}: designs differ"
    assert it3 < it0, (it3, it0)
    again, it3b = _trajectory(3, steps=12, **case)        # fixed-order reductions: bit for bit
    assert it3b == it3
    for k, (a, b) in enumerate(zip(proj, again)):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and torch.equal(a[2], b[2]), k


def test_ihd_dropin_with_history():
    """The reference-shaped IHD with start_history: a sequence of nearby designs gives the
    C_H and kappa_H of the drop-in without it (both at rtol 1e-10), and from the second
    call on the solves start projected on the earlier solutions."""
    import maple_amd
    r = 16
    cfg = maple_amd.make_cfg(r, smooth_iter=[2, 2, 3])
    plain = maple_amd.IHD(cfg, DEV, rtol=1e-10, maxit=4000)
    hist = maple_amd.IHD(cfg, DEV, rtol=1e-10, maxit=4000, start_history=3)
    rng = np.random.default_rng(8)
    x = np.clip(0.3 + rng.uniform(-0.2, 0.2, r ** 3), 1e-3, 1.0)
    it_plain = it_hist = 0
    for k in range(4):
        x = np.clip(x + 0.02 * rng.standard_normal(r ** 3), 1e-3, 1.0)
        out = []
        for ihd in (plain, hist):
            batch = maple_amd.build_unitcell_batch(DEV, r, torch.from_numpy(x).reshape(r, r, r).float())
            rho = torch.from_numpy(1e-9 + (1 - 1e-9) * x ** 3).reshape(-1, 1).to(DEV)
            batch["rho"] = rho
            n = batch["feature"].shape[0]
            CH = ihd.TO_homo_3d(batch, ihd.Ke_3d, maple_amd.assemble_F_fast(batch["node_index"], ihd.Fe_3d, n, rho),
                                ihd.X0_3d, batch_size=1)
            its = int(ihd.cells.last_iters[0, :6].sum())
            kap = ihd.TO_homo_1d(batch, ihd.Ke_1d,
                                 maple_amd.assemble_F_fast_scalar(batch["node_index"], ihd.Fe_1d, n, rho),
                                 ihd.X0_1d, batch_size=1)
            out.append((CH.detach(), kap.detach(), its))
        (CHp, kp, ip), (CHh, kh, ih) = out
        assert float(torch.linalg.norm(CHh - CHp) / torch.linalg.norm(CHp)) <= 1e-8, k
        assert float(torch.linalg.norm(kh - kp) / torch.linalg.norm(kp)) <= 1e-8, k
        if k:
            assert (hist.cells.start_sources("elastic") == 3).all()
            it_plain, it_hist = it_plain + ip, it_hist + ih
    assert it_hist < it_plain, (it_hist, it_plain)
