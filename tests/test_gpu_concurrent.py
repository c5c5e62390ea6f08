"""Concurrent groups of cells (maple_amd.concurrent.ConcurrentOC): every cell follows
the trajectory it has in one lockstep batch, whatever group it runs in; errors on a
group's thread reach the caller."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0) if torch.cuda.is_available() else None


def _make(ncells, r=16):
    import maple_amd
    cb = maple_amd.CellBatch([r] * ncells, nlevels=3, device=DEV, precision="mixed", warm_start=True, islands=True)
    drv = maple_amd.OCDriver(cb, "stiffness:0.5,thermal:0.5", 0.25, rtol=1e-8, maxit=3000)
    return cb, drv


def test_groups_follow_the_lockstep_trajectories():
    from maple_amd.concurrent import ConcurrentOC
    r, n, steps = 16, 4, 3
    x0 = np.concatenate([np.clip(0.25 + np.random.default_rng(k).uniform(-0.1, 0.1, r ** 3), 1e-3, 1)
                         for k in range(n)])
    cb, drv = _make(n, r)
    x = torch.from_numpy(x0).to(DEV)
    for _ in range(steps):
        x, info = drv.step(x)
    ref = x.cpu().numpy()
    cb.close()
    groups = []
    for g in range(2):
        c, d = _make(2, r)
        groups.append((c, d, torch.from_numpy(x0[g * 2 * r ** 3:(g + 1) * 2 * r ** 3].copy()).to(DEV)))
    seen = []
    res = ConcurrentOC(groups).run(steps, on_step=lambda g, k, info: seen.append((g, k)))
    assert sorted(seen) == [(g, k) for g in range(2) for k in range(steps)]
    got = np.concatenate([res[g][0].cpu().numpy() for g in range(2)])
    assert np.abs(got - ref).max() <= 1e-9
    for c, _, _ in groups:
        c.close()


def test_group_errors_reach_the_caller():
    from maple_amd.concurrent import ConcurrentOC
    import maple_amd
    c, d = _make(1)
    bad = torch.zeros(7, dtype=torch.float64, device=DEV)   # wrong size: the step raises
    with pytest.raises(ValueError):
        ConcurrentOC([(c, d, bad)]).run(1)
    c.close()


def test_group_streams_are_reserved_once_and_reused():
    """The first CellBatch of the device reserves the process's concurrent-work streams
    before its context makes its own; later callers (groups, repeated runs) get those
    same streams back instead of fresh ones."""
    from maple_amd import streams
    from maple_amd.concurrent import ConcurrentOC
    cb, drv = _make(1)
    first = streams.group_streams(DEV, streams.RESERVED)
    assert len(first) == streams.RESERVED
    assert len({s.cuda_stream for s in first}) == streams.RESERVED       # distinct streams
    again = streams.group_streams(DEV, 3)
    assert [s.cuda_stream for s in again] == [s.cuda_stream for s in first[:3]]
    x = torch.full((cb.nelem,), 0.25, dtype=torch.float64, device=DEV)
    co1 = ConcurrentOC([(cb, drv, x.clone())])
    co2 = ConcurrentOC([(cb, drv, x.clone())])
    assert co1.streams[0].cuda_stream == co2.streams[0].cuda_stream == first[0].cuda_stream
    more = streams.group_streams(DEV, streams.RESERVED + 2)              # grows on demand, keeps the first ones
    assert [s.cuda_stream for s in more[:streams.RESERVED]] == [s.cuda_stream for s in first]
    cb.close()


def test_headline_composition_small():
    """The headline's composition on a small workload: cells in concurrent groups, mixed
    precision, island correction, Chebyshev-weighted 3+3 sweeps, coarsest inverse every third
    density, the learned start (delta mode) inside a 4-solution projected start.  After a few
    OC steps every cell's C_H and kappa_H match a tight FP64 solve of its density to 1e-5,
    and the solves started from projections."""
    import maple_amd
    from maple_amd import learned as ml
    from maple_amd.concurrent import ConcurrentOC
    r, per, groups, steps = 32, 2, 2, 5
    sets = []
    for g in range(groups):
        cb = maple_amd.CellBatch([r] * per, nlevels=3, device=DEV, precision="mixed", warm_start=True, islands=True)
        cb.set_coarse_refresh(3)
        cb.set_sweeps(3, 3)
        cb.set_chebyshev(0.1)
        cb.set_start_history(4)
        start = ml.LearnedStart(cb, ml.UNet.load(max_batch=per, device=DEV), mode="delta")
        drv = maple_amd.OCDriver(cb, "stiffness:0.5,thermal:0.5", 0.25, move=0.2, xmin=1e-3, filter_k=3, rtol=1e-6,
                                 maxit=3000, start=start)
        x0 = np.concatenate([maple_amd.x0_uniform_noise(r, seed=g * per + k) for k in range(per)])
        sets.append((cb, drv, torch.from_numpy(x0).to(DEV), start))
    res = ConcurrentOC([(cb, drv, x) for cb, drv, x, _ in sets]).run(steps)
    for (cb, _, _, start), (x, info, _) in zip(sets, res):
        assert (cb.start_sources("elastic") == 3).all() and (cb.start_sources("thermal") == 3).all()
        ref = maple_amd.CellBatch([r] * per, nlevels=3, device=DEV)
        CHt, kt, _ = ref.homogenize(info["rho"], rtol=1e-10, maxit=4000)
        for c in range(per):
            CH, kap = torch.from_numpy(info["CH"][c]).to(DEV), torch.from_numpy(info["kappa"][c]).to(DEV)
            assert float(torch.linalg.norm(CH - CHt[c]) / torch.linalg.norm(CHt[c])) <= 1e-5
            assert float(torch.linalg.norm(kap - kt[c]) / torch.linalg.norm(kt[c])) <= 1e-5
        ref.close()
        start.net.close()
        cb.close()
