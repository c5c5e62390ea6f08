"""CG iterations replayed as hipGraphs give bitwise the results of single launches.

Each check runs the same OC trajectory twice from the same start, once with the CG
iterations as single launches and once as captured hipGraphs (maple_set_graphs), and
compares every step's C_H, kappa_H, dJ/drho, design and CG iteration counts bit for bit.
The graph run replays its captured segments many times per solve, with the host's
convergence polls, launch-table compaction and the OC step's eager work in between;
the cases cover one cell, a batch whose cells converge at different iterations (so the
launch tables are compacted and the segments re-captured), the mixed form with the
island correction, and the mixed form's switch to the FP64 V-cycle.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0) if torch.cuda.is_available() else None


def _trajectory(dims, graphs, steps, precision="fp64", islands=False, warm=False, fallback=None, seed=0):
    import maple_amd
    nlev = 3
    cells = maple_amd.CellBatch(dims, nlevels=nlev, device=DEV, precision=precision, islands=islands,
                                warm_start=warm)
    if fallback is not None:
        cells.set_mixed_fallback(fallback)
    cells.set_graphs(graphs)
    rng = np.random.default_rng(seed)
    x = torch.from_numpy(np.concatenate([np.clip(0.3 + rng.uniform(-0.2, 0.2, r ** 3), 1e-3, 1.0)
                                         for r in dims])).to(DEV)
    drv = maple_amd.OCDriver(cells, "stiffness:0.5,thermal:0.5", 0.3, move=0.2, xmin=1e-3, filter_k=3, rtol=1e-8,
                             maxit=2000)
    out = []
    for _ in range(steps):
        x, info = drv.step(x)
        torch.cuda.synchronize()
        out.append((info["CH"].copy(), info["kappa"].copy(), info["dJdrho"].clone(), x.clone(),
                    np.asarray(info["iters"]).copy()))
    counts = cells.graph_counts()
    cells.close()
    return out, counts


@pytest.mark.parametrize("case", [
    dict(dims=[16]),
    dict(dims=[16, 32, 16]),
    dict(dims=[32, 16], precision="mixed", islands=True, warm=True),
    dict(dims=[16, 16], precision="mixed", fallback=3),
])
def test_graph_replay_is_bitwise_eager(case):
    print(f"[graphs] {case}", flush=True)
    eager, ce = _trajectory(graphs=False, steps=4, **case)
    graph, cg = _trajectory(graphs=True, steps=4, **case)
    assert ce == (0, 0)
    # the graph run replayed its segments: far more launches than captures
    assert cg[0] >= 2 and cg[1] >= 4 * cg[0], cg
    for k, (a, b) in enumerate(zip(eager, graph)):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), f"step {k}: C_H / kappa differ"
        assert torch.equal(a[2], b[2]) and torch.equal(a[3], b[3]), f"step {k}: sensitivities / design differ"
        assert np.array_equal(a[4], b[4]), f"step {k}: CG iteration counts differ"
    # the trajectory did iterate: several CG iterations per solve
    assert min(int(np.asarray(e[4]).max()) for e in eager) >= 3
