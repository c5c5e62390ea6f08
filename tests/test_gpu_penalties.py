"""The AM penalties on the GPU against the oracle (oracle/printability.py).

Values within 1e-12 relative and gradients within 1e-11 of the largest entry, on
continuous designs and on quantized designs full of ties (the tie rules of the
adjoints must agree entry by entry), for one cell and for a batch of cells of
different resolutions.
"""
import numpy as np
import pytest
import torch

from oracle import printability as op

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0) if torch.cuda.is_available() else None


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(DEV)


def _designs(r, seed):
    rng = np.random.default_rng(seed)
    cont = rng.uniform(0.0, 1.0, r ** 3)
    quant = rng.choice([1e-9, 0.3, 1.0], size=r ** 3)
    blobs = np.full((r, r, r), 1e-9)
    for _ in range(6):
        a = rng.integers(0, r - 3, 3)
        w = rng.integers(2, 4, 3)
        blobs[a[0]:a[0] + w[0], a[1]:a[1] + w[1], a[2]:a[2] + w[2]] = 1.0
    return [cont, quant, blobs.ravel()]


def _check(P, dP, Po, go, what):
    assert abs(P - Po) <= 1e-12 * max(abs(Po), 1e-300) + 1e-15, f"{what}: P {P} vs {Po}"
    assert np.abs(dP - go).max() <= 1e-11 * max(np.abs(go).max(), 1e-300), f"{what}: gradient"


@pytest.mark.parametrize("deg,periodic", [(45.0, True), (45.0, False), (30.0, True)])
@pytest.mark.parametrize("r", [8, 32])
def test_overhang(deg, periodic, r):
    import maple_amd
    print(f"[overhang] r={r} deg={deg} periodic={periodic}", flush=True)
    cb = maple_amd.CellBatch([r], nlevels=2, device=DEV)
    for k, rho in enumerate(_designs(r, r)):
        P, dP = cb.penalty_overhang(_t(rho), deg, periodic)
        Po, go = op.overhang(rho, r, deg, periodic)
        _check(float(P[0]), dP.cpu().numpy(), Po, go, f"design {k}")
    cb.close()


CAVITY = [
    dict(min_channel_rad=0, rad=1, periodic_xyz=True, seed_faces="all"),
    dict(min_channel_rad=1, rad=1, periodic_xyz=True, seed_faces="z"),
    dict(min_channel_rad=3, rad=1, periodic_xyz=False, seed_faces="all"),
    dict(min_channel_rad=0, rad=0, periodic_xyz=True, seed_faces="x"),
    dict(min_channel_rad=2, rad=1, periodic_xyz=False, seed_faces="xy"),
]


@pytest.mark.parametrize("opts", CAVITY)
def test_closed_void(opts):
    import maple_amd
    r = 12
    print(f"[closed void] {opts}", flush=True)
    cb = maple_amd.CellBatch([r], nlevels=2, device=DEV)
    cb.set_cavity(iters=20, **opts)
    for k, rho in enumerate(_designs(r, 5)):
        P, dP = cb.penalty_cavity(_t(rho))
        Po, go = op.closed_void(rho, r, 0.5, 20.0, opts["min_channel_rad"], opts["rad"], 20, opts["periodic_xyz"],
                                opts["seed_faces"])
        _check(float(P[0]), dP.cpu().numpy(), Po, go, f"design {k}")
    cb.close()


@pytest.mark.parametrize("tiles,focus,mcr", [(2, True, 3), (3, True, 1), (3, False, 0), ((2, 1, 3), True, 1)])
def test_periodic_array(tiles, focus, mcr):
    import maple_amd
    r = 8
    txyz = tiles if isinstance(tiles, tuple) else (tiles,) * 3
    print(f"[periodic array] tiles={txyz} focus={focus} mcr={mcr}", flush=True)
    cb = maple_amd.CellBatch([r], nlevels=2, device=DEV)
    cb.set_cavity(iters=0, min_channel_rad=mcr, tiles=txyz, focus_center=focus)
    for k, rho in enumerate(_designs(r, 7)):
        P, dP = cb.penalty_cavity(_t(rho))
        Po, go = op.periodic_array(rho, r, tiles_xyz=txyz, focus_center=focus, min_channel_rad=mcr, iters=0)
        _check(float(P[0]), dP.cpu().numpy(), Po, go, f"design {k}")
    cb.close()


def test_batch_of_cells():
    """Cells of different resolutions in one context: each cell's penalties are its own."""
    import maple_amd
    sizes = [8, 16, 12]
    rhos = [_designs(r, 11 + i)[i % 3] for i, r in enumerate(sizes)]
    cb = maple_amd.CellBatch(sizes, nlevels=2, device=DEV)
    rho = _t(np.concatenate(rhos))
    P, dP = cb.penalty_overhang(rho)
    cb.set_cavity(iters=0, min_channel_rad=1, tiles=2)
    C, dC = cb.penalty_cavity(rho)
    off = np.concatenate([[0], np.cumsum([r ** 3 for r in sizes])])
    for i, r in enumerate(sizes):
        Po, go = op.overhang(rhos[i], r)
        _check(float(P[i]), dP[off[i]:off[i + 1]].cpu().numpy(), Po, go, f"overhang cell {i}")
        Co, gc = op.periodic_array(rhos[i], r, tiles=2, min_channel_rad=1, iters=0)
        _check(float(C[i]), dC[off[i]:off[i + 1]].cpu().numpy(), Co, gc, f"cavity cell {i}")
    cb.close()


def test_penalty_argument_errors():
    import maple_amd
    cb = maple_amd.CellBatch([8], nlevels=2, device=DEV)
    with pytest.raises(ValueError):
        cb.penalty_cavity(_t(np.ones(512)))          # not configured
    with pytest.raises(ValueError):
        cb.set_cavity(min_channel_rad=4)
    with pytest.raises(ValueError):
        cb.penalty_overhang(_t(np.ones(512)), overhang_deg=5.0)   # support radius 12
    cb.close()
