"""Pins oracle/printability.py (the AM penalties and their adjoints).

* against torch autograd of the same compositions written here with torch's own
  pooling ops in float64 (torch's max-pool backward routes to the first maximum of
  a window, its minimum backward halves ties: the conventions the oracle restates),
  on continuous designs and on quantized designs full of ties;
* against closed forms: solid columns have no overhang, a slab floating at layer k
  costs 1/(r-1), a sealed void cube is trapped (sigmoid(beta t) - sigmoid(-beta t)
  per voxel), an open channel is not;
* against central finite differences on designs without ties.
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import printability as op


def _t_overhang(g, rad, periodic):
    r = g.shape[0]
    sup = g[:, :, 0]
    P = g.new_zeros(())
    for z in range(1, r):
        s = sup[None, None]
        s = F.pad(s, (rad,) * 4, mode="circular") if periodic else F.pad(s, (rad,) * 4, value=0.0)
        s = F.max_pool2d(s, 2 * rad + 1, stride=1)[0, 0]
        P = P + F.relu(g[:, :, z] - s).mean()
        sup = torch.minimum(g[:, :, z], s)
    return P / max(r - 1, 1)


def _t_pool(x, rad, periodic, val, minimum=False):
    t = x[None, None]
    t = F.pad(t, (rad,) * 6, mode="circular") if periodic else F.pad(t, (rad,) * 6, value=val)
    if minimum:
        return -F.max_pool3d(-t, 2 * rad + 1, stride=1)[0, 0]
    return F.max_pool3d(t, 2 * rad + 1, stride=1)[0, 0]


def _t_trapped(void, mcr, rad, nsteps, periodic, seed, region):
    vc = void if mcr == 0 else _t_pool(_t_pool(void, mcr, periodic, 1.0, True), mcr, periodic, 0.0)
    reach = vc * seed
    for _ in range(nsteps if rad > 0 else 0):
        reach = torch.minimum(vc, _t_pool(reach, rad, periodic, 0.0))
    return F.relu(void - reach)[region].mean()


def _grad(fn, rho):
    x = torch.tensor(rho, dtype=torch.float64, requires_grad=True)
    P = fn(x)
    (g,) = torch.autograd.grad(P, x)
    return float(P.detach()), g.numpy().ravel()


def _designs(r, seed):
    rng = np.random.default_rng(seed)
    cont = rng.uniform(0.0, 1.0, r ** 3)
    quant = rng.choice([1e-9, 0.3, 1.0], size=r ** 3)
    return [cont, quant]


@pytest.mark.parametrize("deg", [45.0, 30.0])
@pytest.mark.parametrize("periodic", [True, False])
def test_overhang_matches_autograd(deg, periodic):
    r = 8
    rad = op.support_radius(deg)
    for rho in _designs(r, 1):
        P, g = op.overhang(rho, r, deg, periodic)
        Pt, gt = _grad(lambda x: _t_overhang(x.reshape(r, r, r), rad, periodic), rho)
        assert abs(P - Pt) <= 1e-14 * max(1.0, abs(Pt))
        assert np.abs(g - gt).max() <= 1e-14


@pytest.mark.parametrize("mcr,rad,periodic,faces", [(0, 1, True, "all"), (1, 1, True, "z"), (2, 1, False, "all"),
                                                    (0, 0, True, "x"), (1, 1, False, "xy")])
def test_closed_void_matches_autograd(mcr, rad, periodic, faces):
    r = 8
    for rho in _designs(r, 2):
        P, g = op.closed_void(rho, r, 0.5, 20.0, mcr, rad, 12, periodic, faces)
        seed = torch.from_numpy(op._seed((r, r, r), faces))
        fn = lambda x: _t_trapped(torch.sigmoid(20.0 * (0.5 - x.reshape(r, r, r))), mcr, rad, 12, periodic, seed,
                                  (slice(None),) * 3)
        Pt, gt = _grad(fn, rho)
        assert abs(P - Pt) <= 1e-14 * max(1.0, abs(Pt))
        assert np.abs(g - gt).max() <= 1e-13 * max(1.0, np.abs(gt).max())


@pytest.mark.parametrize("tiles,focus,mcr", [(2, True, 3), (3, True, 1), (3, False, 0), ((2, 1, 3), True, 1)])
def test_periodic_array_matches_autograd(tiles, focus, mcr):
    r = 6
    txyz = tiles if isinstance(tiles, tuple) else (tiles,) * 3
    for rho in _designs(r, 3):
        P, g = op.periodic_array(rho, r, tiles_xyz=txyz, focus_center=focus, min_channel_rad=mcr, iters=10)
        R = tuple(t * r for t in txyz)
        seed = torch.from_numpy(op._seed(R, "all"))
        if focus and min(txyz) >= 3:
            region = tuple(slice((t // 2) * r, (t // 2 + 1) * r) for t in txyz)
        else:
            region = (slice(None),) * 3

        def fn(x):
            v = torch.sigmoid(20.0 * (0.5 - x.reshape(r, r, r))).repeat(*txyz)
            return _t_trapped(v, mcr, 1, 10, False, seed, region)
        Pt, gt = _grad(fn, rho)
        assert abs(P - Pt) <= 1e-14 * max(1.0, abs(Pt))
        assert np.abs(g - gt).max() <= 1e-13 * max(1.0, np.abs(gt).max())


def test_overhang_closed_forms():
    r = 8
    solid = np.ones(r ** 3)
    P, g = op.overhang(solid, r)
    assert P == 0.0
    slab = np.full((r, r, r), 1e-9)
    slab[:, :, 4] = 1.0
    P, _ = op.overhang(slab.ravel(), r)
    assert abs(P - (1.0 - 1e-9) / (r - 1)) <= 1e-15
    # a supported column is free, the same column floating one layer up is not
    col = np.zeros((r, r, r))
    col[3, 3, :] = 1.0
    assert op.overhang(col.ravel(), r)[0] == 0.0
    col[3, 3, 0] = 0.0
    assert op.overhang(col.ravel(), r)[0] > 0.0


def test_cavity_closed_forms():
    r = 10
    shell = np.full((r, r, r), 1e-9)
    shell[2:8, 2:8, 2:8] = 1.0
    shell[4:6, 4:6, 4:6] = 1e-9          # a sealed 2^3 void inside a solid block
    sv, ss = op._sigmoid(20 * (0.5 - 1e-9)), op._sigmoid(20 * (0.5 - 1.0))
    P, _ = op.closed_void(shell.ravel(), r, iters=0)
    assert abs(P - 8 * (sv - ss) / r ** 3) <= 1e-12
    opened = shell.copy()
    opened[4:6, 4:6, 2:4] = 1e-9         # a channel to the outside void
    P, _ = op.closed_void(opened.ravel(), r, iters=0)
    assert P <= 1e-12


def test_finite_differences():
    r = 6
    rng = np.random.default_rng(9)
    rho = rng.uniform(0.05, 0.95, r ** 3)
    h = 1e-7
    for fn in (lambda x: op.overhang(x, r),
               lambda x: op.closed_void(x, r, min_channel_rad=1, iters=8),
               lambda x: op.periodic_array(x, r, tiles=2, min_channel_rad=1, iters=8)):
        P, g = fn(rho)
        for e in rng.choice(r ** 3, 6, replace=False):
            xp, xm = rho.copy(), rho.copy()
            xp[e] += h
            xm[e] -= h
            fd = (fn(xp)[0] - fn(xm)[0]) / (2 * h)
            assert abs(fd - g[e]) <= 1e-6 * max(1.0, abs(g[e]))
