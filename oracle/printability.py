"""Additive-manufacturing penalties and their adjoints (oracle, FP64).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates, in numpy, the three differentiable penalties of the reference's
printability module and the gradients its autograd takes through them:

* ``overhang``       printability_penalty_overhang   (printability.py:91-134)
* ``closed_void``    cavity_penalty_closed_void      (printability.py:139-223)
* ``periodic_array`` cavity_penalty_periodic_array   (printability.py:229-323)
  with the soft opening of printability.py:8-71 (erosion = min-pool padded
  with 1, dilation = max-pool padded with 0).

Grids are (x, y, z) = rho.reshape(r, r, r) of the element vector (element id
(x*r + y)*r + z), build direction +z.  Each function returns (P, dP/drho).

Subgradients are those of the reference's autograd:
* a max-pool sends a window's gradient to the FIRST maximum of the window in
  (x, y, z) raster order (a strict ``>`` scan); a window whose first maximum
  is a padding position drops its gradient;
* min-pool (erosion, -max(-t)) likewise to the first minimum;
* ``minimum(a, b)`` sends the gradient to the smaller operand and halves it
  between the two on a tie;
* ``relu`` passes it where its argument is > 0.

Pinned by tests/test_oracle_printability.py against torch autograd of the same
compositions written with torch's own pooling ops (float64), by finite
differences at points without ties, and by closed forms (fully supported
columns, a floating slab, a sealed cube of void, an open channel).
"""
from __future__ import annotations

import numpy as np


def support_radius(overhang_deg: float) -> int:
    """Lateral support radius per layer for an overhang angle from the build plate
    (printability.py:77-88): ceil(tan(90 - clip(deg, 1, 89))), at least 1."""
    od = float(np.clip(overhang_deg, 1.0, 89.0))
    return max(1, int(np.ceil(np.tan(np.deg2rad(90.0 - od)))))


# ---- pooling with the first-maximum rule ----------------------------------------

def _pad(a, rad, periodic, val):
    if periodic:
        return np.pad(a, rad, mode="wrap")
    return np.pad(a, rad, mode="constant", constant_values=val)


def maxpool(a, rad, periodic, pad_val=0.0, minimum=False):
    """Window max (min) of an nd array over (2 rad + 1)^nd with stride 1.

    Returns (out, src): src[i] is the flat index, in the padded array, of the first
    maximum of window i in raster order."""
    p = _pad(np.asarray(a, dtype=np.float64), rad, periodic, pad_val)
    if minimum:
        p = -p
    shape = a.shape
    K = 2 * rad + 1
    # the windows, last axis in raster order of the offsets
    idx = np.indices(shape).reshape(len(shape), -1)
    offs = np.indices((K,) * len(shape)).reshape(len(shape), -1)
    pos = idx[:, :, None] + offs[:, None, :]                    # (nd, n, K^nd)
    flat = np.ravel_multi_index(tuple(pos), p.shape)            # (n, K^nd)
    vals = p.ravel()[flat]
    k = np.argmax(vals, axis=1)                                  # first maximum
    src = flat[np.arange(flat.shape[0]), k]
    out = p.ravel()[src].reshape(shape)
    return (-out if minimum else out), src.reshape(shape), p.shape


def maxpool_T(g, src, pshape, rad, periodic):
    """Adjoint of maxpool: scatter g to the window sources, fold the padding."""
    gp = np.zeros(int(np.prod(pshape)))
    np.add.at(gp, src.ravel(), g.ravel())
    gp = gp.reshape(pshape)
    nd = g.ndim
    if not periodic:
        return gp[tuple(slice(rad, rad + n) for n in g.shape)].copy()
    out = gp
    for ax in range(nd):
        n = g.shape[ax]
        core = np.take(out, np.arange(rad, rad + n), axis=ax)
        lo = np.take(out, np.arange(0, rad), axis=ax)             # padding left = last rad entries
        hi = np.take(out, np.arange(rad + n, rad + n + rad), axis=ax)
        idx_lo = [slice(None)] * out.ndim
        idx_hi = [slice(None)] * out.ndim
        idx_lo[ax] = slice(n - rad, n)
        idx_hi[ax] = slice(0, rad)
        core[tuple(idx_lo)] += lo
        core[tuple(idx_hi)] += hi
        out = core
    return out


def _min_grads(a, b):
    """d minimum(a, b) / da, / db with the tie halved."""
    wa = np.where(a < b, 1.0, np.where(a == b, 0.5, 0.0))
    return wa, 1.0 - wa


# ---- overhang ------------------------------------------------------------------------

def overhang(rho, r, overhang_deg=45.0, periodic_xy=True):
    """Returns (P, dP/drho) of printability_penalty_overhang (printability.py:91-134)."""
    g = np.asarray(rho, dtype=np.float64).reshape(r, r, r)
    rad = support_radius(overhang_deg)
    c = 1.0 / max(r - 1, 1)
    sup = g[:, :, 0].copy()
    tape = []
    P = 0.0
    for z in range(1, r):
        s, src, psh = maxpool(sup, rad, periodic_xy, 0.0)
        rz = g[:, :, z]
        P += np.maximum(rz - s, 0.0).mean()
        tape.append((s, src, psh))
        sup = np.minimum(rz, s)
    P *= c
    grad = np.zeros_like(g)
    gsup = np.zeros((r, r))
    n2 = float(r * r)
    for z in range(r - 1, 0, -1):
        s, src, psh = tape[z - 1]
        rz = g[:, :, z]
        pos = (rz - s > 0.0).astype(np.float64) * (c / n2)
        wr, ws = _min_grads(rz, s)
        grad[:, :, z] += pos + gsup * wr
        gs = -pos + gsup * ws
        gsup = maxpool_T(gs, src, psh, rad, periodic_xy)
    grad[:, :, 0] += gsup
    return P, grad.ravel()


# ---- trapped voids ---------------------------------------------------------------------

def _seed(shape, faces):
    s = np.zeros(shape)
    f = str(faces).lower()
    if f == "all":
        s[0], s[-1] = 1, 1
        s[:, 0], s[:, -1] = 1, 1
        s[:, :, 0], s[:, :, -1] = 1, 1
    elif f == "x":
        s[0], s[-1] = 1, 1
    elif f == "y":
        s[:, 0], s[:, -1] = 1, 1
    elif f in ("z", "xy"):
        s[:, :, 0], s[:, :, -1] = 1, 1
    else:
        raise ValueError("seed_faces must be one of: all, x, y, z, xy")
    return s


def _trapped(void, seed, mcr, rad, nsteps, periodic, region):
    """Forward and adjoint of the flood fill; returns (P, dP/dvoid)."""
    if mcr > 0:
        er, src_e, psh_e = maxpool(void, mcr, periodic, 1.0, minimum=True)
        vc, src_d, psh_d = maxpool(er, mcr, periodic, 0.0)
    else:
        vc = void
    reach = vc * seed
    tape = []
    if rad > 0:
        for _ in range(nsteps):
            dil, src, psh = maxpool(reach, rad, periodic, 0.0)
            tape.append((dil, src, psh))
            reach = np.minimum(vc, dil)
    trapped = np.maximum(void - reach, 0.0)
    cnt = trapped[region]
    P = float(cnt.mean())
    gt = np.zeros_like(void)
    gt[region] = 1.0 / cnt.size
    gt *= (void - reach > 0.0)
    gvoid = gt.copy()
    greach = -gt
    gvc = np.zeros_like(void)
    for t in range(len(tape) - 1, -1, -1):
        dil, src, psh = tape[t]
        prev_vc = vc
        wv, wd = _min_grads(prev_vc, dil)
        gvc += greach * wv
        greach = maxpool_T(greach * wd, src, psh, rad, periodic)
    gvc += greach * seed
    if mcr > 0:
        ger = maxpool_T(gvc, src_d, psh_d, mcr, periodic)
        gvoid += maxpool_T(ger, src_e, psh_e, mcr, periodic)
    else:
        gvoid += gvc
    return P, gvoid


def _sigmoid(t):
    return 1.0 / (1.0 + np.exp(-t))


def closed_void(rho, r, void_thresh=0.5, beta=20.0, min_channel_rad=0, rad=1, iters=0, periodic_xyz=True,
                seed_faces="all"):
    """Returns (P, dP/drho) of cavity_penalty_closed_void (printability.py:139-223)."""
    g = np.asarray(rho, dtype=np.float64).reshape(r, r, r)
    nsteps = int(iters) if int(iters) > 0 else r
    void = _sigmoid(float(beta) * (float(void_thresh) - g))
    P, gvoid = _trapped(void, _seed(g.shape, seed_faces), int(max(0, min_channel_rad)), int(max(0, rad)), nsteps,
                        bool(periodic_xyz), (slice(None),) * 3)
    return P, (gvoid * (-float(beta)) * void * (1.0 - void)).ravel()


def periodic_array(rho, r, tiles=3, tiles_xyz=None, focus_center=True, void_thresh=0.5, beta=20.0,
                   min_channel_rad=0, rad=1, iters=0, seed_faces="all"):
    """Returns (P, dP/drho) of cavity_penalty_periodic_array (printability.py:229-323)."""
    g = np.asarray(rho, dtype=np.float64).reshape(r, r, r)
    tx, ty, tz = (int(tiles),) * 3 if tiles_xyz is None else tuple(int(t) for t in tiles_xyz)
    tx, ty, tz = max(1, tx), max(1, ty), max(1, tz)
    vuc = _sigmoid(float(beta) * (float(void_thresh) - g))
    void = np.tile(vuc, (tx, ty, tz))
    nsteps = int(iters) if int(iters) > 0 else max(void.shape)
    if focus_center and tx >= 3 and ty >= 3 and tz >= 3:
        region = tuple(slice((t // 2) * r, (t // 2 + 1) * r) for t in (tx, ty, tz))
    else:
        region = (slice(None),) * 3
    P, gvoid = _trapped(void, _seed(void.shape, seed_faces), int(max(0, min_channel_rad)), int(max(0, rad)), nsteps,
                        False, region)
    guc = gvoid.reshape(tx, r, ty, r, tz, r).sum(axis=(0, 2, 4))
    return P, (guc * (-float(beta)) * vuc * (1.0 - vuc)).ravel()
