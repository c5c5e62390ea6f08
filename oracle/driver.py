"""The plain OC loop on the CPU (oracle, FP64); also the timed CPU baseline.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

One OC iteration follows main_IHD_v3.py:620-716 with sym="none" and no
printability penalties:
  x_used = clip(x, xmin, 1); rho = rho_min + (1-rho_min) x_used^p     (625-628)
  kappa_H, C_H from converged solves                                  (639-645)
  J = sum_t w_t term_t                                                 (647)
  dJ/dx = (G6:S6_e + G3:S3_e) * p (1-rho_min) x_used^(p-1)            (688, closed form)
  dJ/dx <- periodic box filter k=3                                    (690-691)
  x <- OC bisection (move, xmin, volfrac on mean(x))                   (693)

Initial designs of the two single-cell workloads:
  ``x0_sphere_void`` : 0.25 everywhere, 0.1 for element centres within
                       0.3 L of the cell centre, + U(-0.02, 0.02) from
                       numpy default_rng(seed), clipped to [0.001, 1];
  ``x0_uniform_noise``: 0.25 + U(-0.05, 0.05) from default_rng(seed),
                       clipped to [0.001, 1].
"""
from __future__ import annotations

import time

import numpy as np

from . import objective as obj
from .homog import CellSolver
from .oc import box_filter, oc_update, simp, simp_grad


def x0_sphere_void(r: int, seed: int = 0, vf=0.25, void=0.1, radius=0.3, noise=0.02):
    c = (np.arange(r) + 0.5) / r - 0.5
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    x = np.full((r, r, r), vf)
    x[X ** 2 + Y ** 2 + Z ** 2 <= radius ** 2] = void
    x = x.ravel() + np.random.default_rng(seed).uniform(-noise, noise, size=r ** 3)
    return np.clip(x, 0.001, 1.0)


def x0_uniform_noise(r: int, seed: int = 0, vf=0.25, noise=0.05):
    x = vf + np.random.default_rng(seed).uniform(-noise, noise, size=r ** 3)
    return np.clip(x, 0.001, 1.0)


def oc_iteration(cs: CellSolver, x, weights, volfrac, move=0.2, xmin=1e-3, p=3.0,
                 rho_min=1e-9, filter_k=3, rtol=None):
    xu = np.clip(x, xmin, 1.0)
    rho = simp(xu, p, rho_min)
    need_ch = any(t in ("stiffness", "shear", "bulk") for t in weights)
    need_k = "thermal" in weights
    CH = S6 = kap = S3 = None
    if need_k:
        kap, S3, _ = cs.thermal(rho, rtol)
    if need_ch:
        CH, S6, _ = cs.elastic(rho, rtol)
    J, G6, G3 = obj.evaluate(weights, CH, kap)
    dJdrho = np.zeros_like(rho)
    if need_ch:
        dJdrho += np.einsum("ij,eij->e", G6, S6)
    if need_k:
        dJdrho += np.einsum("ij,eij->e", G3, S3)
    dJdx = dJdrho * simp_grad(xu, p, rho_min)
    if filter_k and filter_k > 1:
        dJdx = box_filter(dJdx, cs.dims, filter_k)
    xn, lam = oc_update(x, dJdx, volfrac, move, xmin)
    return xn, dict(J=J, CH=CH, kappa=kap, dJdx=dJdx, rho=rho, lam=lam)


def run(r, x0, weights, iters, volfrac=0.25, nlevels=3, rtol=1e-6, move=0.2, xmin=1e-3,
        warmup=0, verbose=False):
    """Run the loop; returns (x, history, seconds per timed iteration)."""
    cs = CellSolver((r, r, r), nlevels=nlevels, rtol=rtol)
    x = np.asarray(x0, dtype=np.float64).copy()
    hist = []
    times = []
    for it in range(warmup + iters):
        t0 = time.perf_counter()
        x, info = oc_iteration(cs, x, weights, volfrac, move, xmin)
        dt = time.perf_counter() - t0
        if it >= warmup:
            times.append(dt)
        hist.append(info["J"])
        if verbose:
            print(f"it={it:04d} J={info['J']:.6e} mean_x={x.mean():.4f} dt={dt:.2f}s", flush=True)
    return x, hist, times
