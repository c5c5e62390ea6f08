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


def run_driver(r, x0, weights, iters, volfrac, symmetry="none", nlevels=3, rtol=1e-10, move=0.2, xmin=1e-3,
               p=3.0, rho_min=1e-9, filter_k=3, tol_x=1e-3, tol_J=1e-4, patience=10, min_iters=10,
               lambda_print=0.0, overhang_deg=45.0, lambda_cavity=0.0, cavity=None):
    """The loop of optimize_inverse_homogenization (main_IHD_v3.py:603-758) without
    meshes: symmetry projection of x (and, by the chain rule, of dJ/dx), the AM
    penalties J_total = J - lambda_print P_print - lambda_cavity P_cavity
    (main_IHD_v3.py:650-688; cavity: keyword arguments of printability.closed_void, or
    of periodic_array when they hold "tiles_xyz"), filter, OC, early stop.  Returns
    (x, rows) with rows (iter, J_total, mean_x_used, mean_rho, dx_max, dx_mean,
    converged, J_raw, P_print, P_cavity)."""
    from . import printability as pr
    from .symmetry import parse_symmetry, symmetrize
    cav = dict(cavity or {})
    cav_fn = pr.periodic_array if "tiles_xyz" in cav else pr.closed_void
    kind, axes = parse_symmetry(symmetry)
    S = (lambda v: symmetrize(v, r, kind, axes)) if kind != "none" else (lambda v: np.asarray(v, dtype=np.float64))
    cs = CellSolver((r, r, r), nlevels=nlevels, rtol=rtol)
    x = np.clip(S(np.asarray(x0, dtype=np.float64)), xmin, 1.0)
    rows = []
    J_prev = None
    stall = 0
    for it in range(iters):
        x_old = x
        xu = np.clip(S(np.clip(x, xmin, 1.0)), xmin, 1.0)
        rho = simp(xu, p, rho_min)
        need_ch = any(t in ("stiffness", "shear", "bulk") for t in weights)
        need_k = "thermal" in weights
        CH = S6 = kap = S3 = None
        if need_k:
            kap, S3, _ = cs.thermal(rho)
        if need_ch:
            CH, S6, _ = cs.elastic(rho)
        J, G6, G3 = obj.evaluate(weights, CH, kap)
        dJdrho = np.zeros_like(rho)
        if need_ch:
            dJdrho += np.einsum("ij,eij->e", G6, S6)
        if need_k:
            dJdrho += np.einsum("ij,eij->e", G3, S3)
        Pp = Pc = 0.0
        if lambda_print > 0.0:
            Pp, g = pr.overhang(rho, r, overhang_deg, True)
            dJdrho = dJdrho - lambda_print * g
        if lambda_cavity > 0.0:
            Pc, g = cav_fn(rho, r, **cav)
            dJdrho = dJdrho - lambda_cavity * g
        J_raw = J
        J = J - lambda_print * Pp - lambda_cavity * Pc
        dJdx = S(dJdrho * simp_grad(xu, p, rho_min))
        if filter_k and filter_k > 1:
            dJdx = box_filter(dJdx, cs.dims, filter_k)
        xn, _ = oc_update(x_old, dJdx, volfrac, move, xmin)
        x_next = np.clip(S(xn), xmin, 1.0)
        dx = np.abs(x_next - x_old)
        dJ_rel = float("inf") if J_prev is None else abs(J - J_prev) / max(abs(J_prev), 1e-12)
        stall = stall + 1 if (it >= min_iters and dx.max() < tol_x and dJ_rel < tol_J) else 0
        x = x_next
        J_prev = J
        rows.append((it + 1, J, xu.mean(), rho.mean(), dx.max(), dx.mean(), stall >= patience, J_raw, Pp, Pc))
        if stall >= patience:
            break
    return x, rows
