"""Drop-in for the reference's OC driver, ``optimize_inverse_homogenization``
(main_IHD_v3.py:492-891), running every per-element step on the device.

Same keyword signature and the same outputs in ``output_dir``: ``iter_log.csv``
(header and columns of main_IHD_v3.py:600-603, 733-736), ``summary.json`` (the
keys of 827-872) and ``rho_final.npy`` (float32 (res, res, res), 812).  The loop
follows 603-758: x -> symmetry projection -> SIMP -> both homogenizations ->
objective -> dJ/dx (the converged sensitivity pulled back through SIMP and the
symmetry projection, which is self-adjoint) -> periodic box filter -> OC ->
symmetry projection of the new x -> early stop after `patience` stalls
(tol_x 1e-3, tol_J 1e-4, min_iters 10).

The printability and cavity penalties (printability.py:91-323) run on the device
with their adjoints and enter as J_total = J_raw - lambda_print P_print -
lambda_cavity P_cavity (main_IHD_v3.py:650-688), with the reference's choice of
the one-cell or the tiled-array cavity form (tiles > 1, main_IHD_v3.py:657-683).
Not on this path: mesh export (``mesh_every`` and ``save_initial_mesh`` are
accepted and write no mesh; the summary says so).
"""
from __future__ import annotations

import csv
import json
import os
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from .cells import CellBatch
from .objective import evaluate_np, parse_objective_spec


def parse_symmetry(symmetry: str) -> Tuple[str, List[int]]:
    """("none" | "mirror" | "cubic", axes) as main_IHD_v3.py:221-255."""
    sym = (symmetry or "none").lower().strip()
    if sym in ("none", ""):
        return "none", []
    if sym in ("octant", "1/8", "one_eighth"):
        sym = "xyz"
    if sym in ("cubic", "iso", "isotropic"):
        return "cubic", [0, 1, 2]
    axes = sorted({a for a, ch in enumerate("xyz") if ch in sym})
    return ("mirror", axes) if axes else ("none", [])


def symmetrize_vec(x_vec: torch.Tensor, elem_coords: Optional[torch.Tensor], res: int, kind: str,
                   axes: List[int], cells: Optional[CellBatch] = None) -> torch.Tensor:
    """Symmetry projection of an (N, 1) element vector (main_IHD_v3.py:318-327), on the device."""
    if kind == "none":
        return x_vec.reshape(-1, 1)
    if cells is None:
        from .optim import _batch_for
        cells = _batch_for(res, x_vec.device)
    return cells.symmetrize(x_vec.reshape(-1), kind, axes).reshape(-1, 1).to(x_vec.dtype)


@torch.no_grad()
def make_random_x_init(res: int, volfrac: float, xmin: float, device, iters: int = 5) -> torch.Tensor:
    """Random x in [xmin, 1] with mean volfrac (main_IHD_v3.py:383-402); torch's RNG."""
    x = torch.rand((res ** 3, 1), device=device, dtype=torch.float32)
    for _ in range(int(iters)):
        x = (x - x.mean() + float(volfrac)).clamp(xmin, 1.0)
    return (x - x.mean() + float(volfrac)).clamp(xmin, 1.0)


@dataclass
class RunResult:
    J_raw_init: float
    J_raw_final: float
    J_total_init: float
    J_total_final: float
    P_print_init: float
    P_print_final: float
    P_cavity_init: float
    P_cavity_final: float
    iters_run: int
    converged: bool
    rho_final: np.ndarray
    log_csv: Optional[str]
    summary_json: Optional[str]
    meshes_dir: str
    rho_final_npy: str
    final_obj: str


def _evaluate(cells: CellBatch, weights: Dict[str, float], rho: torch.Tensor, rtol: float, maxit: int):
    need_el = any(t in ("stiffness", "shear", "bulk") for t in weights)
    need_th = "thermal" in weights
    phys = tuple(p for p, n in (("elastic", need_el), ("thermal", need_th)) if n)
    CH, kap, _ = cells.homogenize(rho.reshape(-1).to(torch.float64), phys, rtol=rtol, maxit=maxit)
    CHh = CH[0].cpu().numpy() if CH is not None else None
    kh = kap[0].cpu().numpy() if kap is not None else None
    J, G6, G3 = evaluate_np(weights, CHh, kh)
    return J, G6, G3, need_el, need_th


def parse_tiles_xyz(tiles_xyz: Optional[str], tiles: int) -> Tuple[int, int, int]:
    """Tiling of the array-cavity check: "tx,ty,tz" (or spaces) overriding `tiles`
    (main_IHD_v3.py:101-115)."""
    if tiles_xyz is None:
        t = max(1, int(tiles))
        return t, t, t
    parts = [p for p in str(tiles_xyz).replace(" ", ",").split(",") if p != ""]
    if len(parts) != 3:
        raise ValueError("--cavity_tiles_xyz must have 3 integers, e.g. 3,3,3")
    return tuple(max(1, int(v)) for v in parts)


class Penalties:
    """The penalty terms of one batch: P values per cell and the summed gradient
    lambda_print dP_print/drho + lambda_cavity dP_cavity/drho (torch device arrays)."""

    def __init__(self, cells: CellBatch, lambda_print: float = 0.0, overhang_deg: float = 45.0,
                 lambda_cavity: float = 0.0, cavity_thresh: float = 0.5, cavity_beta: float = 20.0,
                 min_channel_rad: int = 0, cavity_rad: int = 1, cavity_seed_faces: str = "all",
                 cavity_iters: int = 128, cavity_tiles: int = 1, cavity_tiles_xyz: Optional[str] = None,
                 cavity_focus_center: bool = True):
        self.cells = cells
        self.lp, self.lc = float(lambda_print), float(lambda_cavity)
        self.deg = float(overhang_deg)
        if self.lc > 0.0:
            tiles = parse_tiles_xyz(cavity_tiles_xyz, cavity_tiles)
            cells.set_cavity(void_thresh=cavity_thresh, beta=cavity_beta, min_channel_rad=min_channel_rad,
                             rad=cavity_rad, iters=cavity_iters, periodic_xyz=True, seed_faces=cavity_seed_faces,
                             tiles=tiles, focus_center=cavity_focus_center)

    @property
    def active(self) -> bool:
        return self.lp > 0.0 or self.lc > 0.0

    def __call__(self, rho: torch.Tensor, grad: bool = True):
        """Returns (P_print (ncells,), P_cavity (ncells,), lambda-weighted dP/drho or None)."""
        n = self.cells.ncells
        zero = np.zeros(n)
        Pp, Pc, g = zero, zero, None
        if self.lp > 0.0:
            P, dP = self.cells.penalty_overhang(rho, self.deg, True, grad)
            Pp = P.cpu().numpy()
            g = self.lp * dP if grad else None
        if self.lc > 0.0:
            P, dP = self.cells.penalty_cavity(rho, grad)
            Pc = P.cpu().numpy()
            if grad:
                g = self.lc * dP if g is None else g.add_(dP, alpha=self.lc)
        return Pp, Pc, g


def optimize_inverse_homogenization(ihd_module, batch: Dict[str, torch.Tensor], *, res: int, objective: str,
                                    init_mode: str, volfrac: float, penal: float, rho_min: float, move: float,
                                    xmin: float, iters: int, seed: int, symmetry: str, lambda_print: float = 0.0,
                                    overhang_deg: float = 45.0, lambda_cavity: float = 0.0,
                                    cavity_thresh: float = 0.5, cavity_beta: float = 20.0, min_channel_rad: int = 0,
                                    cavity_rad: int = 1, cavity_seed_faces: str = "all", cavity_iters: int = 128,
                                    cavity_tiles: int = 1, cavity_tiles_xyz: Optional[str] = None,
                                    cavity_focus_center: bool = True, filter_k: int = 3, output_dir: str = ".",
                                    mesh_every: int = 0, mesh_res: int = 0, mesh_iso: float = 0.5, mesh_pad: int = 1,
                                    save_iter_log: bool = True, save_summary: bool = True,
                                    save_initial_mesh: bool = False, verbose: bool = True,
                                    rtol: Optional[float] = None, maxit: Optional[int] = None) -> RunResult:
    cells: CellBatch = ihd_module.cells
    if cells.ncells != 1 or cells.dims[0] != (res, res, res):
        raise ValueError(f"ihd_module must hold one {res}^3 cell")
    rtol = float(ihd_module.rtol if rtol is None else rtol)
    maxit = int(ihd_module.maxit if maxit is None else maxit)
    device = cells.device
    torch.manual_seed(seed)
    np.random.seed(seed)
    os.makedirs(output_dir, exist_ok=True)
    meshes_dir = os.path.join(output_dir, "meshes")
    log_csv = os.path.join(output_dir, "iter_log.csv") if save_iter_log else None
    summary_json = os.path.join(output_dir, "summary.json") if save_summary else None
    final_rho_npy = os.path.join(output_dir, "rho_final.npy")
    final_obj = os.path.join(meshes_dir, "final.obj")
    weights = parse_objective_spec(objective)
    kind, axes = parse_symmetry(symmetry)
    S = (lambda v: cells.symmetrize(v, kind, axes)) if kind != "none" else (lambda v: v)  # noqa: E731

    if init_mode == "noise":
        x = make_random_x_init(res, volfrac, xmin, device).reshape(-1).to(torch.float64)
    else:
        rho0 = batch["rho"].reshape(-1).to(torch.float64).clamp(rho_min, 1.0)
        x = ((rho0 - rho_min) / (1.0 - rho_min)).clamp(0.0, 1.0).pow(1.0 / penal).clamp(xmin, 1.0)
        if volfrac <= 0:
            volfrac = float(x.mean())
    if kind != "none":
        x = S(x).clamp(xmin, 1.0)

    simp = lambda xx: rho_min + (1.0 - rho_min) * xx ** penal  # noqa: E731
    pen = Penalties(cells, lambda_print, overhang_deg, lambda_cavity, cavity_thresh, cavity_beta, min_channel_rad,
                    cavity_rad, cavity_seed_faces, cavity_iters, cavity_tiles, cavity_tiles_xyz, cavity_focus_center)
    rho_init = simp(x).clamp(rho_min, 1.0)
    J_raw_init = _evaluate(cells, weights, rho_init, rtol, maxit)[0]
    Pp, Pc, _ = pen(rho_init, grad=False)
    P_print_init, P_cavity_init = float(Pp[0]), float(Pc[0])
    J_total_init = float(J_raw_init) - float(lambda_print) * P_print_init - float(lambda_cavity) * P_cavity_init
    if log_csv is not None:
        with open(log_csv, "w", newline="") as f:
            csv.writer(f).writerow(["iter", "J_total", "J_raw", "P_print", "P_cavity", "mean_x", "mean_rho",
                                    "dx_max", "dx_mean", "time_sec"])
    t0 = time.time()
    J_prev = None
    stall, converged, it_run = 0, False, 0
    tol_x, tol_J, patience, min_iters = 1e-3, 1e-4, 10, 10
    for it in range(iters):
        it_run = it + 1
        x_old = x
        xu = S(x.clamp(xmin, 1.0)).clamp(xmin, 1.0)
        rho = simp(xu)
        J, G6, G3, need_el, need_th = _evaluate(cells, weights, rho, rtol, maxit)
        d = cells.sensitivity(torch.from_numpy(G6).to(device).reshape(1, 6, 6) if need_el else None,
                              torch.from_numpy(G3).to(device).reshape(1, 3, 3) if need_th else None)
        Pp, Pc, gpen = pen(rho)
        P_print, P_cavity = float(Pp[0]), float(Pc[0])
        J_total = float(J) - float(lambda_print) * P_print - float(lambda_cavity) * P_cavity
        if gpen is not None:
            d = d - gpen
        dJdx = S(d * (penal * (1.0 - rho_min)) * xu ** (penal - 1.0))
        if filter_k and int(filter_k) > 1:
            dJdx = cells.filter(dJdx, int(filter_k))
        x_new = cells.oc_update(x_old, dJdx, volfrac, move, xmin)
        x_next = S(x_new).clamp(xmin, 1.0)
        dx = (x_next - x_old).abs()
        dx_max, dx_mean = float(dx.max()), float(dx.mean())
        dJ_rel = float("inf") if J_prev is None else abs(J_total - J_prev) / max(abs(J_prev), 1e-12)
        stall = stall + 1 if (it >= min_iters and dx_max < tol_x and dJ_rel < tol_J) else 0
        x = x_next
        J_prev = J_total
        x_mean, rho_mean = float(xu.mean()), float(rho.mean())
        if log_csv is not None:
            with open(log_csv, "a", newline="") as f:
                csv.writer(f).writerow([it_run, J_prev, float(J), P_print, P_cavity, x_mean, rho_mean, dx_max,
                                        dx_mean, time.time() - t0])
        if verbose:
            print(f"it={it_run:04d}  J_total={J_prev:.6e}  J_raw={float(J):.6e}  P_print={P_print:.6e}  "
                  f"P_cavity={P_cavity:.6e}  mean_x={x_mean:.4f}  mean_rho={rho_mean:.4f}  dx_max={dx_max:.3e}  "
                  f"dx_mean={dx_mean:.3e}  stall={stall}/{patience}", flush=True)
        if stall >= patience:
            converged = True
            if verbose:
                print(f"[EARLY STOP] Converged at it={it_run} for {patience} consecutive checks.", flush=True)
            break

    x = S(x).clamp(xmin, 1.0)
    rho_final = simp(x).clamp(rho_min, 1.0)
    J_raw_final = _evaluate(cells, weights, rho_final, rtol, maxit)[0]
    Pp, Pc, _ = pen(rho_final, grad=False)
    P_print_final, P_cavity_final = float(Pp[0]), float(Pc[0])
    J_total_final = float(J_raw_final) - float(lambda_print) * P_print_final - float(lambda_cavity) * P_cavity_final
    rho_final_grid = rho_final.reshape(res, res, res).cpu().numpy().astype(np.float32)
    np.save(final_rho_npy, rho_final_grid)
    summary = {
        "objective": objective, "init_mode": init_mode, "symmetry": str(symmetry),
        "lambda_print": float(lambda_print), "overhang_deg": float(overhang_deg),
        "lambda_cavity": float(lambda_cavity), "cavity_thresh": float(cavity_thresh),
        "cavity_beta": float(cavity_beta), "min_channel_rad": int(min_channel_rad), "cavity_rad": int(cavity_rad),
        "cavity_seed_faces": str(cavity_seed_faces), "cavity_iters": int(cavity_iters),
        "cavity_tiles": int(cavity_tiles),
        "cavity_tiles_xyz": (None if cavity_tiles_xyz is None else str(cavity_tiles_xyz)),
        "cavity_focus_center": bool(cavity_focus_center), "resolution": res, "iters_requested": iters,
        "iters_run": int(it_run), "converged": bool(converged), "seed": int(seed), "volfrac_target": float(volfrac),
        "penal": float(penal), "rho_min": float(rho_min), "move": float(move), "xmin": float(xmin),
        "mesh_every": int(mesh_every), "mesh_res": int(mesh_res), "mesh_iso": float(mesh_iso),
        "mesh_pad": int(mesh_pad), "J_raw_init": float(J_raw_init), "J_raw_final": float(J_raw_final),
        "J_total_init": J_total_init, "J_total_final": J_total_final, "P_print_init": P_print_init,
        "P_print_final": P_print_final, "P_cavity_init": P_cavity_init, "P_cavity_final": P_cavity_final,
        "log_csv": (None if log_csv is None else os.path.abspath(log_csv)),
        "rho_final_npy": os.path.abspath(final_rho_npy), "meshes_dir": os.path.abspath(meshes_dir),
        "meshes": "not written (mesh export is not on this path)",
        "time_sec": float(time.time() - t0),
    }
    if summary_json is not None:
        with open(summary_json, "w", encoding="utf-8") as f:
            json.dump(summary, f, ensure_ascii=False, indent=2)
    return RunResult(J_raw_init=float(J_raw_init), J_raw_final=float(J_raw_final), J_total_init=J_total_init,
                     J_total_final=J_total_final, P_print_init=P_print_init, P_print_final=P_print_final,
                     P_cavity_init=P_cavity_init, P_cavity_final=P_cavity_final, iters_run=int(it_run),
                     converged=bool(converged),
                     rho_final=rho_final_grid, log_csv=log_csv, summary_json=summary_json, meshes_dir=meshes_dir,
                     rho_final_npy=final_rho_npy, final_obj=final_obj)
