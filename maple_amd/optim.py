"""Optimizer-side pieces of the reference's driver, on the GPU.

* Module-level drop-ins with the reference's names and arguments:
  ``avg_filter3d_x_periodic`` (main_IHD_v3.py:342-359), ``oc_update``
  (main_IHD_v3.py:362-380), ``build_unitcell_batch`` (main_IHD_v3.py:162-215),
  ``assemble_F_fast`` / ``assemble_F_fast_scalar`` (_utils.py:453-507).
* ``OCDriver``: the OC loop of optimize_inverse_homogenization
  (main_IHD_v3.py:620-716, symmetry "none", the AM penalties when given) over a
  ``CellBatch`` of any number of cells, everything on the device except the
  6x6 objective algebra (symmetry projection: ``maple_amd.driver`` and
  ``maple_amd.densify``).
* Initial designs of the single-cell workloads.
"""
from __future__ import annotations

import weakref
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from .cells import CellBatch, max_levels
from .objective import evaluate_np, parse_objective_spec

_REGISTRY: "weakref.WeakValueDictionary" = weakref.WeakValueDictionary()


def _batch_for(res: int, device) -> CellBatch:
    key = (str(torch.device(device)), int(res))
    cb = _REGISTRY.get(key)
    if cb is None:
        cb = CellBatch([int(res)], nlevels=max_levels(int(res)), device=device)
        _REGISTRY[key] = cb
        _KEEP.append(cb)
    return cb


_KEEP: List[CellBatch] = []


def register(cells: CellBatch):
    """Let the module-level drop-ins reuse an existing single-cell context."""
    if cells.ncells == 1 and len(set(cells.dims[0])) == 1:
        _REGISTRY[(str(cells.device), cells.dims[0][0])] = cells


def _check_coords(elem_coords: Optional[torch.Tensor], res: int):
    if elem_coords is None:
        return
    N = res ** 3
    if elem_coords.shape[0] != N:
        raise ValueError(f"elem_coords must describe the full {res}^3 grid")


def avg_filter3d_x_periodic(x_vec: torch.Tensor, elem_coords: Optional[torch.Tensor], res: int, k: int = 3):
    """Periodic box filter of a (N,1) vector laid out on (res,res,res) (main_IHD_v3.py:342-359)."""
    if int(k) <= 0:
        raise ValueError(f"filter size must be positive, got {k}")
    if int(k) % 2 == 0:
        raise ValueError(f"filter size must be odd, got {k}")
    _check_coords(elem_coords, res)
    cb = _batch_for(res, x_vec.device)
    return cb.filter(x_vec, int(k)).reshape(-1, 1).to(x_vec.dtype)


@torch.no_grad()
def oc_update(x: torch.Tensor, dJdx: torch.Tensor, volfrac: float, move: float = 0.1, xmin: float = 1e-3,
              nbisect: int = 60):
    """OC for maximisation with the reference's 60-step bisection (main_IHD_v3.py:362-380)."""
    n = x.numel()
    res = round(n ** (1.0 / 3.0))
    if res ** 3 != n:
        raise ValueError("x must hold a cubic cell")
    cb = _batch_for(res, x.device)
    return cb.oc_update(x, dJdx, volfrac, move, xmin, nbisect).reshape(x.shape).to(x.dtype)


def build_unitcell_batch(device, res: int, rho_grid: torch.Tensor) -> Dict[str, torch.Tensor]:
    """The reference's periodic unit-cell dict (main_IHD_v3.py:162-215)."""
    if tuple(rho_grid.shape) != (res, res, res):
        raise ValueError(f"rho_grid must be {(res, res, res)}, got {tuple(rho_grid.shape)}")
    dev = torch.device(device)
    ar = torch.arange(res, device=dev)
    grid = torch.stack(torch.meshgrid(ar, ar, ar, indexing="ij"), dim=-1).reshape(-1, 3).long()
    hex8 = torch.tensor([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
                        dtype=torch.int64, device=dev)
    coo = (grid[:, None, :] + hex8[None]) % res
    node_index = (coo[..., 0] * res + coo[..., 1]) * res + coo[..., 2]
    rg = rho_grid.to(dev)
    feats = rg.reshape(-1)[node_index].to(torch.float32)
    b0 = torch.zeros((grid.shape[0], 1), dtype=torch.long, device=dev)
    coord = torch.cat([b0, grid], dim=1)
    return {
        "coord": coord,
        "feature": feats,
        "node_index": node_index,
        "elem_coords": coord.clone(),
        "rho": rg.reshape(-1, 1),
        "voxel_batch_id": torch.zeros((grid.shape[0],), dtype=torch.long, device=dev),
    }


def assemble_F_fast(node_index: torch.Tensor, Fe: torch.Tensor, n_nodes: int, rho: torch.Tensor):
    """Elastic RHS (N, 3, 6) (_utils.py:480-507), assembled on the device."""
    res = round(n_nodes ** (1.0 / 3.0))
    cb = _batch_for(res, rho.device)
    F = cb.rhs("elastic", rho.reshape(-1))                   # (6, 3, N)
    return F.permute(2, 1, 0).contiguous()


def assemble_F_fast_scalar(node_index: torch.Tensor, Fe: torch.Tensor, n_nodes: int, rho: torch.Tensor):
    """Thermal RHS (N, 3) (_utils.py:453-478), assembled on the device."""
    if Fe.dim() != 2 or Fe.shape[0] != 8:
        raise ValueError(f"Expected Fe shape (8,3), got {tuple(Fe.shape)}")
    if node_index.dim() != 2 or node_index.shape[1] != 8:
        raise ValueError(f"Expected node_index shape (E,8), got {tuple(node_index.shape)}")
    res = round(n_nodes ** (1.0 / 3.0))
    cb = _batch_for(res, rho.device)
    F = cb.rhs("thermal", rho.reshape(-1))                   # (3, 1, N)
    return F[:, 0, :].T.contiguous()


# ---- workloads -------------------------------------------------------------

def x0_sphere_void(r: int, seed: int = 0, vf=0.25, void=0.1, radius=0.3, noise=0.02) -> np.ndarray:
    """0.25 everywhere, 0.1 for element centres within 0.3 L of the centre, + U(-0.02, 0.02)
    from numpy default_rng(seed), clipped to [0.001, 1]."""
    c = (np.arange(r) + 0.5) / r - 0.5
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    x = np.full((r, r, r), vf)
    x[X ** 2 + Y ** 2 + Z ** 2 <= radius ** 2] = void
    x = x.ravel() + np.random.default_rng(seed).uniform(-noise, noise, size=r ** 3)
    return np.clip(x, 0.001, 1.0)


def x0_uniform_noise(r: int, seed: int = 0, vf=0.25, noise=0.05) -> np.ndarray:
    """vf + U(-0.05, 0.05) from numpy default_rng(seed), clipped to [0.001, 1]."""
    x = vf + np.random.default_rng(seed).uniform(-noise, noise, size=r ** 3)
    return np.clip(x, 0.001, 1.0)


def simplex_grid(step: float = 0.05):
    """The (w_E, w_G, w_kappa) grid on the simplex with the given step (231 points at 5 %)."""
    n = int(round(1.0 / step))
    return [(i / n, j / n, (n - i - j) / n) for i in range(n + 1) for j in range(n + 1 - i)]


def batched_workload(sizes=((32, 16), (48, 32), (64, 16)), base_seed: int = 0):
    """The mixed-resolution batch: per cell its resolution, seed (base + index), objective
    weights from the 5 % simplex grid, volume fraction U[0.05, 0.40] and x0 = vf + U(-0.05, 0.05)
    clipped to [0.001, 1] (numpy default_rng(seed))."""
    grid = simplex_grid(0.05)
    cells = []
    idx = 0
    for r, count in sizes:
        for _ in range(count):
            seed = base_seed + idx
            rng = np.random.default_rng(seed)
            wE, wG, wK = grid[int(rng.integers(len(grid)))]
            vf = float(rng.uniform(0.05, 0.40))
            x0 = np.clip(vf + rng.uniform(-0.05, 0.05, r ** 3), 0.001, 1.0)
            w = {k: v for k, v in (("stiffness", wE), ("shear", wG), ("thermal", wK)) if v > 0}
            cells.append({"r": r, "seed": seed, "weights": w, "volfrac": vf, "x0": x0})
            idx += 1
    return cells


# ---- the OC loop -------------------------------------------------------------

class OCDriver:
    """Plain OC loop over every cell of a CellBatch (one OC step = all cells)."""

    def __init__(self, cells: CellBatch, objectives: Sequence, volfrac: Sequence[float], penal: float = 3.0,
                 rho_min: float = 1e-9, move: float = 0.2, xmin: float = 1e-3, filter_k: int = 3,
                 rtol: float = 1e-6, maxit: int = 1000, physics: Optional[Sequence[str]] = None, start=None,
                 strict: bool = True, penalties=None):
        """physics: solve these physics whatever the objectives need (None: what they need).
        start: callable(rho, physics) that supplies start candidates before each
        homogenization (e.g. maple_amd.learned.LearnedStart).
        strict: raise when a solve stops before rtol (maxit or CG breakdown); with
        strict=False the step goes on with the fields reached and reports the (cell,
        case) pairs that stopped early in info["unconverged"].
        penalties: a maple_amd.driver.Penalties of this batch (AM overhang and trapped-void
        terms, J_total = J - lambda P, main_IHD_v3.py:650-688)."""
        self.cells = cells
        n = cells.ncells
        objs = list(objectives) if not isinstance(objectives, (str, dict)) else [objectives] * n
        if len(objs) != n:
            raise ValueError("one objective per cell")
        self.weights = [parse_objective_spec(o) if isinstance(o, str) else dict(o) for o in objs]
        self.volfrac = np.broadcast_to(np.asarray(volfrac, dtype=np.float64), (n,)).copy()
        self.penal, self.rho_min, self.move, self.xmin = float(penal), float(rho_min), float(move), float(xmin)
        self.filter_k, self.rtol, self.maxit = int(filter_k), float(rtol), int(maxit)
        self.need_el = any(t in ("stiffness", "shear", "bulk") for w in self.weights for t in w)
        self.need_th = any("thermal" in w for w in self.weights)
        self.physics = tuple(p for p, need in (("elastic", self.need_el), ("thermal", self.need_th)) if need)
        if physics is not None:
            self.physics = tuple(physics)
            if (self.need_el and "elastic" not in self.physics) or (self.need_th and "thermal" not in self.physics):
                raise ValueError("the objectives need a physics that is not solved")
        self.start = start
        self.strict = bool(strict)
        self.penalties = penalties if (penalties is not None and penalties.active) else None

    def design_to_rho(self, x: torch.Tensor) -> torch.Tensor:
        xu = x.clamp(self.xmin, 1.0)
        return self.rho_min + (1.0 - self.rho_min) * xu ** self.penal

    @torch.no_grad()
    def step(self, x: torch.Tensor):
        xu = x.clamp(self.xmin, 1.0)
        rho = self.rho_min + (1.0 - self.rho_min) * xu ** self.penal
        if self.start is not None:
            self.start(rho, self.physics)
        CH, kap, iters = self.cells.homogenize(rho, self.physics, rtol=self.rtol, maxit=self.maxit,
                                               strict=self.strict)
        status = self.cells.last_status
        CHh = CH.cpu().numpy() if CH is not None else None
        kh = kap.cpu().numpy() if kap is not None else None
        n = self.cells.ncells
        J = np.zeros(n)
        G6 = np.zeros((n, 6, 6))
        G3 = np.zeros((n, 3, 3))
        for i, w in enumerate(self.weights):
            J[i], G6[i], G3[i] = evaluate_np(w, CHh[i] if CHh is not None else None, kh[i] if kh is not None else None)
        dev = self.cells.device
        dJdrho = self.cells.sensitivity(torch.from_numpy(G6).to(dev) if self.need_el else None,
                                        torch.from_numpy(G3).to(dev) if self.need_th else None)
        Pp = Pc = np.zeros(n)
        J_total = J.copy()
        if self.penalties is not None:
            Pp, Pc, gpen = self.penalties(rho)
            J_total = J - self.penalties.lp * Pp - self.penalties.lc * Pc
            dJdrho = dJdrho - gpen
        dJdx = dJdrho * (self.penal * (1.0 - self.rho_min)) * xu ** (self.penal - 1.0)
        if self.filter_k > 1:
            dJdx = self.cells.filter(dJdx, self.filter_k)
        xnew = self.cells.oc_update(x, dJdx, self.volfrac, self.move, self.xmin)
        cols = [k for p, ks in (("elastic", range(6)), ("thermal", range(6, 9))) if p in self.physics for k in ks]
        unconv = int((iters[:, cols] >= self.maxit).sum()) if status else 0
        return xnew, {"J": J, "J_total": J_total, "P_print": Pp, "P_cavity": Pc, "CH": CHh, "kappa": kh,
                      "iters": iters.copy(), "rho": rho, "dJdx": dJdx,
                      "dJdrho": dJdrho, "G6": G6, "G3": G3, "status": int(status), "unconverged": unconv}
