"""The Pareto-densification worker's optimization workload, batched on one GPU.

The reference runs one task per worker process (densify_pareto_property_network_32.py:
1055-1255): a warm-start density at 32^3, then optimize_inverse_homogenization with the
voxel initialisation, `iso` symmetry, the overhang and trapped-void penalties at
lambda = 0.9 and 2000 OC iterations (densify:184-201), then the homogenized terms of the
final density, one row of dataset_log.csv per task.  ``DensifyBatch`` advances many
tasks as one CellBatch: every OC step is one set of launches for all cells, and a cell
that meets the early-stop test (main_IHD_v3.py:695-716) keeps its design while the
others go on.

The warm start comes from the reference's Stage-II generator, which is not on this
path: the caller supplies each task's density (``rho_warm``).

Outputs, as the worker and the Stage-II preparation write them:
* per task ``sample_dir``: condition.npy (7 float32: w_E, w_G, w_kappa, vf and the
  three target properties), rho_warmstart_32.npy, rho_final.npy (float32 (r, r, r)),
  iter_log.csv (the driver's columns);
* ``write_dataset_log``: dataset_log.csv with the worker's FIELDNAMES (densify:35-82);
* ``write_stage2_npz``: Stage-II samples (prepare_stage2_training_data.py:330-354):
  condition (7,) float32, seed and source_index int64, rho_target_32 (1, 32, 32, 32)
  float16 (the final density block-averaged to 32^3).
"""
from __future__ import annotations

import csv
import os
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np
import torch

from .cells import CellBatch, max_levels
from .driver import Penalties, parse_symmetry
from .objective import evaluate_np, parse_objective_spec

# dataset_log.csv columns of the densify worker (densify_pareto_property_network_32.py:35-82)
FIELDNAMES = [
    "sample_id", "front_objectives", "weight_tag", "weight_idx", "w_stiffness", "w_shear", "w_thermal", "objective",
    "vf_target", "vf_final", "seed", "sampling_mode", "front_gap", "front_interval_rank", "tri_front_edge",
    "condition_npy", "condition_stiffness", "condition_shear", "condition_thermal", "condition_kappa",
    "source_lower_w", "source_upper_w", "source_alpha", "source_lower_seed", "source_upper_seed", "source_lower_vf",
    "source_upper_vf", "source_lower_stiffness", "source_lower_shear", "source_lower_thermal", "source_lower_kappa",
    "source_upper_stiffness", "source_upper_shear", "source_upper_thermal", "source_upper_kappa", "rho_net_32_npy",
    "rho_warmstart_32_npy", "stiffness", "shear", "thermal", "kappa", "sample_dir", "rho_final_npy", "final_obj",
    "checkpoint", "source_csv",
]

ITER_LOG_HEADER = ["iter", "J_total", "J_raw", "P_print", "P_cavity", "mean_x", "mean_rho", "dx_max", "dx_mean",
                   "time_sec"]

# the worker's optimisation settings (densify:184-201)
DENSIFY_DEFAULTS = dict(iters=2000, penal=3.0, rho_min=1e-5, move=0.05, xmin=1e-5, filter_k=3, symmetry="iso",
                        lambda_print=0.9, overhang_deg=45.0, lambda_cavity=0.9, cavity_thresh=0.5, cavity_beta=20.0,
                        min_channel_rad=3, cavity_rad=1, cavity_seed_faces="all", cavity_iters=128, cavity_tiles=2,
                        cavity_tiles_xyz=None, cavity_focus_center=True)


@dataclass
class Task:
    """One densification task: its warm-start density and the metadata of its row."""
    rho_warm: np.ndarray                  # (r, r, r) or r^3
    objective: str                        # e.g. "stiffness:0.4,shear:0.3,thermal:0.3"
    vf_target: float
    seed: int
    sample_dir: str
    condition: Optional[np.ndarray] = None   # 7 float32; built from the weights and vf if None
    meta: Dict[str, object] = field(default_factory=dict)   # further dataset_log.csv fields


def objective_terms(CH: Optional[np.ndarray], kappa: Optional[np.ndarray]) -> Dict[str, float]:
    """The stiffness, shear and thermal terms of a homogenized design
    (evaluate_homogenized_terms, main_IHD_v3.py:432-466)."""
    out = {}
    if CH is not None:
        out["stiffness"] = evaluate_np({"stiffness": 1.0}, CH, None)[0]
        out["shear"] = evaluate_np({"shear": 1.0}, CH, None)[0]
    if kappa is not None:
        out["thermal"] = evaluate_np({"thermal": 1.0}, None, kappa)[0]
    return out


class DensifyBatch:
    """Tasks of one resolution advanced together (one CellBatch; one launch set per OC step)."""

    def __init__(self, tasks: Sequence[Task], device=None, precision: str = "mixed", rtol: float = 1e-6,
                 maxit: int = 5000, nlevels: Optional[int] = None, islands: bool = True, start_history: int = 4,
                 **settings):
        if not tasks:
            raise ValueError("no tasks")
        self.s = dict(DENSIFY_DEFAULTS)
        unknown = set(settings) - set(self.s)
        if unknown:
            raise ValueError(f"unknown settings {sorted(unknown)}")
        self.s.update(settings)
        self.tasks = list(tasks)
        sizes = {int(round(np.asarray(t.rho_warm).size ** (1.0 / 3.0))) for t in self.tasks}
        if len(sizes) != 1:
            raise ValueError("the tasks of one batch share one resolution")
        self.r = sizes.pop()
        if self.r ** 3 != np.asarray(self.tasks[0].rho_warm).size:
            raise ValueError("rho_warm must hold a cubic grid")
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        n = len(self.tasks)
        self.cells = CellBatch([self.r] * n, nlevels=nlevels or max_levels(self.r), device=self.device,
                               precision=precision, warm_start=True, islands=islands)
        # each solve starts from the energy-optimal combination of the task's last solutions
        self.cells.set_start_history(start_history)
        self.rtol, self.maxit = float(rtol), int(maxit)
        self.weights = [parse_objective_spec(t.objective) for t in self.tasks]
        self.kind, self.axes = parse_symmetry(self.s["symmetry"])
        s = self.s
        self.pen = Penalties(self.cells, s["lambda_print"], s["overhang_deg"], s["lambda_cavity"], s["cavity_thresh"],
                             s["cavity_beta"], s["min_channel_rad"], s["cavity_rad"], s["cavity_seed_faces"],
                             s["cavity_iters"], s["cavity_tiles"], s["cavity_tiles_xyz"], s["cavity_focus_center"])

    def close(self):
        self.cells.close()

    def _S(self, v: torch.Tensor) -> torch.Tensor:
        return self.cells.symmetrize(v, self.kind, self.axes) if self.kind != "none" else v

    def _simp(self, x):
        return self.s["rho_min"] + (1.0 - self.s["rho_min"]) * x ** self.s["penal"]

    def initial_design(self) -> torch.Tensor:
        """Voxel initialisation of every task (main_IHD_v3.py:549-557): x = ((rho0 - rho_min) /
        (1 - rho_min))^(1/p), clamped, projected on the symmetry."""
        s = self.s
        rho0 = torch.from_numpy(np.concatenate([np.asarray(t.rho_warm, dtype=np.float32).ravel()
                                                for t in self.tasks])).to(self.device, torch.float64)
        rho0 = rho0.clamp(s["rho_min"], 1.0)
        x = ((rho0 - s["rho_min"]) / (1.0 - s["rho_min"])).clamp(0.0, 1.0).pow(1.0 / s["penal"]).clamp(s["xmin"], 1.0)
        return self._S(x).clamp(s["xmin"], 1.0)

    def run(self, iters: Optional[int] = None, on_step: Optional[Callable[[int, dict], None]] = None,
            write_files: bool = True) -> List[Dict[str, object]]:
        """Run the loop; returns one result dict per task (the dataset_log.csv row)."""
        s = self.s
        n, r, E = len(self.tasks), self.r, self.r ** 3
        iters = int(s["iters"] if iters is None else iters)
        phys = ("elastic", "thermal")
        vf = np.array([t.vf_target for t in self.tasks], dtype=np.float64)
        x = self.initial_design()
        xmin, penal, rho_min = s["xmin"], s["penal"], s["rho_min"]
        stall = np.zeros(n, dtype=np.int64)
        done = np.zeros(n, dtype=bool)
        iters_run = np.zeros(n, dtype=np.int64)
        J_prev = [None] * n
        logs: List[List[list]] = [[] for _ in range(n)]
        t0 = time.time()
        for it in range(iters):
            if done.all():
                break
            x_old = x
            xu = self._S(x.clamp(xmin, 1.0)).clamp(xmin, 1.0)
            rho = self._simp(xu)
            CH, kap, _ = self.cells.homogenize(rho, phys, rtol=self.rtol, maxit=self.maxit)
            CHh, kh = CH.cpu().numpy(), kap.cpu().numpy()
            J = np.zeros(n)
            G6 = np.zeros((n, 6, 6))
            G3 = np.zeros((n, 3, 3))
            for i, w in enumerate(self.weights):
                J[i], G6[i], G3[i] = evaluate_np(w, CHh[i], kh[i])
            d = self.cells.sensitivity(torch.from_numpy(G6).to(self.device), torch.from_numpy(G3).to(self.device))
            Pp, Pc, gpen = self.pen(rho)
            if gpen is not None:
                d = d - gpen
            J_total = J - s["lambda_print"] * Pp - s["lambda_cavity"] * Pc
            dJdx = self._S(d * (penal * (1.0 - rho_min)) * xu ** (penal - 1.0))
            if s["filter_k"] and int(s["filter_k"]) > 1:
                dJdx = self.cells.filter(dJdx, int(s["filter_k"]))
            x_new = self.cells.oc_update(x_old, dJdx, vf, s["move"], xmin)
            x_next = self._S(x_new).clamp(xmin, 1.0)
            # finished cells keep their design (their solves then start converged)
            keep = torch.from_numpy(np.repeat(done, E)).to(self.device)
            x_next = torch.where(keep, x_old, x_next)
            dx = (x_next - x_old).abs().reshape(n, E)
            dx_max = dx.max(dim=1).values.cpu().numpy()
            dx_mean = dx.mean(dim=1).cpu().numpy()
            x_mean = xu.reshape(n, E).mean(dim=1).cpu().numpy()
            rho_mean = rho.reshape(n, E).mean(dim=1).cpu().numpy()
            now = time.time() - t0
            for i in range(n):
                if done[i]:
                    continue
                dJ_rel = float("inf") if J_prev[i] is None else abs(J_total[i] - J_prev[i]) / max(abs(J_prev[i]), 1e-12)
                stall[i] = stall[i] + 1 if (it >= 10 and dx_max[i] < 1e-3 and dJ_rel < 1e-4) else 0
                J_prev[i] = float(J_total[i])
                iters_run[i] = it + 1
                logs[i].append([it + 1, float(J_total[i]), float(J[i]), float(Pp[i]), float(Pc[i]), float(x_mean[i]),
                                float(rho_mean[i]), float(dx_max[i]), float(dx_mean[i]), now])
                if stall[i] >= 10:
                    done[i] = True
            x = x_next
            if on_step is not None:
                on_step(it, {"J_total": J_total, "J": J, "P_print": Pp, "P_cavity": Pc, "done": done.copy()})
        # final designs and their homogenized terms (float32 densities, as the worker saves them)
        xf = self._S(x).clamp(xmin, 1.0)
        rho_f = self._simp(xf).clamp(rho_min, 1.0).to(torch.float32).to(torch.float64)
        CH, kap, _ = self.cells.homogenize(rho_f, phys, rtol=self.rtol, maxit=self.maxit)
        CHh, kh = CH.cpu().numpy(), kap.cpu().numpy()
        rho_np = rho_f.reshape(n, r, r, r).cpu().numpy().astype(np.float32)
        rows = []
        for i, t in enumerate(self.tasks):
            terms = objective_terms(CHh[i], kh[i])
            row = self._row(i, t, rho_np[i], terms, int(iters_run[i]), bool(done[i]))
            if write_files:
                self._write_sample(t, rho_np[i], logs[i], row)
            rows.append(row)
        return rows

    def _condition(self, t: Task) -> np.ndarray:
        if t.condition is not None:
            return np.asarray(t.condition, dtype=np.float32).reshape(7)
        w = parse_objective_spec(t.objective)
        props = [float(t.meta.get(k, 0.0) or 0.0) for k in ("stiffness", "shear", "thermal")]
        return np.array([w.get("stiffness", 0.0), w.get("shear", 0.0), w.get("thermal", 0.0), t.vf_target] + props,
                        dtype=np.float32)

    def _row(self, i: int, t: Task, rho: np.ndarray, terms: Dict[str, float], iters_run: int,
             converged: bool) -> Dict[str, object]:
        w = parse_objective_spec(t.objective)
        cond = self._condition(t)
        d = os.path.abspath(t.sample_dir)
        row = {k: "" for k in FIELDNAMES}
        row.update({k: v for k, v in t.meta.items() if k in row})
        row.update({
            "w_stiffness": w.get("stiffness", 0.0), "w_shear": w.get("shear", 0.0), "w_thermal": w.get("thermal", 0.0),
            "objective": t.objective, "vf_target": float(t.vf_target), "vf_final": float(rho.mean()),
            "seed": int(t.seed), "condition_npy": os.path.join(d, "condition.npy"),
            "condition_stiffness": float(cond[4]), "condition_shear": float(cond[5]),
            "condition_thermal": float(cond[6]), "condition_kappa": float(cond[6]),
            "rho_warmstart_32_npy": os.path.join(d, "rho_warmstart_32.npy"),
            "stiffness": terms["stiffness"], "shear": terms["shear"], "thermal": terms["thermal"],
            "kappa": terms["thermal"], "sample_dir": d, "rho_final_npy": os.path.join(d, "rho_final.npy"),
            "final_obj": os.path.join(d, "meshes", "final.obj"),
        })
        row["_iters_run"] = iters_run
        row["_converged"] = converged
        return row

    def _write_sample(self, t: Task, rho: np.ndarray, log: List[list], row: Dict[str, object]):
        os.makedirs(t.sample_dir, exist_ok=True)
        np.save(os.path.join(t.sample_dir, "condition.npy"), self._condition(t))
        np.save(os.path.join(t.sample_dir, "rho_warmstart_32.npy"),
                np.asarray(t.rho_warm, dtype=np.float32).reshape(self.r, self.r, self.r))
        np.save(os.path.join(t.sample_dir, "rho_final.npy"), rho.astype(np.float32))
        with open(os.path.join(t.sample_dir, "iter_log.csv"), "w", newline="") as f:
            wr = csv.writer(f)
            wr.writerow(ITER_LOG_HEADER)
            wr.writerows(log)


def write_dataset_log(path: str, rows: Sequence[Dict[str, object]], append: bool = False):
    """dataset_log.csv with the densify worker's columns (densify:35-82)."""
    new = not (append and os.path.exists(path))
    with open(path, "a" if append else "w", newline="") as f:
        wr = csv.DictWriter(f, fieldnames=FIELDNAMES, extrasaction="ignore")
        if new:
            wr.writeheader()
        for row in rows:
            wr.writerow(row)


def downsample_block_average(vol: np.ndarray, out_res: int) -> np.ndarray:
    """Mean over factor^3 blocks (prepare_stage2_training_data.py:194-204)."""
    arr = np.asarray(vol, dtype=np.float32)
    n = arr.shape[0]
    if n == out_res:
        return arr.astype(np.float32)
    if n % out_res:
        raise ValueError(f"input resolution {n} must be divisible by output resolution {out_res}")
    f = n // out_res
    return arr.reshape(out_res, f, out_res, f, out_res, f).mean(axis=(1, 3, 5)).astype(np.float32)


def write_stage2_npz(path: str, condition: np.ndarray, seed: int, source_index: int, rho_final: np.ndarray,
                     out_res: int = 32, dtype=np.float16):
    """One Stage-II training sample (prepare_stage2_training_data.py:330-354)."""
    np.savez_compressed(path, condition=np.asarray(condition, dtype=np.float32).reshape(7),
                        seed=np.array(int(seed), dtype=np.int64), source_index=np.array(int(source_index), dtype=np.int64),
                        rho_target_32=downsample_block_average(np.asarray(rho_final), out_res)[None].astype(dtype))


def stage2_sample_name(index: int, row: Dict[str, object]) -> str:
    """File name of a Stage-II sample (prepare_stage2_training_data.py:262-272)."""
    seed_text = str(row.get("seed", "")).strip()
    seed = int(float(seed_text)) if seed_text else int(index)
    parts = [f"{k}-{str(row.get(k, '')).strip()}" for k in ("vf_percent", "weight_idx", "w1_idx", "w2_idx", "w3_idx")
             if str(row.get(k, "")).strip()]
    prefix = "_".join(parts) if parts else f"sample-{index:06d}"
    return f"{prefix}_seed-{seed:010d}.npz"
