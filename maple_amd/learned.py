"""Learned warm start for the cell solves.

A 3-level periodic U-Net (channels 16-32-64, about 0.35 M parameters, BF16
weights and activations with FP32 accumulation; the kernels are in
csrc/unet.hip) maps a density feature on an n^3 grid to the 21 characteristic
fields (6 elastic cases x 3 components, 3 thermal cases) on the same grid.
Used inside the OC loop, its prediction is a *candidate* start of every cell
solve: per (cell, case) the solver starts from whichever of zero, the previous
iterate's fields and the prediction has the smallest residual, and every solve
still converges to the same rtol, so the network can only save iterations.

* ``UNet``: the network on the device (forward, loss/gradient, Adam step).
* ``density_feature``: rho -> the network input (rho^(1/3), 2x2x2-averaged for a
  cell twice the network's grid).
* ``fields_to_channels`` / ``channels_to_fields``: solver layout <-> network layout.
* ``generate_pairs``: (feature, fields) training pairs from batched OC runs of the
  solver itself.
* ``train``: minibatch Adam on those pairs, all on the GPU.
* ``LearnedStart``: the predictor the OC driver calls before each homogenization.

The reference has no learned start: its warm-start hook ``init_list``
(EBE_GMG.py:648-757) is never fed (model.py:120, 149).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch
import torch.nn.functional as F

from . import _lib as L
from .cells import CellBatch

NCH = 21
# name, cin, cout, kernel size (storage order of the parameter vector)
LAYERS = [("e1a", 1, 16, 3), ("e1b", 16, 16, 3), ("e2a", 16, 32, 3), ("e2b", 32, 32, 3), ("b1", 32, 64, 3),
          ("b2", 64, 64, 3), ("d2a", 96, 32, 3), ("d2b", 32, 32, 3), ("d1a", 48, 16, 3), ("d1b", 16, 16, 3),
          ("head", 16, NCH, 1)]
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "warmstart_unet32.npz")


def num_params() -> int:
    return sum(co * k ** 3 * ci + co for _, ci, co, k in LAYERS)


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights (std sqrt(2 / fan_in); head sqrt(1 / fan_in)), zero biases, numpy default_rng(seed)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, ci, co, k in LAYERS:
        fan_in = ci * k ** 3
        std = math.sqrt((1.0 if name == "head" else 2.0) / fan_in)
        parts.append(rng.standard_normal(co * k ** 3 * ci) * std)
        parts.append(np.zeros(co))
    return np.concatenate(parts).astype(np.float32)


class UNet:
    """The warm-start network on one device for n^3 grids and batches up to max_batch."""

    def __init__(self, n: int = 32, max_batch: int = 16, device=None, theta: Optional[np.ndarray] = None,
                 seed: int = 0, scales: Optional[np.ndarray] = None):
        self.lib = L.load()
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("UNet needs a CUDA (HIP) device; there is no CPU path")
        self.n, self.max_batch = int(n), int(max_batch)
        h = C.c_void_p()
        L.check(self.lib.maple_unet_create(self.device.index or 0, self.n, self.max_batch, C.byref(h)))
        self._h = h
        self.nparam = int(self.lib.maple_unet_num_params())
        self.set_params(init_params(seed) if theta is None else theta)
        self.scales = np.ones(NCH, dtype=np.float32) if scales is None else np.asarray(scales, dtype=np.float32)
        self.weighted = False

    def close(self):
        if getattr(self, "_h", None):
            self.lib.maple_unet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def set_params(self, theta: np.ndarray):
        th = np.ascontiguousarray(np.asarray(theta, dtype=np.float32).ravel())
        if th.size != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters, got {th.size}")
        L.check(self.lib.maple_unet_set_params(self._h, th.ctypes.data_as(C.POINTER(C.c_float))))

    def set_loss(self, weighted: bool):
        """Voxel weight f^3 + 1e-3 in the training loss (f: input feature) or none."""
        L.check(self.lib.maple_unet_set_loss(self._h, int(bool(weighted))))
        self.weighted = bool(weighted)

    def params(self) -> np.ndarray:
        th = np.empty(self.nparam, dtype=np.float32)
        L.check(self.lib.maple_unet_get_params(self._h, th.ctypes.data_as(C.POINTER(C.c_float))))
        return th

    def _feat(self, feat: torch.Tensor) -> torch.Tensor:
        B = feat.shape[0]
        if feat.numel() != B * self.n ** 3 or B > self.max_batch or feat.device != self.device:
            raise ValueError(f"feature must be (B <= {self.max_batch}, {self.n}^3) on {self.device}")
        return feat.reshape(B, -1).to(torch.float32).contiguous()

    def _target(self, target: torch.Tensor, B: int) -> torch.Tensor:
        if target.dtype != torch.bfloat16 or target.numel() != B * self.n ** 3 * NCH or target.device != self.device:
            raise ValueError(f"target must be bfloat16 ({B}, {self.n}^3, {NCH}) on {self.device}")
        return target.contiguous()

    def forward(self, feat: torch.Tensor) -> torch.Tensor:
        """Raw network output (B, n^3, 21) in normalised units."""
        f = self._feat(feat)
        out = torch.empty(f.shape[0], self.n ** 3, NCH, dtype=torch.float32, device=self.device)
        L.check(self.lib.maple_unet_forward(self._h, C.c_void_p(f.data_ptr()), f.shape[0], C.c_void_p(out.data_ptr()),
                                            self._stream()))
        return out

    def loss_grad(self, feat: torch.Tensor, target: torch.Tensor):
        f = self._feat(feat)
        t = self._target(target, f.shape[0])
        g = torch.empty(self.nparam, dtype=torch.float32, device=self.device)
        loss = C.c_double(0.0)
        L.check(self.lib.maple_unet_loss_grad(self._h, C.c_void_p(f.data_ptr()), C.c_void_p(t.data_ptr()), f.shape[0],
                                              C.byref(loss), C.c_void_p(g.data_ptr()), self._stream()))
        return float(loss.value), g

    def train_step(self, feat: torch.Tensor, target: torch.Tensor, lr: float) -> float:
        f = self._feat(feat)
        t = self._target(target, f.shape[0])
        loss = C.c_double(0.0)
        L.check(self.lib.maple_unet_train_step(self._h, C.c_void_p(f.data_ptr()), C.c_void_p(t.data_ptr()),
                                               f.shape[0], float(lr), C.byref(loss), self._stream()))
        return float(loss.value)

    def save(self, path: str, **meta):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, theta=self.params(), scales=self.scales, n=np.int64(self.n), weighted=np.int64(self.weighted),
                 **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS, max_batch: int = 16, device=None) -> "UNet":
        z = np.load(path, allow_pickle=False)
        return cls(n=int(z["n"]), max_batch=max_batch, device=device, theta=z["theta"], scales=z["scales"])


# ---- layouts -----------------------------------------------------------------

def density_feature(rho_cell: torch.Tensor, r: int, n: int) -> torch.Tensor:
    """rho^(1/3) of one r^3 cell (about the design variable x under SIMP p = 3),
    averaged over (r/n)^3 blocks down to the network's n^3 grid."""
    x = rho_cell.reshape(1, 1, r, r, r).clamp_min(0).to(torch.float32) ** (1.0 / 3.0)
    if r != n:
        x = F.avg_pool3d(x, r // n)
    return x.reshape(-1)


def fields_to_channels(u_el: torch.Tensor, u_th: torch.Tensor) -> torch.Tensor:
    """Solver fields of one cell ((6, 3, N), (3, 1, N)) -> (N, 21) de-meaned channels."""
    N = u_el.shape[-1]
    ch = torch.cat([u_el.reshape(18, N), u_th.reshape(3, N)], dim=0)
    ch = ch - ch.mean(dim=1, keepdim=True)
    return ch.T.contiguous()


def channels_to_fields(ch: torch.Tensor):
    """(N, 21) -> ((6, 3, N), (3, 1, N)) float64."""
    N = ch.shape[0]
    t = ch.T.to(torch.float64)
    return t[:18].reshape(6, 3, N), t[18:].reshape(3, 1, N)


# ---- training data -----------------------------------------------------------------

def _objective(rng) -> Dict[str, float]:
    """Config-1 objective (bulk modulus) for half of the cells, weights from the 5 % simplex grid otherwise."""
    if rng.uniform() < 0.5:
        return {"bulk": 1.0}
    from .optim import simplex_grid
    grid = simplex_grid(0.05)
    wE, wG, wK = grid[int(rng.integers(len(grid)))]
    w = {k: v for k, v in (("stiffness", wE), ("shear", wG), ("thermal", wK)) if v > 0}
    return w or {"bulk": 1.0}


def generate_pairs(n_cells: int, iters: int = 50, res: int = 32, seed: int = 0, batch_cells: int = 64,
                   device=None, log=print):
    """(feature, target) pairs from batched OC runs of the solver at res^3.

    Cell k of the run: numpy default_rng(seed + k) draws vf ~ U[0.01, 0.40], the
    objective, and the initial design: the config-1 sphere void (0.4 vf inside a
    0.3 L sphere, vf outside, + U(-0.02, 0.02)) for even k, vf + U(-0.05, 0.05) for
    odd k, clipped to [0.001, 1].  Every OC iterate of every cell gives one pair:
    the density feature and the converged fields of both physics (rtol 1e-6) as
    de-meaned channels.  Returns (feat float32 (P, res^3), target bfloat16
    (P, res^3, 21) in physical units), P = n_cells * iters.
    """
    from .optim import OCDriver, x0_sphere_void
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    N = res ** 3
    P = n_cells * iters
    feats = torch.empty(P, N, dtype=torch.float32, device=dev)
    targets = torch.empty(P, N, NCH, dtype=torch.bfloat16, device=dev)
    p = 0
    for b0 in range(0, n_cells, batch_cells):
        ks = list(range(b0, min(n_cells, b0 + batch_cells)))
        objs, vfs, x0s = [], [], []
        for k in ks:
            rng = np.random.default_rng(seed + k)
            vf = float(rng.uniform(0.01, 0.40))
            objs.append(_objective(rng))
            vfs.append(vf)
            if k % 2 == 0:
                x0 = x0_sphere_void(res, seed=seed + k, vf=vf, void=0.4 * vf)
            else:
                x0 = np.clip(vf + rng.uniform(-0.05, 0.05, N), 0.001, 1.0)
            x0s.append(x0)
        cells = CellBatch([res] * len(ks), nlevels=3, device=dev, warm_start=True)
        drv = OCDriver(cells, objs, vfs, move=0.2, rtol=1e-6, maxit=5000, physics=("elastic", "thermal"))
        x = torch.from_numpy(np.concatenate(x0s)).to(dev)
        t0 = time.time()
        for it in range(iters):
            x_new, info = drv.step(x)
            rho = info["rho"]
            ue, ut = cells.fields("elastic"), cells.fields("thermal")
            for c in range(len(ks)):
                o0, o1 = int(cells.offsets[c]), int(cells.offsets[c + 1])
                feats[p] = density_feature(rho[o0:o1], res, res)
                targets[p] = fields_to_channels(ue[:, :, o0:o1], ut[:, :, o0:o1]).to(torch.bfloat16)
                p += 1
            x = x_new
        torch.cuda.synchronize(dev)
        log(f"pairs: cells {ks[0]}..{ks[-1]}, {iters} OC iterations, {time.time() - t0:.1f} s", flush=True)
        cells.close()
    return feats, targets


def channel_scales(targets: torch.Tensor) -> np.ndarray:
    """Per-channel RMS of de-meaned targets (the network predicts fields / scale)."""
    acc = torch.zeros(NCH, dtype=torch.float64, device=targets.device)
    for i in range(0, targets.shape[0], 256):
        acc += targets[i:i + 256].to(torch.float64).pow(2).sum(dim=(0, 1))
    ms = acc / (targets.shape[0] * targets.shape[1])
    return ms.sqrt().clamp_min(1e-12).cpu().numpy().astype(np.float32)


def train(net: UNet, feat: torch.Tensor, target: torch.Tensor, epochs: int = 8, batch: int = 16, lr: float = 2e-3,
          val_frac: float = 0.05, seed: int = 0, log=print):
    """Minibatch Adam with a cosine learning-rate schedule; the last val_frac of the
    pairs is held out.  target is in physical units; it is normalised with
    net.scales and rounded to bfloat16.  Returns the history of (epoch, train, val)."""
    P = feat.shape[0]
    nval = max(batch, int(P * val_frac))
    ntr = P - nval
    sc = torch.from_numpy(net.scales).to(target.device)
    tgt = torch.empty_like(target, dtype=torch.bfloat16)
    for i in range(0, P, 256):
        tgt[i:i + 256] = (target[i:i + 256].to(torch.float32) / sc).to(torch.bfloat16)
    g = torch.Generator(device="cpu").manual_seed(seed)
    steps = epochs * (ntr // batch)
    hist = []
    k = 0
    for ep in range(epochs):
        perm = torch.randperm(ntr, generator=g)
        t0 = time.time()
        tl = 0.0
        for i in range(ntr // batch):
            idx = perm[i * batch:(i + 1) * batch].to(feat.device)
            lr_k = lr * 0.5 * (1 + math.cos(math.pi * k / steps))
            tl += net.train_step(feat[idx], tgt[idx], lr_k)
            k += 1
        vl = evaluate(net, feat[ntr:], tgt[ntr:], batch)
        hist.append((ep, tl / max(1, ntr // batch), vl))
        log(f"epoch {ep}: train loss {hist[-1][1]:.4e}  val loss {vl:.4e}  ({time.time() - t0:.1f} s)", flush=True)
    return hist


def evaluate(net: UNet, feat: torch.Tensor, tgt_bf16: torch.Tensor, batch: int = 16) -> float:
    tot, cnt = 0.0, 0
    for i in range(0, feat.shape[0] - batch + 1, batch):
        l, _ = net.loss_grad(feat[i:i + batch], tgt_bf16[i:i + batch])
        tot += l
        cnt += 1
    return tot / max(cnt, 1)


# ---- use in the OC loop -------------------------------------------------------------

class LearnedStart:
    """Supplies the network's fields as the start candidate of every solve of a CellBatch.

    Cells of the network's resolution get the prediction on their own grid; cells
    of twice that resolution get it on their level-1 grid, prolongated to level 0
    by the multigrid's periodic trilinear P.  Other cells get no candidate (zero).

    mode "direct": the candidate is the prediction itself.  mode "delta": from the
    second call on, the candidate is the previous solve's fields plus the change
    the network predicts between the previous density and this one,
    u_prev + N(rho_t) - N(rho_{t-1}); the network's systematic error is mostly the
    same for two neighbouring densities and cancels in the difference, while the
    fields carry everything it cannot resolve.  The solve still starts from the best
    of zero, the previous fields and this candidate.
    """

    def __init__(self, cells: CellBatch, net: UNet, mode: str = "direct"):
        if mode not in ("direct", "delta"):
            raise ValueError("mode must be 'direct' or 'delta'")
        self.cells, self.net = cells, net
        self.mode = mode
        self.prev = None
        n = net.n
        self.l0 = [i for i, d in enumerate(cells.dims) if d == (n, n, n)]
        self.l1 = [i for i, d in enumerate(cells.dims) if d == (2 * n, 2 * n, 2 * n)]
        if self.l1 and cells.nlevels < 2:
            raise ValueError("cells twice the network's grid need a level-1 grid")
        self.scales = torch.from_numpy(net.scales).to(cells.device)
        # level-1 node offsets of every cell
        sz1 = [int(np.prod([v // 2 for v in d])) for d in cells.dims]
        self.off1 = np.concatenate([[0], np.cumsum(sz1)]).astype(np.int64)

    def _predict(self, feats: List[torch.Tensor]) -> torch.Tensor:
        outs = []
        for i in range(0, len(feats), self.net.max_batch):
            outs.append(self.net.forward(torch.stack(feats[i:i + self.net.max_batch])))
        return torch.cat(outs) * self.scales

    @torch.no_grad()
    def __call__(self, rho: torch.Tensor, physics: Sequence[str] = ("elastic", "thermal")):
        if not self.l0 and not self.l1:
            return
        el0, th0 = self.predict(rho)
        prev, self.prev = self.prev, (el0, th0)
        if self.mode == "delta" and prev is not None:
            if "elastic" in physics:
                self.cells.set_start("elastic", self.cells.fields("elastic") + (el0 - prev[0]))
            if "thermal" in physics:
                self.cells.set_start("thermal", self.cells.fields("thermal") + (th0 - prev[1]))
            return
        if "elastic" in physics:
            self.cells.set_start("elastic", el0)
        if "thermal" in physics:
            self.cells.set_start("thermal", th0)

    @torch.no_grad()
    def predict(self, rho: torch.Tensor):
        """Predicted level-0 fields ((6, 3, Etot), (3, 1, Etot)); zero for cells the network does not cover."""
        cells, n = self.cells, self.net.n
        dev = cells.device
        el0 = torch.zeros(6, 3, cells.nelem, dtype=torch.float64, device=dev)
        th0 = torch.zeros(3, 1, cells.nelem, dtype=torch.float64, device=dev)
        idx = self.l0 + self.l1
        if not idx:
            return el0, th0
        feats = []
        for ci in idx:
            r = cells.dims[ci][0]
            o0, o1 = int(cells.offsets[ci]), int(cells.offsets[ci + 1])
            feats.append(density_feature(rho[o0:o1], r, n))
        pred = self._predict(feats)
        if self.l1:
            N1 = int(self.off1[-1])
            el1 = torch.zeros(6, 3, N1, dtype=torch.float64, device=dev)
            th1 = torch.zeros(3, 1, N1, dtype=torch.float64, device=dev)
        for k, ci in enumerate(idx):
            e, t = channels_to_fields(pred[k])
            if ci in self.l0:
                o0, o1 = int(cells.offsets[ci]), int(cells.offsets[ci + 1])
                el0[:, :, o0:o1] = e
                th0[:, :, o0:o1] = t
            else:
                o0, o1 = int(self.off1[ci]), int(self.off1[ci + 1])
                el1[:, :, o0:o1] = e
                th1[:, :, o0:o1] = t
        if self.l1:
            el0 += cells.prolong("elastic", 1, el1)
            th0 += cells.prolong("thermal", 1, th1)
        return el0, th0
