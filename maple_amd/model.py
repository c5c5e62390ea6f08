"""Drop-in for the reference's homogenization module ``IHD`` (model.py:12-159).

Same constructor (cfg with ``resolution`` and optionally ``GMG.smooth_iter``,
``GMG.max_cycle``; device), same attributes read by the optimizer driver
(``Ke_3d, Fe_3d, X0_3d, Ke_1d, Fe_1d, X0_1d``, main_IHD_v3.py:419-425,
640-645) and the same positional calls::

    CH    = ihd.TO_homo_3d(batch, Ke, F, X0, batch_size)   # (6, 6)
    kappa = ihd.TO_homo_1d(batch, Ke, F, X0, batch_size)   # (3, 3)

The solve runs in the gfx950 library (multigrid-preconditioned CG to
``rtol``), not the reference's single unconverged V-cycle.  C_H and kappa are
differentiable with respect to ``batch["rho"]``: the backward returns the
total derivative <dJ/dC_H, S_e> (converged sensitivity, main_IHD_v3.py:688)
for rho and None for F, which already depends on rho through
assemble_F_fast (model.py:98-159 semantics; F is re-assembled on the device
from rho, so the F argument only has to have the reference's shape).
Precision, rtol, warm start and the start history (each solve started from the
energy-optimal combination of the last solutions) are constructor keywords; the
positional signature is unchanged.  Results are float64.
"""
from __future__ import annotations

from types import SimpleNamespace
from typing import Optional

import torch
import torch.nn as nn

from .cells import CellBatch, hex8_matrices, max_levels


class _Homogenize(torch.autograd.Function):
    """C_H (kappa_H) of the first cell of a batch of `nb` cells, as the reference's
    TO_homo_3d / TO_homo_1d return it (homogenized_batch_3d(...)[0], model.py:128, 158);
    every cell of the batch is solved (the reference solves the whole batch too)."""

    @staticmethod
    def forward(ctx, rho, F, owner, physics, nb):
        ctx.owner = owner
        ctx.physics = physics
        ctx.nb = nb
        rho64 = rho.detach().reshape(-1).to(torch.float64).contiguous()
        out = owner._solve(rho64, physics, nb)
        ctx.gen = owner.batch(nb).generation[physics]
        ctx.save_for_backward(rho64)
        ctx.rho_dtype = rho.dtype
        ctx.rho_shape = rho.shape
        return out

    @staticmethod
    def backward(ctx, grad):
        (rho64,) = ctx.saved_tensors
        owner, physics, nb = ctx.owner, ctx.physics, ctx.nb
        cells = owner.batch(nb)
        if cells.generation[physics] != ctx.gen:
            owner._solve(rho64, physics, nb)      # the context moved on: re-solve this rho
        g = grad.detach().to(torch.float64).contiguous()
        m = 6 if physics == "elastic" else 3
        G = torch.zeros(nb, m, m, dtype=torch.float64, device=g.device)
        G[0] = g.reshape(m, m)                    # only the first cell's C_H is returned
        if physics == "elastic":
            d = cells.sensitivity(G6=G)
        else:
            d = cells.sensitivity(G3=G)
        return d.reshape(ctx.rho_shape).to(ctx.rho_dtype), None, None, None, None


class IHD(nn.Module):
    def __init__(self, cfg, device, *, rtol: float = 1e-6, maxit: int = 1000, nlevels: Optional[int] = None,
                 precision: str = "fp64", warm_start: bool = False, omega: Optional[float] = None, nsmooth: int = 2,
                 E0: float = 1.0, nu: float = 0.3, kappa0: float = 1.0, islands: bool = False,
                 start_history: int = 0):
        super().__init__()
        self.cfg = cfg
        self.r = self.resolution = int(cfg.resolution)
        self.device = torch.device(device)
        Ke24, Fe24, X024, Ke8, Fe8, X08 = hex8_matrices(self.r, E0, nu, kappa0)
        lam = nu / (1 + nu) / (1 - 2 * nu) * E0
        mu = E0 / (2 * (1 + nu))
        C = torch.zeros(6, 6, dtype=torch.float64)
        C[:3, :3] = lam
        C[0, 0] = C[1, 1] = C[2, 2] = lam + 2 * mu
        C[3, 3] = C[4, 4] = C[5, 5] = mu
        self.C = C.to(self.device)
        t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=self.device)  # noqa: E731
        self.Ke_3d, self.Fe_3d, self.X0_3d = t(Ke24), t(Fe24), t(X024)
        self.Ke_1d, self.Fe_1d, self.X0_1d = t(Ke8), t(Fe8), t(X08)
        if nlevels is None:
            gmg = getattr(cfg, "GMG", None)
            want = len(getattr(gmg, "smooth_iter", [])) if gmg is not None else 0
            nlevels = min(max(want, 2), max_levels(self.r, 2)) if want else max_levels(self.r)
        self.nlevels = int(nlevels)
        self.rtol = float(rtol)
        self.maxit = int(maxit)
        self.cells = CellBatch([self.r], nlevels=self.nlevels, device=self.device, precision=precision,
                               omega=omega, nsmooth=nsmooth, warm_start=warm_start, islands=islands)
        self._precision, self._omega, self._nsmooth = precision, omega, nsmooth
        self._warm, self._islands = warm_start, islands
        self._history = int(start_history)
        if self._history:
            self.cells.set_start_history(self._history)
        self._batches = {}
        self._custom = None
        if (E0, nu, kappa0) != (1.0, 0.3, 1.0):
            self._custom = (Ke24, Fe24, X024, Ke8, Fe8, X08)
            self.cells.set_element_matrices(0, *self._custom)
        self.last_iters = None

    def batch(self, nb: int) -> CellBatch:
        """The context of `nb` cells of this resolution (made on first use, then kept)."""
        nb = int(nb)
        if nb == 1:
            return self.cells
        cb = self._batches.get(nb)
        if cb is None:
            cb = CellBatch([self.r] * nb, nlevels=self.nlevels, device=self.device, precision=self._precision,
                           omega=self._omega, nsmooth=self._nsmooth, warm_start=self._warm, islands=self._islands)
            for i in range(nb):
                if self._custom is not None:
                    cb.set_element_matrices(i, *self._custom)
            if self._history:
                cb.set_start_history(self._history)
            self._batches[nb] = cb
        return cb

    def _solve(self, rho64: torch.Tensor, physics: str, nb: int = 1) -> torch.Tensor:
        CH, kap, it = self.batch(nb).homogenize(rho64, (physics,), rtol=self.rtol, maxit=self.maxit)
        self.last_iters = it
        return CH[0] if physics == "elastic" else kap[0]

    def _check_batch(self, batch, F, ncase, batch_size):
        nb = int(batch_size)
        if nb < 1:
            raise ValueError("batch_size must be >= 1")
        rho = batch["rho"]
        if rho.numel() != nb * self.r ** 3:
            raise ValueError(f"batch['rho'] has {rho.numel()} values, expected {nb * self.r ** 3} "
                             f"({nb} cells of {self.r}^3)")
        if F is not None and F.numel() != nb * self.r ** 3 * ncase:
            raise ValueError(f"F has {F.numel()} values, expected {nb * self.r ** 3 * ncase}")
        if rho.device != self.device:
            raise ValueError(f"batch['rho'] is on {rho.device}, IHD is on {self.device}")
        return rho, nb

    def TO_homo_3d(self, batch, Ke, F, X0, batch_size):
        """C_H (6, 6) of the first cell of the batch (model.py:98-130)."""
        rho, nb = self._check_batch(batch, F, 18, batch_size)
        return _Homogenize.apply(rho, F, self, "elastic", nb)

    def TO_homo_1d(self, batch, Ke, F, X0, batch_size):
        """kappa_H (3, 3) of the first cell of the batch (model.py:131-159)."""
        rho, nb = self._check_batch(batch, F, 3, batch_size)
        return _Homogenize.apply(rho, F, self, "thermal", nb)


def make_cfg(resolution: int, smooth_iter=(2, 2, 3, 4), max_cycle: int = 1):
    """The SimpleNamespace cfg of the reference's CLI (main_IHD_v3.py:1031-1037)."""
    return SimpleNamespace(resolution=int(resolution),
                           GMG=SimpleNamespace(smooth_iter=list(smooth_iter), max_cycle=int(max_cycle)))
