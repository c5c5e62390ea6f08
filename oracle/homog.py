"""Energy-form homogenization and closed-form element sensitivities (oracle, FP64).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``element_energies``: S_e = L_e^T Ke L_e with L_e = X0 - u_e
  (model.py:39-43 for C_H, model.py:77-93 for kappa); the unique 6x6 or 3x3
  per-element matrices.
* ``homogenize``: C_H = sum_e rho_e S_e (volume 1, model.py:34, 43, 93).
* dC_H/drho_e = S_e is the converged limit of the reference's autograd
  sensitivity (main_IHD_v3.py:688): C_H is the minimum over periodic u of
  the element-energy sum, so the envelope theorem applies.

Pinned by tests/test_oracle_homog.py: homogeneous cell gives C (and kappa=I),
the z-laminate closed forms (Backus / harmonic-arithmetic means) at a
1e9 contrast, Euler homogeneity sum_e rho_e S_e = C_H, central finite
differences of C_H with respect to rho_e, translation invariance, and
Reuss <= C_H <= Voigt.
"""
from __future__ import annotations

import numpy as np

from . import fe
from .grid import assemble_rhs, gather_elements
from .mg import Hierarchy, pcg


def element_energies(u: np.ndarray, Ke: np.ndarray, X0: np.ndarray, dims) -> np.ndarray:
    """S_e (E, nc, nc) for u (N, nd, nc)."""
    nc = u.shape[2]
    ue = np.stack([gather_elements(u[:, :, j], dims) for j in range(nc)], axis=-1)   # (E, 8nd, nc)
    L = X0[None] - ue
    return np.einsum("eai,ab,ebj->eij", L, Ke, L, optimize=True)


def homogenize(rho, S):
    return np.einsum("e,eij->ij", np.asarray(rho).ravel(), S)


class CellSolver:
    """Both physics for one cell: element matrices and converged solves."""

    def __init__(self, dims, E0=1.0, nu=0.3, kappa0=1.0, nlevels=3, rtol=1e-10,
                 omega=0.6, nsmooth=2, maxit=1000):
        self.dims = tuple(dims)
        r = self.dims[0]
        if any(d != r for d in self.dims):
            raise ValueError("the reference discretization is cubic (h = 1/r)")
        self.C = fe.isotropic_elastic_tensor(E0, nu)
        self.Ke3, self.Fe3, self.X03 = fe.hexahedron(r, self.C)
        Ke1, Fe1, X01 = fe.hexahedron_scalar(r, kappa0)
        self.Ke1, self.Fe1, self.X01 = Ke1, Fe1, X01
        self.nlevels = nlevels
        self.rtol = rtol
        self.omega = omega
        self.nsmooth = nsmooth
        self.maxit = maxit
        self.iters = {}

    def solve(self, rho, physics: str, rtol=None, x0=None):
        """x0 (N, nd, ncase): CG start.  The result is whatever meets this solver's own
        residual test ||f - K u|| <= rtol ||f|| on its own operator, from any start; a
        start only changes how many iterations that takes."""
        rho = np.asarray(rho, dtype=np.float64).ravel()
        Ke, Fe = (self.Ke3, self.Fe3) if physics == "elastic" else (self.Ke1, self.Fe1)
        H = Hierarchy(rho, Ke, self.dims, self.nlevels, omega=self.omega, nsmooth=self.nsmooth)
        F = assemble_rhs(rho, Fe, self.dims)
        fscale = np.linalg.norm(Fe) * np.sqrt(rho.size) * np.max(np.abs(rho))
        u, it = pcg(H, F, rtol=self.rtol if rtol is None else rtol, maxit=self.maxit, x0=x0,
                    fscale=fscale)
        self.iters[physics] = it
        return u

    def elastic(self, rho, rtol=None, x0=None):
        """Returns (C_H, S_e, u)."""
        u = self.solve(rho, "elastic", rtol, x0)
        S = element_energies(u, self.Ke3, self.X03, self.dims)
        return homogenize(rho, S), S, u

    def thermal(self, rho, rtol=None, x0=None):
        """Returns (kappa_H, S_e, u)."""
        u = self.solve(rho, "thermal", rtol, x0)
        S = element_energies(u, self.Ke1, self.X01, self.dims)
        return homogenize(rho, S), S, u
