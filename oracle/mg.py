"""Converged periodic cell solve: Galerkin multigrid-preconditioned CG (oracle, FP64).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The operator is the reference's (EBE_GMG.py:412-430): K = sum_e A_e^T rho_e Ke A_e
on the periodic grid.  The coarse operators are the reference's element-wise
Galerkin projection (``sparse_coarsen_ke``, EBE_GMG.py:226-337): a coarse
element matrix is sum_c P_c^T K_child(c) P_c over its 8 children with the
trilinear local interpolation (EBE_GMG.py:202-224).  The grid transfer is the
periodic trilinear prolongation with weights prod_d (1 - |rem_d/2 - off_d|)
(EBE_GMG.py:531-647).  The solver form is a converged PCG (symmetric
V-cycle with weighted Jacobi, exact coarsest solve with the translation
constraint, CG with rtol on ||r||/||f||); the reference itself runs a single
unconverged V-cycle, which is not a parity target.

``direct_solve`` (dense / sparse LU with the rigid translations pinned) is
the independent check of ``pcg`` on small grids.

Pinned by tests/test_oracle_mg.py: P is the adjoint pair of the restriction,
the element-wise Galerkin operator equals the assembled P^T K P, PCG equals
the direct solve, and the laminate closed forms.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .fe import trilinear_child_prolongation
from .grid import assemble_csr, connectivity, ebe_matvec, scatter_elements


def coarsen_dims(dims):
    return tuple(d // 2 for d in dims)


def child_index_grid(cdims):
    """For each coarse element (C-order), the ids of its 8 children, ordered by c = cx*4+cy*2+cz."""
    cx, cy, cz = cdims
    fx, fy, fz = 2 * cx, 2 * cy, 2 * cz
    X, Y, Z = np.meshgrid(np.arange(cx), np.arange(cy), np.arange(cz), indexing="ij")
    X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
    out = np.empty((X.size, 8), dtype=np.int64)
    for c in range(8):
        ox, oy, oz = (c >> 2) & 1, (c >> 1) & 1, c & 1
        out[:, c] = ((2 * X + ox) * fy + (2 * Y + oy)) * fz + (2 * Z + oz)
    return out


def galerkin_elements(fine_mats: np.ndarray, cdims, Pc: np.ndarray) -> np.ndarray:
    """Coarse element matrices sum_c Pc^T K_child Pc  (EBE_GMG.py:226-337)."""
    ch = child_index_grid(cdims)
    m = fine_mats.shape[1]
    out = np.zeros((ch.shape[0], m, m))
    for c in range(8):
        out += np.einsum("ai,eab,bj->eij", Pc[c], fine_mats[ch[:, c]], Pc[c], optimize=True)
    return out


def level1_elements(rho: np.ndarray, Ke: np.ndarray, cdims, Pc: np.ndarray) -> np.ndarray:
    """Coarse element matrices of the first coarse level directly from rho: sum_c rho_c G_c."""
    ch = child_index_grid(cdims)
    G = np.einsum("cai,ab,cbj->cij", Pc, Ke, Pc)
    return np.einsum("ec,cij->eij", rho[ch], G, optimize=True)


def prolongation(fdims, cdims, nd: int) -> sp.csr_matrix:
    """Periodic trilinear P (fine <- coarse), weights prod(1 - |rem/2 - off|) (EBE_GMG.py:531-647)."""
    fx, fy, fz = fdims
    cx, cy, cz = cdims
    X, Y, Z = np.meshgrid(np.arange(fx), np.arange(fy), np.arange(fz), indexing="ij")
    f = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    base = f // 2
    rem = f % 2
    rows, cols, vals = [], [], []
    fid = (f[:, 0] * fy + f[:, 1]) * fz + f[:, 2]
    for k in range(8):
        off = np.array([(k >> 2) & 1, (k >> 1) & 1, k & 1])
        t = (base + off) % np.array([cx, cy, cz])
        w = np.prod(1.0 - np.abs(0.5 * rem - off), axis=1)
        keep = w != 0.0
        cid = (t[:, 0] * cy + t[:, 1]) * cz + t[:, 2]
        for d in range(nd):
            rows.append(fid[keep] * nd + d)
            cols.append(cid[keep] * nd + d)
            vals.append(w[keep])
    nf = fx * fy * fz * nd
    nc = cx * cy * cz * nd
    P = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nf, nc))
    return P.tocsr()


def translation_basis(nnodes: int, nd: int) -> np.ndarray:
    Z = np.zeros((nnodes * nd, nd))
    for d in range(nd):
        Z[d::nd, d] = 1.0 / np.sqrt(nnodes)
    return Z


def chebyshev_weights(k, lo, rising=True):
    """Sweep weights of the Chebyshev smoother on [lo, 1]: the reciprocals of the k
    Chebyshev nodes of that interval, so that prod_i (1 - w_i t) is the degree-k
    Chebyshev polynomial of [lo, 1] normalised to 1 at t = 0 (rising or falling order)."""
    i = np.arange(k)
    nodes = 0.5 * (1.0 + lo) + 0.5 * (1.0 - lo) * np.cos((2.0 * i + 1.0) * np.pi / (2.0 * k))
    w = np.sort(1.0 / nodes)
    return w if rising else w[::-1]


class Hierarchy:
    """Multigrid hierarchy for one density field and one physics (nd = 3 elastic, 1 thermal).

    weights: optional (pre, post) sequences of per-sweep smoother weights (the
    Chebyshev schedule); by default every sweep takes omega."""

    def __init__(self, rho, Ke, dims, nlevels, omega=0.6, nsmooth=1, weights=None):
        self.nd = Ke.shape[0] // 8
        nd = self.nd
        self.rho = np.asarray(rho, dtype=np.float64).ravel()
        self.Ke = Ke
        self.omega = omega
        self.nsmooth = nsmooth
        self.wpre, self.wpost = weights if weights is not None else ([omega] * nsmooth, [omega] * nsmooth)
        self.dims = [tuple(dims)]
        for _ in range(nlevels - 1):
            self.dims.append(coarsen_dims(self.dims[-1]))
        Pc = trilinear_child_prolongation(nd)
        # level 0 diagonal
        ediag = np.tile(np.diag(Ke)[None, :], (self.rho.size, 1)) * self.rho[:, None]
        self.diag = [scatter_elements(ediag, dims, nd).ravel()]
        self.A = [None]
        self.P = []
        mats = None
        for l in range(1, nlevels):
            cd = self.dims[l]
            mats = level1_elements(self.rho, Ke, cd, Pc) if l == 1 else galerkin_elements(mats, cd, Pc)
            A = assemble_csr(mats, connectivity(cd), int(np.prod(cd)), nd)
            self.A.append(A)
            self.diag.append(A.diagonal())
            self.P.append(prolongation(self.dims[l - 1], cd, nd))
        # coarsest: dense with the translation constraint (K + gamma Z Z^T)
        if nlevels == 1:
            raise ValueError("need at least 2 levels")
        Ac = self.A[-1].toarray()
        nn = int(np.prod(self.dims[-1]))
        Zb = translation_basis(nn, nd)
        gamma = np.mean(np.diag(Ac))
        self.Zc = Zb
        self.coarse_chol = sla.cho_factor(Ac + gamma * Zb @ Zb.T)

    def apply(self, l, x):
        if l == 0:
            return ebe_matvec(x.reshape(-1, self.nd), self.rho, self.Ke, self.dims[0]).ravel()
        return self.A[l] @ x

    def vcycle(self, l, b):
        if l == len(self.dims) - 1:
            bb = b - self.Zc @ (self.Zc.T @ b)
            return sla.cho_solve(self.coarse_chol, bb)
        dinv = 1.0 / self.diag[l]
        x = self.wpre[0] * dinv * b
        for w in self.wpre[1:]:
            x = x + w * dinv * (b - self.apply(l, x))
        r = b - self.apply(l, x)
        e = self.vcycle(l + 1, self.P[l].T @ r)
        x = x + self.P[l] @ e
        for w in self.wpost:
            x = x + w * dinv * (b - self.apply(l, x))
        return x


ATOL_REL = 1e-14


def residual_target(bnorm, rtol, fscale):
    """Stop when ||r|| <= max(rtol ||f||, 1e-14 fscale).

    fscale = ||Fe||_F sqrt(E) max(rho) is the size of the right-hand side of
    a generic design; the absolute floor only matters for load cases whose
    right-hand side vanishes analytically (e.g. in-plane shear of a
    laminate), where ||f|| is round-off and a relative test cannot be met.
    """
    return max(rtol * bnorm, ATOL_REL * fscale)


def pcg(H: Hierarchy, f: np.ndarray, rtol=1e-6, maxit=500, x0=None, fscale=None):
    """Preconditioned CG on the singular consistent periodic system, per load case.

    f: (N, nd, ncase).  Returns u (N, nd, ncase) with zero mean per component,
    and the iteration count per case.
    """
    N = f.shape[0]
    nd = f.shape[1]
    nc = f.shape[2]
    if fscale is None:
        fscale = np.linalg.norm(f) / np.sqrt(nc)
    u = np.zeros_like(f) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    iters = []
    for j in range(nc):
        b = f[:, :, j].ravel()
        x = u[:, :, j].ravel().copy()
        bn = np.linalg.norm(b)
        tgt = residual_target(bn, rtol, fscale)
        r = b - H.apply(0, x)
        it = 0
        if np.linalg.norm(r) <= tgt:
            xr = x.reshape(N, nd)
            u[:, :, j] = xr - xr.mean(axis=0, keepdims=True)
            iters.append(0)
            continue
        z = H.vcycle(0, r)
        p = z.copy()
        rz = r @ z
        while np.linalg.norm(r) > tgt and it < maxit:
            q = H.apply(0, p)
            alpha = rz / (p @ q)
            x += alpha * p
            r -= alpha * q
            it += 1
            if np.linalg.norm(r) <= tgt:
                break
            z = H.vcycle(0, r)
            rz_new = r @ z
            p = z + (rz_new / rz) * p
            rz = rz_new
        xr = x.reshape(N, nd)
        u[:, :, j] = xr - xr.mean(axis=0, keepdims=True)
        iters.append(it)
    return u, iters


def direct_solve(rho, Ke, dims, f):
    """Sparse LU with node 0 pinned (removes the periodic translations); zero-mean result."""
    nd = Ke.shape[0] // 8
    conn = connectivity(dims)
    N = int(np.prod(dims))
    mats = Ke[None] * np.asarray(rho, dtype=np.float64).reshape(-1, 1, 1)
    K = assemble_csr(mats, conn, N, nd).tocsc()
    keep = np.arange(nd, N * nd)
    Kr = K[keep][:, keep]
    lu = spla.splu(Kr.tocsc())
    u = np.zeros_like(f)
    for j in range(f.shape[2]):
        x = np.zeros(N * nd)
        x[keep] = lu.solve(f[:, :, j].ravel()[keep])
        xr = x.reshape(N, nd)
        u[:, :, j] = xr - xr.mean(axis=0, keepdims=True)
    return u
