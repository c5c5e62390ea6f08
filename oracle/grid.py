"""Periodic hex8 voxel grid, RHS assembly and the element-by-element product (oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Layout (main_IHD_v3.py:175-197): element and node ids are both
``(x*ny + y)*nz + z`` (z contiguous); element (x,y,z) owns the nodes
(x,y,z)+HEX8[k] with a periodic wrap.  The DOF index is ``ndof*node + comp``
(EBE_GMG.py:796-798); vectors are stored (N, ndof) as in the reference.

* ``connectivity``      -> node_index (E, 8)      (main_IHD_v3.py:188-197)
* ``assemble_rhs``      -> F (N, ndof, ncase)     (_utils.py:453-507)
* ``ebe_matvec``        -> sum_e A_e^T (rho_e Ke) A_e u  (EBE_GMG.py:412-430)
* ``assemble_csr``      -> the same operator as a scipy CSR matrix

Pinned by tests/test_oracle_grid.py: connectivity against a brute-force
loop, CSR and matrix-free products agree, the operator annihilates rigid
translations, F sums to zero per component, and K applied to the
interpolated linear field equals F for a homogeneous cell.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .fe import HEX8


def connectivity(dims) -> np.ndarray:
    nx, ny, nz = dims
    X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
    conn = np.empty((X.size, 8), dtype=np.int64)
    for k in range(8):
        ox, oy, oz = HEX8[k]
        conn[:, k] = (((X + ox) % nx) * ny + (Y + oy) % ny) * nz + (Z + oz) % nz
    return conn


def _roll_to_node(field, k, dims):
    """Move per-element values that belong to corner k onto the node grid."""
    ox, oy, oz = HEX8[k]
    return np.roll(field, shift=(ox, oy, oz), axis=(0, 1, 2))


def _roll_from_node(field, k, dims):
    ox, oy, oz = HEX8[k]
    return np.roll(field, shift=(-ox, -oy, -oz), axis=(0, 1, 2))


def assemble_rhs(rho: np.ndarray, Fe: np.ndarray, dims) -> np.ndarray:
    """F[n, c, j] = sum_{e ∋ n} rho_e Fe[ndof*k(n,e)+c, j]  (_utils.py:480-507)."""
    nx, ny, nz = dims
    nd = Fe.shape[0] // 8
    nc = Fe.shape[1]
    rg = rho.reshape(nx, ny, nz)
    F = np.zeros((nx, ny, nz, nd, nc))
    for k in range(8):
        blk = Fe[nd * k: nd * (k + 1), :]                  # (nd, nc)
        F += _roll_to_node(rg[..., None, None] * blk[None, None, None], k, dims)
    return F.reshape(-1, nd, nc)


def gather_elements(u: np.ndarray, dims) -> np.ndarray:
    """u (N, nd) -> u_e (E, 8*nd) in element DOF order."""
    nx, ny, nz = dims
    nd = u.shape[1]
    ug = u.reshape(nx, ny, nz, nd)
    parts = [_roll_from_node(ug, k, dims) for k in range(8)]
    return np.concatenate(parts, axis=-1).reshape(-1, 8 * nd)


def scatter_elements(fe: np.ndarray, dims, nd: int) -> np.ndarray:
    """f_e (E, 8*nd) -> v (N, nd) summed over the elements sharing each node."""
    nx, ny, nz = dims
    fg = fe.reshape(nx, ny, nz, 8 * nd)
    v = np.zeros((nx, ny, nz, nd))
    for k in range(8):
        v += _roll_to_node(fg[..., nd * k: nd * (k + 1)], k, dims)
    return v.reshape(-1, nd)


def ebe_matvec(u: np.ndarray, rho: np.ndarray, Ke: np.ndarray, dims) -> np.ndarray:
    """v = sum_e A_e^T (rho_e Ke) A_e u, u of shape (N, nd)."""
    nd = u.shape[1]
    ue = gather_elements(u, dims)
    fe = (ue @ Ke.T) * rho.reshape(-1, 1)
    return scatter_elements(fe, dims, nd)


def element_dofs(conn: np.ndarray, nd: int) -> np.ndarray:
    return (conn[:, :, None] * nd + np.arange(nd)[None, None, :]).reshape(conn.shape[0], -1)


def assemble_csr(elem_mats: np.ndarray, conn: np.ndarray, nnodes: int, nd: int) -> sp.csr_matrix:
    """Assemble per-element matrices (E, 8nd, 8nd) into a global CSR matrix."""
    dofs = element_dofs(conn, nd)
    m = dofs.shape[1]
    rows = np.repeat(dofs, m, axis=1).ravel()
    cols = np.tile(dofs, (1, m)).ravel()
    K = sp.coo_matrix((elem_mats.reshape(-1), (rows, cols)), shape=(nnodes * nd, nnodes * nd))
    return K.tocsr()
