"""Element matrices of the periodic voxel discretization (oracle, FP64).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows the reference's ``isotropic_elastic_tensor`` (_utils.py:39-48),
``hexahedron`` (_utils.py:49-109) and ``hexahedron_scalar``
(_utils.py:116-196):

* node order of the hex8 element: (0,0,0),(1,0,0),(1,1,0),(0,1,0),
  (0,0,1),(1,0,1),(1,1,1),(0,1,1)  (_utils.py:62-64);
* Voigt order [xx, yy, zz, xy, yz, xz] with engineering shear
  (B rows at _utils.py:91-100);
* 3x3x3 Gauss-Legendre quadrature on a cube of side h = 1/r (_utils.py:51-58);
* X0 solves Ke[f,f] X0[f] = Fe[f] with DOFs {0,1,2,4,5,11} fixed
  (_utils.py:105-108); scalar X0 anchors node 0 (_utils.py:189-195).

Pinned by tests/test_oracle_fe.py: symmetry and PSD, nullspace dimension
(6 rigid modes / 1 constant), Ke X0 = Fe, the scaling Ke ~ h, Fe ~ h^2,
agreement with an independently written 2-point tensor-product rule,
the constant-strain patch test, and (through the whole chain) the
laminate closed forms in tests/test_oracle_homog.py.
"""
from __future__ import annotations

import numpy as np

HEX8 = np.array(
    [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.int64)

# Signs of the local coordinate of each corner, in [-1, 1] form.
_SIGNS = 2.0 * HEX8 - 1.0

ELASTIC_FIXED_DOFS = (0, 1, 2, 4, 5, 11)


def isotropic_elastic_tensor(E: float = 1.0, nu: float = 0.3) -> np.ndarray:
    """6x6 isotropic stiffness in Voigt order [xx,yy,zz,xy,yz,xz] (_utils.py:39-48)."""
    lam = nu / (1.0 + nu) / (1.0 - 2.0 * nu) * E
    mu = 1.0 / (2.0 * (1.0 + nu)) * E
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[0, 0] = C[1, 1] = C[2, 2] = lam + 2.0 * mu
    C[3, 3] = C[4, 4] = C[5, 5] = mu
    return C


def _gauss3():
    p = np.sqrt(3.0 / 5.0)
    return np.array([-p, 0.0, p]), np.array([5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0])


def _grad_ref(xi, eta, zeta):
    """d N_a / d(xi,eta,zeta) for the 8 trilinear shape functions, shape (3,8)."""
    sx, sy, sz = _SIGNS[:, 0], _SIGNS[:, 1], _SIGNS[:, 2]
    fx = 1.0 + sx * xi
    fy = 1.0 + sy * eta
    fz = 1.0 + sz * zeta
    return np.stack([sx * fy * fz, fx * sy * fz, fx * fy * sz]) / 8.0


def _strain_matrix(g):
    """B (6x24) from physical gradients g (3,8), rows [xx,yy,zz,xy,yz,xz]."""
    B = np.zeros((6, 24))
    for a in range(8):
        gx, gy, gz = g[:, a]
        c = 3 * a
        B[0, c + 0] = gx
        B[1, c + 1] = gy
        B[2, c + 2] = gz
        B[3, c + 0] = gy
        B[3, c + 1] = gx
        B[4, c + 1] = gz
        B[4, c + 2] = gy
        B[5, c + 0] = gz
        B[5, c + 2] = gx
    return B


def hexahedron(r: int, C: np.ndarray):
    """Elastic Ke (24,24), Fe (24,6), X0 (24,6) for a cube of side 1/r (_utils.py:49-109)."""
    h = 1.0 / r
    half = h / 2.0
    pts, wts = _gauss3()
    Ke = np.zeros((24, 24))
    Fe = np.zeros((24, 6))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                g = _grad_ref(pts[i], pts[j], pts[k]) / half   # Jacobian = half * I
                B = _strain_matrix(g)
                w = wts[i] * wts[j] * wts[k] * half ** 3
                Ke += w * B.T @ C @ B
                Fe += w * B.T @ C
    free = np.ones(24, dtype=bool)
    free[list(ELASTIC_FIXED_DOFS)] = False
    X0 = np.zeros((24, 6))
    X0[free] = np.linalg.solve(Ke[np.ix_(free, free)], Fe[free])
    return Ke, Fe, X0


def hexahedron_scalar(r: int, kappa: float = 1.0):
    """Thermal Ke (8,8), Fe (8,3), X0 (8,3) for a cube of side 1/r (_utils.py:116-196)."""
    h = 1.0 / r
    half = h / 2.0
    pts, wts = _gauss3()
    Ke = np.zeros((8, 8))
    Fe = np.zeros((8, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                g = _grad_ref(pts[i], pts[j], pts[k]) / half
                w = wts[i] * wts[j] * wts[k] * half ** 3 * kappa
                Ke += w * g.T @ g
                Fe += w * g.T
    X0 = np.zeros((8, 3))
    X0[1:] = np.linalg.solve(Ke[1:, 1:], Fe[1:])
    return Ke, Fe, X0


def trilinear_child_prolongation(ndof: int):
    """Local interpolation from a coarse element's 8 corners to child c's 8 corners.

    Returns P (8, 8*ndof, 8*ndof): P[c] maps the coarse element's DOFs to the
    DOFs of its child c, where child c sits at offset ((c>>2)&1, (c>>1)&1, c&1)
    in the 2x2x2 split (EBE_GMG.py:261, 284-294) and interpolation is trilinear
    (EBE_GMG.py:202-224).
    """
    P = np.zeros((8, 8 * ndof, 8 * ndof))
    for c in range(8):
        off = np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1])
        for a in range(8):
            pos = (off + HEX8[a]) / 2.0          # child corner in coarse [0,1]^3
            for b in range(8):
                wb = np.prod(1.0 - np.abs(pos - HEX8[b]))
                for d in range(ndof):
                    P[c, a * ndof + d, b * ndof + d] = wb
    return P
