"""Pins oracle.fe (element matrices) against mathematics written independently here.

The independent restatements below use the full fourth-order isotropic
tensor C_ijkl = lam d_ij d_kl + mu (d_ik d_jl + d_il d_jk) and index
notation instead of the Voigt B-matrix, a 2-point tensor-product rule
instead of the 3-point one, and the literal corner coordinates of the
reference (_utils.py:62-64).
"""
import itertools

import numpy as np
import pytest

from oracle import fe

CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                    [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=float)
VOIGT = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]


def _grads_unit_cube(p):
    """Physical gradients of the 8 trilinear functions on [0,h]^3 at point p in [0,1]^3 (h=1)."""
    g = np.zeros((8, 3))
    for a, c in enumerate(CORNERS):
        f = [(p[d] if c[d] else 1.0 - p[d]) for d in range(3)]
        s = [(1.0 if c[d] else -1.0) for d in range(3)]
        g[a] = [s[0] * f[1] * f[2], f[0] * s[1] * f[2], f[0] * f[1] * s[2]]
    return g


def _tensor_ke(h, lam, mu):
    gp = (1.0 - 1.0 / np.sqrt(3.0)) / 2.0
    pts = [gp, 1.0 - gp]
    d = np.eye(3)
    Cijkl = (lam * np.einsum("ij,kl->ijkl", d, d)
             + mu * (np.einsum("ik,jl->ijkl", d, d) + np.einsum("il,jk->ijkl", d, d)))
    Ke = np.zeros((8, 3, 8, 3))
    Fe = np.zeros((8, 3, 6))
    for p in itertools.product(pts, pts, pts):
        g = _grads_unit_cube(np.array(p)) / h        # d/dx on the cube of side h
        w = (h ** 3) / 8.0
        # K[a,i,b,k] = int C_ijkl dN_a/dx_j dN_b/dx_l
        Ke += w * np.einsum("ijkl,aj,bl->aibk", Cijkl, g, g)
        for J, (m, n) in enumerate(VOIGT):
            eps = np.zeros((3, 3))
            if m == n:
                eps[m, m] = 1.0
            else:                       # engineering shear: gamma = 1
                eps[m, n] = eps[n, m] = 0.5
            sig = np.einsum("ijkl,kl->ij", Cijkl, eps)
            Fe[:, :, J] += w * g @ sig.T
    return Ke.reshape(24, 24), Fe.reshape(24, 6)


@pytest.mark.parametrize("r", [1, 4, 32])
def test_elastic_matches_tensor_form(r):
    E, nu = 1.0, 0.3
    C = fe.isotropic_elastic_tensor(E, nu)
    lam = nu * E / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    Ke, Fe, X0 = fe.hexahedron(r, C)
    Kt, Ft = _tensor_ke(1.0 / r, lam, mu)
    assert np.abs(Ke - Kt).max() <= 1e-14 * np.abs(Kt).max()
    assert np.abs(Fe - Ft).max() <= 1e-14 * np.abs(Ft).max()


def test_elastic_properties():
    C = fe.isotropic_elastic_tensor(1.0, 0.3)
    Ke, Fe, X0 = fe.hexahedron(8, C)
    assert np.allclose(Ke, Ke.T, atol=1e-16)
    w = np.linalg.eigvalsh(Ke)
    assert w.min() > -1e-14 * w.max()
    assert np.sum(w < 1e-10 * w.max()) == 6            # 3 translations + 3 rotations
    # Ke X0 = Fe on every DOF (X0 = linear field + rigid motion)
    assert np.abs(Ke @ X0 - Fe).max() < 1e-14 * np.abs(Fe).max()
    assert np.all(X0[list(fe.ELASTIC_FIXED_DOFS)] == 0.0)
    # patch test: the nodal interpolant of eps0 . x reproduces Fe
    h = 1.0 / 8
    ulin = np.zeros((24, 6))
    for J, (m, n) in enumerate(VOIGT):
        eps = np.zeros((3, 3))
        if m == n:
            eps[m, m] = 1.0
        else:
            eps[m, n] = eps[n, m] = 0.5
        ulin[:, J] = (CORNERS * h @ eps.T).reshape(-1)
    assert np.abs(Ke @ ulin - Fe).max() < 1e-14 * np.abs(Fe).max()
    # X0 - ulin is a rigid motion
    assert np.abs(Ke @ (X0 - ulin)).max() < 1e-14 * np.abs(Fe).max()


def test_scaling_with_h():
    C = fe.isotropic_elastic_tensor(1.0, 0.3)
    K1, F1, X1 = fe.hexahedron(1, C)
    K5, F5, X5 = fe.hexahedron(5, C)
    assert np.allclose(K5, K1 / 5, rtol=1e-13, atol=0)
    assert np.allclose(F5, F1 / 25, rtol=1e-13, atol=1e-18)
    assert np.abs(X5 - X1 / 5).max() < 1e-14


def test_scalar_closed_form():
    """Q1 Laplacian on a cube of side h: (h/12) * [4 diag, 0 edge, -1 face diagonal, -1 body diagonal]."""
    for r in (1, 3, 64):
        Ke, Fe, X0 = fe.hexahedron_scalar(r)
        h = 1.0 / r
        ref = np.zeros((8, 8))
        for a in range(8):
            for b in range(8):
                dist = int(np.abs(CORNERS[a] - CORNERS[b]).sum())
                ref[a, b] = {0: 4.0, 1: 0.0, 2: -1.0, 3: -1.0}[dist] * h / 12.0
        assert np.abs(Ke - ref).max() < 1e-15
        # Fe = int grad N dV: each corner gets sign * h^2 / 4 per direction
        sgn = 2.0 * CORNERS - 1.0
        assert np.abs(Fe - sgn * h * h / 4.0).max() < 1e-15
        assert np.abs(Ke @ X0 - Fe).max() < 1e-15
        # X0 is the linear field x_d, shifted so node 0 is 0
        assert np.abs(X0 - CORNERS * h).max() < 1e-14


def test_child_prolongation_partition_of_unity():
    P = fe.trilinear_child_prolongation(3)
    for c in range(8):
        assert np.allclose(P[c].sum(axis=1), 1.0)
    # child 0, corner 0 coincides with coarse corner 0; corner 6 is the coarse centre
    P1 = fe.trilinear_child_prolongation(1)
    assert P1[0][0, 0] == 1.0
    assert np.allclose(P1[0][6], 1.0 / 8.0)
