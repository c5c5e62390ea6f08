"""Objective terms and their gradients dJ/dC_H, dJ/dkappa (oracle, FP64).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Terms (objective.py:11-41): with A = (C_H + C_H^T)/2 + 1e-9 I and S = A^-1,
  stiffness = mean_{i<3} 1/S_ii,  shear = mean_{i in 3..5} 1/S_ii,
  thermal = trace(kappa).
Extension (not in the reference, named by the config-1 workload):
  bulk = (1/9) sum_{i,j<3} C_H[i,j]   (Voigt bulk modulus).
J = sum_t w_t term_t with the weights as given (objective.py:222-244).

Gradient of 1/S_ii with respect to A is S[:,i] S[i,:] / S_ii^2 (from dS = -S dA S);
symmetrised because A depends on C_H only through its symmetric part.

Pinned by tests/test_oracle_objective.py: closed-form values on the
isotropic C (E=1, nu=0.3 gives stiffness 1 and shear 1/(2(1+nu))) and
central finite differences of every gradient.
"""
from __future__ import annotations

import numpy as np

STIFF_IDX = (0, 1, 2)
SHEAR_IDX = (3, 4, 5)


def _compliance(CH):
    A = 0.5 * (CH + CH.T) + 1e-9 * np.eye(6)
    return np.linalg.inv(A)


def _inv_diag_mean(CH, idx):
    S = _compliance(CH)
    val = np.mean([1.0 / S[i, i] for i in idx])
    G = np.zeros((6, 6))
    for i in idx:
        G += np.outer(S[:, i], S[i, :]) / S[i, i] ** 2
    G /= len(idx)
    return val, 0.5 * (G + G.T)


def stiffness(CH):
    return _inv_diag_mean(CH, STIFF_IDX)


def shear(CH):
    return _inv_diag_mean(CH, SHEAR_IDX)


def bulk(CH):
    G = np.zeros((6, 6))
    G[:3, :3] = 1.0 / 9.0
    return float(np.sum(CH[:3, :3]) / 9.0), G


def thermal(kappa):
    return float(np.trace(kappa)), np.eye(3)


def evaluate(weights: dict, CH=None, kappa=None):
    """Returns (J, G6 = dJ/dC_H, G3 = dJ/dkappa)."""
    J = 0.0
    G6 = np.zeros((6, 6))
    G3 = np.zeros((3, 3))
    for term, w in weights.items():
        if term == "thermal":
            v, g = thermal(kappa)
            G3 += w * g
        else:
            v, g = {"stiffness": stiffness, "shear": shear, "bulk": bulk}[term](CH)
            G6 += w * g
        J += w * v
    return J, G6, G3
