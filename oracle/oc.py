"""Periodic box filter and OC bisection update (oracle, FP64).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``box_filter``: circular pad k//2 then a ones(k^3)/k^3 correlation
  (main_IHD_v3.py:342-359), written as the mean of k^3 periodic shifts.
* ``oc_update``: the reference's 60-step bisection (main_IHD_v3.py:362-380),
  line for line: dJ clamped at 1e-12, l1=1e-12, l2=1e6, lam=(l1+l2)/2,
  x' = clip(clip(x sqrt(dJ/lam), x-move, x+move), xmin, 1),
  mean(x') > volfrac -> l1 = lam else l2 = lam; x' of the last lam returned.
* ``simp``: rho = rho_min + (1-rho_min) x^p  (main_IHD_v3.py:628), and its
  derivative for the chain rule dJ/dx = dJ/drho * drho/dx.

Pinned by tests/test_oracle_oc.py: the filter against scipy.ndimage
(wrap mode), mean preservation and constants; OC invariants (move and box
bounds, mean within the final bisection width) and the bisection path
against an independent bracketing on the monotone map lam -> mean(x'(lam)).
"""
from __future__ import annotations

import numpy as np


def box_filter(v: np.ndarray, dims, k: int = 3) -> np.ndarray:
    if k <= 0 or k % 2 == 0:
        raise ValueError(f"filter size must be positive and odd, got {k}")
    g = np.asarray(v, dtype=np.float64).reshape(dims)
    p = k // 2
    acc = np.zeros_like(g)
    for dx in range(-p, p + 1):
        for dy in range(-p, p + 1):
            for dz in range(-p, p + 1):
                acc += np.roll(g, shift=(dx, dy, dz), axis=(0, 1, 2))
    return (acc / float(k ** 3)).reshape(np.shape(v))


def oc_update(x, dJ, volfrac, move=0.1, xmin=1e-3, nbisect=60):
    x = np.asarray(x, dtype=np.float64)
    dJ = np.maximum(np.asarray(dJ, dtype=np.float64), 1e-12)
    l1, l2 = 1e-12, 1e6
    xn = x
    lam = None
    for _ in range(nbisect):
        lam = 0.5 * (l1 + l2)
        xn = x * np.sqrt(dJ / lam)
        xn = np.clip(xn, x - move, x + move)
        xn = np.clip(xn, xmin, 1.0)
        if xn.mean() > volfrac:
            l1 = lam
        else:
            l2 = lam
    return xn, lam


def simp(x, p=3.0, rho_min=1e-9):
    return rho_min + (1.0 - rho_min) * x ** p


def simp_grad(x, p=3.0, rho_min=1e-9):
    return (1.0 - rho_min) * p * x ** (p - 1.0)
