"""Writes tests/golden/sphere_void_32.npz with the FP64 oracle.

Input: the sphere-void design of the 32^3 CPU workload (0.25, 0.1 inside
radius 0.3 L, U(-0.02, 0.02) from default_rng(0), clipped to [0.001, 1]),
rho = 1e-9 + (1 - 1e-9) x^3.  Outputs at rtol 1e-12: C_H, kappa_H, and
dJ/drho for J = 0.5 stiffness + 0.5 thermal, with its G6 = dJ/dC_H and
G3 = dJ/dkappa.  Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import driver, homog, objective, oc  # noqa: E402

r = 32
x0 = driver.x0_sphere_void(r, seed=0)
rho = oc.simp(x0)
cs = homog.CellSolver((r, r, r), nlevels=3, rtol=1e-12)
CH, S6, _ = cs.elastic(rho)
kap, S3, _ = cs.thermal(rho)
J, G6, G3 = objective.evaluate({"stiffness": 0.5, "thermal": 0.5}, CH, kap)
dJdrho = np.einsum("ij,eij->e", G6, S6) + np.einsum("ij,eij->e", G3, S3)
out = os.path.join(ROOT, "tests", "golden", "sphere_void_32.npz")
np.savez(out, x0=x0, CH=CH, kappa=kap, G6=G6, G3=G3, dJdrho=dJdrho, J=J, iters=np.array(cs.iters["elastic"]))
print("wrote", out, "sha256(x0) =", hashlib.sha256(x0.tobytes()).hexdigest())
