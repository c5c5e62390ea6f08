"""CG breakdown is reported, never taken for convergence.

A negated element matrix makes the operator negative semidefinite, so the first
search direction has p^T K p < 0 on every load case.  The solve must return
MAPLE_EBREAKDOWN (raised as maple_amd.Breakdown, or kept in last_status with
strict=False), not MAPLE_OK with unconverged fields.  Every value stays finite:
the smoother and the coarsest inverse of a negative definite operator are
ordinary arithmetic.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0) if torch.cuda.is_available() else None


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_negative_operator_reports_breakdown(precision):
    import maple_amd
    from maple_amd import _lib as L
    print(f"[breakdown] {precision}", flush=True)
    r = 8
    cb = maple_amd.CellBatch([r], nlevels=3, device=DEV, precision=precision)
    Ke, Fe, X0, _, _, _ = maple_amd.hex8_matrices(r)
    cb.set_element_matrices(0, -Ke, Fe, X0)
    rho = torch.from_numpy(np.full(r ** 3, 0.5) + np.random.default_rng(0).uniform(-0.2, 0.2, r ** 3)).to(DEV)
    with pytest.raises(L.Breakdown):
        cb.homogenize(rho, ("elastic",), rtol=1e-8, maxit=50)
    CH, _, it = cb.homogenize(rho, ("elastic",), rtol=1e-8, maxit=50, strict=False)
    assert cb.last_status == L.MAPLE_EBREAKDOWN
    assert torch.isfinite(CH).all()
    assert (it[0, :6] < 50).all()   # stopped at the breakdown, not at maxit
    # the positive operator of the same context still converges
    cb.set_element_matrices(0, Ke, Fe, X0)
    CH, _, it = cb.homogenize(rho, ("elastic",), rtol=1e-8, maxit=200)
    assert cb.last_status == L.MAPLE_OK
    assert (it[0, :6] > 0).all()
    cb.close()
