"""The FP32 level-0 elastic operator on the matrix cores (variant 9) against the oracle.

The element product Ke (rho_e u_e) of EBEOperator.matvec (EBE_GMG.py:412-430) in FP32
on v_mfma_f32_16x16x4f32.  Tolerances: FP32 arithmetic on FP32-rounded inputs, so the
operator is compared with the FP64 oracle at 1e-5 of the largest entry (measured:
~1e-7) and with the vector-FMA FP32 kernel of the same context at 1e-5; solves that
use it as the preconditioner's operator keep the FP64 form's 1e-9 bar on C_H and
kappa_H (CG runs in FP64 around the FP32 V-cycle).
"""
import numpy as np
import pytest
import torch

from oracle import fe as ofe, grid as ogrid, homog as ohomog

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0) if torch.cuda.is_available() else None
MFMA = 9


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(DEV)


def _rho(n, seed, contrast=1e-9):
    x = np.random.default_rng(seed).uniform(0.0, 1.0, n)
    return contrast + (1 - contrast) * x ** 3


def _batch(dims, nlevels=3, precision="mixed"):
    import maple_amd
    return maple_amd.CellBatch(dims, nlevels=nlevels, device=DEV, precision=precision)


@pytest.mark.parametrize("dims", [[16], [32], [64], [48], [16, 48, 32, 64]])
def test_mfma_operator_matches_oracle(dims):
    """Level-0 FP32 product K u of every cell of the batch against the oracle's EBE product."""
    import maple_amd
    print(f"[mfma operator] dims={dims}", flush=True)
    cb = _batch(dims)
    cb.set_variant(maple_amd._lib.MAPLE_KERNEL_FINE_OPERATOR, MFMA)
    n = sum(r ** 3 for r in dims)
    rho = _rho(n, sum(dims))
    keep = cb.mg_setup("elastic", _t(rho), "fp32")
    u = np.random.default_rng(3).standard_normal((6, 3, n))
    v = cb.mg_apply("elastic", 0, _t(u), mode=0, precision="fp32").cpu().numpy()
    off = 0
    for r in dims:
        Ke = ofe.hexahedron(r, ofe.isotropic_elastic_tensor(1, 0.3))[0]
        sl = slice(off, off + r ** 3)
        for j in range(6):
            ref = ogrid.ebe_matvec(u[j][:, sl].T.copy(), rho[sl], Ke, (r, r, r)).T
            err = np.abs(v[j][:, sl] - ref).max() / np.abs(ref).max()
            assert err <= 1e-5, (r, j, err)
        off += r ** 3
    del keep
    cb.close()


@pytest.mark.parametrize("dims", [[64], [32, 48]])
@pytest.mark.parametrize("mode", [0, 2, 3])
def test_mfma_modes_match_vector_kernel(dims, mode):
    """Residual (mode 2) and Jacobi sweep (mode 3) of the MFMA kernel against the vector-FMA
    FP32 kernel on the same hierarchy (same FP32 diagonal, same inputs)."""
    import maple_amd
    print(f"[mfma modes] dims={dims} mode={mode}", flush=True)
    cb = _batch(dims, nlevels=4 if min(dims) >= 32 else 3)
    n = sum(r ** 3 for r in dims)
    rho = _rho(n, 7)
    keep = cb.mg_setup("elastic", _t(rho), "fp32")
    rng = np.random.default_rng(mode)
    u, f = _t(rng.standard_normal((6, 3, n))), _t(rng.standard_normal((6, 3, n)))
    out = {}
    for var in (5, MFMA):
        cb.set_variant(maple_amd._lib.MAPLE_KERNEL_FINE_OPERATOR, var)
        out[var] = cb.mg_apply("elastic", 0, u, f, mode=mode, precision="fp32").cpu().numpy()
    scale = np.abs(out[5]).max()
    assert np.abs(out[MFMA] - out[5]).max() <= 1e-5 * scale
    del keep
    cb.close()


@pytest.mark.parametrize("dims", [[16]])
def test_mfma_homogenize_matches_oracle(dims):
    """Mixed-form solve with the MFMA operator in the FP32 V-cycle: C_H and kappa_H at the
    FP64 bar against the oracle."""
    import maple_amd
    r = dims[0]
    print(f"[mfma homogenize] r={r}", flush=True)
    cb = _batch(dims)
    cb.set_variant(maple_amd._lib.MAPLE_KERNEL_FINE_OPERATOR, MFMA)
    rho = _rho(r ** 3, 11, contrast=1e-6)
    CH, kap, it = cb.homogenize(_t(rho), rtol=1e-10)
    cs = ohomog.CellSolver((r, r, r), nlevels=3, rtol=1e-12)
    CHo, _, _ = cs.elastic(rho)
    ko, _, _ = cs.thermal(rho)
    assert np.linalg.norm(CH[0].cpu().numpy() - CHo) <= 1e-9 * np.linalg.norm(CHo)
    assert np.linalg.norm(kap[0].cpu().numpy() - ko) <= 1e-9 * np.linalg.norm(ko)
    cb.close()


def test_mfma_batch_matches_single_cells():
    """A mixed-resolution batch solved with the MFMA operator gives every cell the C_H of
    its own single-cell solve (vector-FMA operator), and a second run is bitwise equal."""
    import maple_amd
    dims = [16, 48, 32]
    print(f"[mfma batch] dims={dims}", flush=True)
    n = sum(r ** 3 for r in dims)
    rho = _rho(n, 5, contrast=1e-6)
    cb = _batch(dims, nlevels=3)
    cb.set_variant(maple_amd._lib.MAPLE_KERNEL_FINE_OPERATOR, MFMA)
    CH, kap, _ = cb.homogenize(_t(rho), rtol=1e-10)
    CH2, kap2, _ = cb.homogenize(_t(rho), rtol=1e-10)
    assert torch.equal(CH, CH2) and torch.equal(kap, kap2)
    cb.close()
    off = 0
    for i, r in enumerate(dims):
        one = _batch([r], nlevels=3)
        C1, k1, _ = one.homogenize(_t(rho[off:off + r ** 3]), rtol=1e-10)
        assert torch.linalg.norm(CH[i] - C1[0]) <= 1e-9 * torch.linalg.norm(C1[0])
        assert torch.linalg.norm(kap[i] - k1[0]) <= 1e-9 * torch.linalg.norm(k1[0])
        one.close()
        off += r ** 3
