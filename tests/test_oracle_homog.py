"""Pins oracle.homog (C_H, kappa_H, S_e) with closed forms and invariants."""
import numpy as np
import pytest

from oracle import fe, homog


def _lame(E=1.0, nu=0.3):
    return nu * E / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def test_homogeneous_cell():
    r = 4
    cs = homog.CellSolver((r, r, r), nlevels=2, rtol=1e-12)
    for c in (1.0, 0.37):
        CH, S, u = cs.elastic(np.full(r ** 3, c))
        assert np.abs(CH - c * fe.isotropic_elastic_tensor(1.0, 0.3)).max() < 1e-14
        kap, S3, u3 = cs.thermal(np.full(r ** 3, c))
        assert np.abs(kap - c * np.eye(3)).max() < 1e-14


@pytest.mark.parametrize("r,contrast", [(4, 1e-3), (8, 1e-9)])
def test_z_laminate_closed_form(r, contrast):
    """Layers normal to z: FE is exact (piecewise-linear solution)."""
    rng = np.random.default_rng(r)
    prof = rng.uniform(0.2, 1.0, r)
    prof[::3] = contrast
    rho = np.tile(prof, r * r)                    # element id = (x*r+y)*r+z, z fastest
    cs = homog.CellSolver((r, r, r), nlevels=2 if r == 4 else 3, rtol=1e-12)
    CH, S, u = cs.elastic(rho)
    lam, mu = _lame()
    a = lam + 2 * mu
    h = 1.0 / np.mean(1.0 / prof)
    m = np.mean(prof)
    ref = np.zeros((6, 6))
    ref[0, 0] = ref[1, 1] = 4 * mu * (lam + mu) / a * m + lam ** 2 / a * h
    ref[0, 1] = ref[1, 0] = 2 * mu * lam / a * m + lam ** 2 / a * h
    ref[2, 2] = a * h
    ref[0, 2] = ref[2, 0] = ref[1, 2] = ref[2, 1] = lam * h
    ref[3, 3] = mu * m               # xy: in-plane shear, arithmetic mean
    ref[4, 4] = ref[5, 5] = mu * h   # yz, xz: out-of-plane shear, harmonic mean
    assert np.linalg.norm(CH - ref) < 1e-11 * np.linalg.norm(ref)
    kap, S3, u3 = cs.thermal(rho)
    kref = np.diag([m, m, h])
    assert np.linalg.norm(kap - kref) < 1e-11 * np.linalg.norm(kref)


def _random_rho(r, seed):
    rng = np.random.default_rng(seed)
    return 1e-3 + rng.uniform(0, 1, r ** 3) ** 2


def test_euler_identity_and_symmetry():
    r = 8
    rho = _random_rho(r, 5)
    cs = homog.CellSolver((r, r, r), nlevels=3, rtol=1e-12)
    CH, S, u = cs.elastic(rho)
    assert np.abs(CH - CH.T).max() < 1e-13 * np.abs(CH).max()
    # degree-1 homogeneity: C_H(t rho) = t C_H(rho)
    CH2, _, _ = cs.elastic(2.5 * rho)
    assert np.linalg.norm(CH2 - 2.5 * CH) < 1e-11 * np.linalg.norm(CH)
    # element energies are PSD
    w = np.linalg.eigvalsh(S)
    assert w.min() > -1e-12 * w.max()


def test_finite_difference_sensitivity():
    r = 4
    rho = _random_rho(r, 7)
    cs = homog.CellSolver((r, r, r), nlevels=2, rtol=1e-13)
    CH, S, u = cs.elastic(rho)
    kap, S3, u3 = cs.thermal(rho)
    for e in (0, 13, 63):
        d = 1e-5 * rho[e]
        rp, rm = rho.copy(), rho.copy()
        rp[e] += d
        rm[e] -= d
        fd = (cs.elastic(rp)[0] - cs.elastic(rm)[0]) / (2 * d)
        assert np.abs(fd - S[e]).max() < 1e-6 * np.abs(S[e]).max()
        fdk = (cs.thermal(rp)[0] - cs.thermal(rm)[0]) / (2 * d)
        assert np.abs(fdk - S3[e]).max() < 1e-6 * np.abs(S3[e]).max()


def test_translation_invariance_and_bounds():
    r = 8
    rho = _random_rho(r, 11)
    cs = homog.CellSolver((r, r, r), nlevels=3, rtol=1e-12)
    CH, S, u = cs.elastic(rho)
    shifted = np.roll(rho.reshape(r, r, r), (3, -2, 5), axis=(0, 1, 2)).ravel()
    CHs, _, _ = cs.elastic(shifted)
    assert np.linalg.norm(CHs - CH) < 1e-11 * np.linalg.norm(CH)
    C0 = fe.isotropic_elastic_tensor(1.0, 0.3)
    voigt = rho.mean() * C0
    reuss = C0 / np.mean(1.0 / rho)
    assert np.linalg.eigvalsh(voigt - CH).min() > -1e-12
    assert np.linalg.eigvalsh(CH - reuss).min() > -1e-12
    # adding a rigid translation to u leaves the energies unchanged
    ut = u + np.array([0.3, -0.1, 0.7])[None, :, None]
    St = homog.element_energies(ut, cs.Ke3, cs.X03, cs.dims)
    assert np.abs(St - S).max() < 1e-13 * np.abs(S).max()
