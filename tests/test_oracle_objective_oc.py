"""Pins oracle.objective and oracle.oc."""
import numpy as np
import pytest
import scipy.ndimage as ndi
import scipy.optimize as so

from oracle import fe, objective as ob, oc


def test_objective_isotropic_values():
    C = fe.isotropic_elastic_tensor(1.0, 0.3)
    E, _ = ob.stiffness(C)
    G, _ = ob.shear(C)
    assert abs(E - 1.0) < 1e-8                 # 1/S_00 = Young's modulus
    assert abs(G - 1.0 / 2.6) < 1e-8           # 1/S_33 = shear modulus
    K, _ = ob.bulk(C)
    lam = 0.3 / (1.3 * 0.4)
    mu = 1 / 2.6
    assert abs(K - (lam + 2 * mu / 3)) < 1e-14  # Voigt bulk of an isotropic C is lam + 2mu/3
    assert ob.thermal(np.diag([1.0, 2.0, 3.0]))[0] == 6.0


@pytest.mark.parametrize("term", ["stiffness", "shear", "bulk"])
def test_objective_gradients_fd(term):
    rng = np.random.default_rng(2)
    A = rng.standard_normal((6, 6))
    CH = fe.isotropic_elastic_tensor(1.0, 0.3) * 0.2 + 0.01 * (A @ A.T)
    f = {"stiffness": ob.stiffness, "shear": ob.shear, "bulk": ob.bulk}[term]
    _, G = f(CH)
    h = 1e-6
    for i in range(6):
        for j in range(6):
            D = np.zeros((6, 6))
            D[i, j] = h
            fd = (f(CH + D)[0] - f(CH - D)[0]) / (2 * h)
            assert abs(fd - G[i, j]) < 1e-6 * max(1.0, abs(G).max())


def test_weighted_sum():
    C = fe.isotropic_elastic_tensor(1.0, 0.3)
    J, G6, G3 = ob.evaluate({"stiffness": 0.4, "shear": 0.3, "thermal": 0.3}, C, np.eye(3))
    assert abs(J - (0.4 * 1.0 + 0.3 / 2.6 + 0.3 * 3.0)) < 1e-8
    assert np.allclose(G3, 0.3 * np.eye(3))


@pytest.mark.parametrize("k", [1, 3, 5])
def test_box_filter_against_ndimage(k):
    rng = np.random.default_rng(k)
    dims = (6, 7, 8)
    v = rng.standard_normal(int(np.prod(dims)))
    ref = ndi.uniform_filter(v.reshape(dims), size=k, mode="wrap").ravel()
    out = oc.box_filter(v, dims, k)
    assert np.abs(out - ref).max() < 1e-14
    assert abs(out.mean() - v.mean()) < 1e-14
    assert np.allclose(oc.box_filter(np.full(v.size, 2.5), dims, k), 2.5)
    with pytest.raises(ValueError):
        oc.box_filter(v, dims, 2)


def test_oc_invariants_and_root():
    rng = np.random.default_rng(4)
    n = 4096
    x = np.clip(0.25 + rng.uniform(-0.1, 0.1, n), 1e-3, 1.0)
    dJ = rng.uniform(0.0, 2.0, n) ** 2
    vf, move, xmin = 0.25, 0.2, 1e-3
    xn, lam = oc.oc_update(x, dJ, vf, move, xmin)
    assert np.all(np.abs(xn - x) <= move + 1e-15)
    assert np.all((xn >= xmin) & (xn <= 1.0))
    assert abs(xn.mean() - vf) < 1e-9

    def g(l):
        y = np.clip(np.clip(x * np.sqrt(np.maximum(dJ, 1e-12) / l), x - move, x + move), xmin, 1.0)
        return y.mean() - vf

    root = so.brentq(g, 1e-12, 1e6, xtol=1e-15, rtol=4 * np.finfo(float).eps)
    assert abs(lam - root) <= 2e-12 + 1e-12 * root


def test_simp_derivative():
    x = np.linspace(0.01, 1.0, 50)
    h = 1e-7
    fd = (oc.simp(x + h) - oc.simp(x - h)) / (2 * h)
    assert np.allclose(oc.simp_grad(x), fd, rtol=1e-7)
