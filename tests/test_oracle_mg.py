"""Pins oracle.mg: transfer operators, Galerkin coarse operators, converged solve."""
import numpy as np
import pytest

from oracle import fe, grid, homog, mg


def _setup(r=8, seed=0):
    rng = np.random.default_rng(seed)
    C = fe.isotropic_elastic_tensor(1.0, 0.3)
    Ke, Fe, X0 = fe.hexahedron(r, C)
    rho = rng.uniform(0.0, 1.0, r ** 3) ** 3 + 1e-9
    return Ke, Fe, X0, rho


def test_prolongation_weights():
    P = mg.prolongation((8, 8, 8), (4, 4, 4), 1)
    assert np.allclose(P @ np.ones(64), 1.0)                 # constants reproduced
    # a coarse-node-coincident fine node takes exactly that coarse value
    fid = (2 * 8 + 4) * 8 + 6                                 # fine (2,4,6) = 2 * coarse (1,2,3)
    cid = (1 * 4 + 2) * 4 + 3
    row = P.getrow(fid).toarray().ravel()
    assert row[cid] == 1.0 and row.sum() == 1.0
    # fine (3,3,3) is the centre of the coarse cell (1..2)^3: 8 weights 1/8
    row = P.getrow((3 * 8 + 3) * 8 + 3).toarray().ravel()
    assert np.count_nonzero(row) == 8 and np.allclose(row[row != 0], 0.125)
    # periodic wrap: fine x=7 interpolates coarse x=3 and x=0
    row = P.getrow((7 * 8 + 0) * 8 + 0).toarray().ravel()
    assert np.isclose(row[(3 * 4 + 0) * 4 + 0], 0.5) and np.isclose(row[0], 0.5)


@pytest.mark.parametrize("r", [4, 8])
def test_elementwise_galerkin_equals_assembled(r):
    Ke, Fe, X0, rho = _setup(r)
    dims = (r, r, r)
    cd = (r // 2,) * 3
    K = grid.assemble_csr(Ke[None] * rho[:, None, None], grid.connectivity(dims), r ** 3, 3)
    P = mg.prolongation(dims, cd, 3)
    ref = (P.T @ K @ P).toarray()
    Pc = fe.trilinear_child_prolongation(3)
    mats = mg.level1_elements(rho, Ke, cd, Pc)
    A1 = grid.assemble_csr(mats, grid.connectivity(cd), int(np.prod(cd)), 3).toarray()
    assert np.abs(A1 - ref).max() < 1e-13 * np.abs(ref).max()
    # two levels down through element matrices equals P2^T P1^T K P1 P2
    if r >= 8:
        cd2 = (r // 4,) * 3
        mats2 = mg.galerkin_elements(mats, cd2, Pc)
        A2 = grid.assemble_csr(mats2, grid.connectivity(cd2), int(np.prod(cd2)), 3).toarray()
        P2 = mg.prolongation(cd, cd2, 3)
        ref2 = P2.T.toarray() @ ref @ P2.toarray()
        assert np.abs(A2 - ref2).max() < 1e-13 * np.abs(ref2).max()


@pytest.mark.parametrize("physics", ["elastic", "thermal"])
def test_pcg_equals_direct(physics):
    r = 8
    Ke, Fe, X0, rho = _setup(r, seed=3)
    cs = homog.CellSolver((r, r, r), nlevels=3, rtol=1e-12)
    if physics == "elastic":
        CH, S, u = cs.elastic(rho)
        Kd, Fd, Xd = cs.Ke3, cs.Fe3, cs.X03
    else:
        CH, S, u = cs.thermal(rho)
        Kd, Fd, Xd = cs.Ke1, cs.Fe1, cs.X01
    F = grid.assemble_rhs(rho, Fd, cs.dims)
    ud = mg.direct_solve(rho, Kd, cs.dims, F)
    Sd = homog.element_energies(ud, Kd, Xd, cs.dims)
    CHd = homog.homogenize(rho, Sd)
    assert np.linalg.norm(CH - CHd) < 1e-10 * np.linalg.norm(CHd)
    assert max(cs.iters[physics]) < 200


@pytest.mark.parametrize("k,lo", [(1, 0.1), (2, 0.1), (3, 0.1), (3, 0.05), (4, 0.2)])
def test_chebyshev_weights_give_the_chebyshev_polynomial(k, lo):
    """prod_i (1 - w_i t) is T_k((1 + lo - 2t) / (1 - lo)) / T_k((1 + lo) / (1 - lo)) (numpy's
    Chebyshev series), so its largest magnitude on [lo, 1] is 1 / T_k((1 + lo) / (1 - lo))."""
    from numpy.polynomial import chebyshev as C
    w = mg.chebyshev_weights(k, lo)
    assert np.all(np.diff(w) > 0) and np.all(mg.chebyshev_weights(k, lo, rising=False) == w[::-1])
    t = np.linspace(0.0, 1.0, 401)
    p = np.prod([1.0 - wi * t for wi in w], axis=0)
    Tk = np.zeros(k + 1)
    Tk[k] = 1.0
    want = C.chebval((1.0 + lo - 2.0 * t) / (1.0 - lo), Tk) / C.chebval((1.0 + lo) / (1.0 - lo), Tk)
    assert np.abs(p - want).max() < 1e-12
    inside = (t >= lo)
    assert np.abs(p[inside]).max() <= 1.0 / C.chebval((1.0 + lo) / (1.0 - lo), Tk) + 1e-12
