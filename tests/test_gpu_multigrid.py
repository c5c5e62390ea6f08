"""The library's multigrid internals against the oracle's dense Galerkin construction.

Each piece of the V-cycle is compared on its own, so a wrong stencil entry,
diagonal, transfer weight or coarsest inverse fails here instead of showing up
only as extra CG iterations:

* Galerkin stencils of every coarse level (sparse_coarsen_ke, EBE_GMG.py:226-337)
  against the oracle's element-wise P^T K P assembled into a matrix;
* smoother diagonals, point Jacobi (_compute_diag_inverse, EBE_GMG.py:375-393)
  and l1-Jacobi, and one smoother sweep (smooth_jacobi, EBE_GMG.py:395-411);
* level operators A_l u;
* restriction P^T (restrict_PT, EBE_GMG.py:59-79) against the oracle's
  prolongation matrix transposed, and <R r, e> = <r, P e> with the library's P;
* one V-cycle (GMGSolver._v_cycle, EBE_GMG.py:886-984) against the oracle's
  V-cycle with the same smoother;
* the coarsest dense inverse: A_c^-1 A_c = I - Z Z^T (Z the translations) with
  a partial last 64-tile.

Bars: FP64 1e-13 relative to the largest entry for assembled quantities, 1e-11
for the V-cycle (it compounds several levels), 1e-9 for the coarsest inverse
(its conditioning); FP32 hierarchy 2e-6 for entries, 2e-4 for the V-cycle.
"""
import numpy as np
import pytest
import torch

from oracle import fe as ofe, grid as ogrid, mg as omg

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0) if torch.cuda.is_available() else None


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(DEV)


def _rho(r, seed, lo):
    x = np.random.default_rng(seed).uniform(0.0, 1.0, r ** 3)
    return lo + (1 - lo) * x ** 3


def _Ke(physics, r):
    if physics == "elastic":
        return ofe.hexahedron(r, ofe.isotropic_elastic_tensor(1, 0.3))[0]
    return ofe.hexahedron_scalar(r)[0]


def _stencil_matrix(S, dims):
    """Dense matrix (nd*node + comp ordering) of a (27, nd, nd, N) stencil; offsets that
    wrap onto the same node (a side of 2) add up, as in an assembled operator."""
    nd = S.shape[1]
    nx, ny, nz = dims
    N = nx * ny * nz
    X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
    n = (X * ny + Y) * nz + Z
    A = np.zeros((N * nd, N * nd))
    for d in range(27):
        dx, dy, dz = d // 9 - 1, (d // 3) % 3 - 1, d % 3 - 1
        m = (((X + dx) % nx) * ny + (Y + dy) % ny) * nz + (Z + dz) % nz
        for i in range(nd):
            for j in range(nd):
                np.add.at(A, (n * nd + i, m * nd + j), S[d, i, j])
    return A


def _planes_to_dofs(v):
    """(ncase, nd, N) -> (N*nd, ncase) in the oracle's DOF order."""
    nc, nd, N = v.shape
    return v.transpose(2, 1, 0).reshape(N * nd, nc)


def _dofs_to_planes(v, nd):
    Nnd, nc = v.shape
    return v.reshape(Nnd // nd, nd, nc).transpose(2, 1, 0)


def _batch(r, nlev, precision="fp64", smoother="l1-jacobi"):
    import maple_amd
    return maple_amd.CellBatch([r], nlevels=nlev, device=DEV, precision=precision, smoother=smoother)


@pytest.mark.parametrize("physics", ["elastic", "thermal"])
@pytest.mark.parametrize("r,nlev", [(8, 3), (16, 3)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_galerkin_stencils(physics, r, nlev, precision):
    print(f"[stencils] {physics} r={r} {precision}", flush=True)
    rho = _rho(r, r, 1e-3)
    H = omg.Hierarchy(rho, _Ke(physics, r), (r, r, r), nlev)
    cb = _batch(r, nlev, precision)
    cb.mg_setup(physics, _t(rho), precision)
    tol = 1e-13 if precision == "fp64" else 2e-6
    for l in range(1, nlev):
        S = cb.mg_stencil(physics, l, precision).cpu().numpy()
        A = _stencil_matrix(S, H.dims[l])
        ref = H.A[l].toarray()
        assert np.abs(A - ref).max() <= tol * np.abs(ref).max(), f"level {l}"
        # symmetric and translation-free, as P^T K P must be
        assert np.abs(A - A.T).max() <= tol * np.abs(ref).max()
    cb.close()


@pytest.mark.parametrize("physics", ["elastic", "thermal"])
@pytest.mark.parametrize("smoother", ["jacobi", "l1-jacobi"])
def test_smoother_diagonals_and_sweep(physics, smoother):
    print(f"[diagonals] {physics} {smoother}", flush=True)
    r, nlev = 16, 3
    rho = _rho(r, 3, 1e-3)
    Ke = _Ke(physics, r)
    nd = Ke.shape[0] // 8
    H = omg.Hierarchy(rho, Ke, (r, r, r), nlev)
    cb = _batch(r, nlev, smoother=smoother)
    rt = cb.mg_setup(physics, _t(rho))
    omega = 1.6 if smoother == "l1-jacobi" else 0.6
    for l in range(nlev - 1):
        if smoother == "jacobi":
            diag = H.diag[l]
        elif l == 0:
            # element-wise absolute row sums sum_{e owns n} rho_e sum_j |Ke[row, j]|
            rows = np.tile(np.abs(Ke).sum(axis=1)[None, :], (r ** 3, 1)) * rho[:, None]
            diag = ogrid.scatter_elements(rows, (r, r, r), nd).ravel()
        else:
            diag = np.asarray(abs(H.A[l]).sum(axis=1)).ravel()
        dinv = cb.mg_dinv(physics, l).cpu().numpy()          # (nd, N)
        ref = (1.0 / diag).reshape(-1, nd).T
        assert np.abs(dinv - ref).max() <= 1e-13 * np.abs(ref).max(), f"level {l}"
        # one sweep out = u + omega Dinv (f - A u)
        N = H.dims[l][0] * H.dims[l][1] * H.dims[l][2]
        nc = 6 if nd == 3 else 3
        rng = np.random.default_rng(l)
        u = rng.standard_normal((nc, nd, N))
        f = rng.standard_normal((nc, nd, N))
        out = cb.mg_apply(physics, l, _t(u), _t(f), mode=3).cpu().numpy()
        ud, fd = _planes_to_dofs(u), _planes_to_dofs(f)
        Au = np.stack([H.apply(l, ud[:, j]) for j in range(nc)], axis=1)
        want = _dofs_to_planes(ud + omega * (1.0 / diag)[:, None] * (fd - Au), nd)
        assert np.abs(out - want).max() <= 1e-12 * np.abs(want).max(), f"sweep level {l}"
        res = cb.mg_apply(physics, l, _t(u), _t(f), mode=2).cpu().numpy()
        assert np.abs(res - _dofs_to_planes(fd - Au, nd)).max() <= 1e-12 * np.abs(fd - Au).max()
        Av = cb.mg_apply(physics, l, _t(u), mode=0).cpu().numpy()
        assert np.abs(Av - _dofs_to_planes(Au, nd)).max() <= 1e-13 * np.abs(Au).max()
    del rt
    cb.close()


@pytest.mark.parametrize("physics", ["elastic", "thermal"])
def test_restriction_is_the_adjoint_of_prolongation(physics):
    print(f"[restriction] {physics}", flush=True)
    r, nlev = 16, 4
    nd, nc = (3, 6) if physics == "elastic" else (1, 3)
    cb = _batch(r, nlev)
    rng = np.random.default_rng(5)
    dims = (r, r, r)
    for l in range(1, nlev):
        cdims = tuple(d // 2 for d in dims)
        Nf, Nc = int(np.prod(dims)), int(np.prod(cdims))
        rf = rng.standard_normal((nc, nd, Nf))
        bc = cb.mg_restrict(physics, l, _t(rf)).cpu().numpy()
        P = omg.prolongation(dims, cdims, nd)
        want = _dofs_to_planes(P.T @ _planes_to_dofs(rf), nd)
        assert np.abs(bc - want).max() <= 1e-13 * np.abs(want).max(), f"level {l}"
        ec = rng.standard_normal((nc, nd, Nc))
        pe = cb.prolong(physics, l, _t(ec)).cpu().numpy()
        for j in range(nc):
            lhs = float(np.sum(bc[j] * ec[j]))
            rhs = float(np.sum(rf[j] * pe[j]))
            assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1e-300) + 1e-12, f"level {l} case {j}"
        dims = cdims
    cb.close()


@pytest.mark.parametrize("physics", ["elastic", "thermal"])
@pytest.mark.parametrize("smoother", ["jacobi", "l1-jacobi"])
def test_vcycle_matches_oracle(physics, smoother):
    print(f"[vcycle] {physics} {smoother}", flush=True)
    r, nlev = 16, 3
    rho = _rho(r, 11, 1e-2)
    Ke = _Ke(physics, r)
    nd = Ke.shape[0] // 8
    nc = 6 if nd == 3 else 3
    omega = 1.6 if smoother == "l1-jacobi" else 0.6
    H = omg.Hierarchy(rho, Ke, (r, r, r), nlev, omega=omega, nsmooth=2)
    if smoother == "l1-jacobi":
        rows = np.tile(np.abs(Ke).sum(axis=1)[None, :], (r ** 3, 1)) * rho[:, None]
        H.diag[0] = ogrid.scatter_elements(rows, (r, r, r), nd).ravel()
        for l in range(1, nlev - 1):
            H.diag[l] = np.asarray(abs(H.A[l]).sum(axis=1)).ravel()
    cb = _batch(r, nlev, smoother=smoother)
    cb.set_solver(omega, 2, False)
    rt = cb.mg_setup(physics, _t(rho))
    rng = np.random.default_rng(2)
    rv = rng.standard_normal((nc, nd, r ** 3))
    rv -= rv.mean(axis=2, keepdims=True)   # no translation content (a consistent residual)
    z = cb.mg_vcycle(physics, _t(rv)).cpu().numpy()
    rd = _planes_to_dofs(rv)
    want = _dofs_to_planes(np.stack([H.vcycle(0, rd[:, j]) for j in range(nc)], axis=1), nd)
    assert np.linalg.norm(z - want) <= 1e-11 * np.linalg.norm(want)
    del rt
    cb.close()


@pytest.mark.parametrize("physics", ["elastic", "thermal"])
@pytest.mark.parametrize("sweeps,lo", [(2, 0.1), (3, 0.05)])
def test_chebyshev_vcycle_matches_oracle(physics, sweeps, lo):
    """V-cycle with the Chebyshev sweep weights (l1-Jacobi spectrum, [lo, 1]) against the
    oracle's V-cycle with the weights of oracle.mg.chebyshev_weights."""
    print(f"[vcycle chebyshev] {physics} sweeps={sweeps} lo={lo}", flush=True)
    r, nlev = 16, 3
    rho = _rho(r, 13, 1e-3)
    Ke = _Ke(physics, r)
    nd = Ke.shape[0] // 8
    nc = 6 if nd == 3 else 3
    H = omg.Hierarchy(rho, Ke, (r, r, r), nlev, nsmooth=sweeps,
                      weights=(omg.chebyshev_weights(sweeps, lo), omg.chebyshev_weights(sweeps, lo, rising=False)))
    rows = np.tile(np.abs(Ke).sum(axis=1)[None, :], (r ** 3, 1)) * rho[:, None]
    H.diag[0] = ogrid.scatter_elements(rows, (r, r, r), nd).ravel()
    for l in range(1, nlev - 1):
        H.diag[l] = np.asarray(abs(H.A[l]).sum(axis=1)).ravel()
    cb = _batch(r, nlev, smoother="l1-jacobi")
    cb.set_sweeps(sweeps, sweeps)
    cb.set_chebyshev(lo)
    rt = cb.mg_setup(physics, _t(rho))
    rng = np.random.default_rng(3)
    rv = rng.standard_normal((nc, nd, r ** 3))
    rv -= rv.mean(axis=2, keepdims=True)
    z = cb.mg_vcycle(physics, _t(rv)).cpu().numpy()
    rd = _planes_to_dofs(rv)
    want = _dofs_to_planes(np.stack([H.vcycle(0, rd[:, j]) for j in range(nc)], axis=1), nd)
    assert np.linalg.norm(z - want) <= 1e-11 * np.linalg.norm(want)
    del rt
    cb.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_chebyshev_solve_matches_oracle(precision):
    """Converged homogenization with the Chebyshev schedule (3 sweeps a side, [0.1, 1]): C_H
    and kappa_H at the FP64 bar against the oracle."""
    import maple_amd
    from oracle import homog as ohomog
    r = 16
    print(f"[chebyshev solve] {precision}", flush=True)
    rho = _rho(r, 17, 1e-9)
    cb = maple_amd.CellBatch([r], nlevels=3, device=DEV, precision=precision)
    cb.set_sweeps(3, 3)
    cb.set_chebyshev(0.1)
    CH, kap, _ = cb.homogenize(_t(rho), rtol=1e-10)
    cs = ohomog.CellSolver((r, r, r), nlevels=3, rtol=1e-12)
    CHo, _, _ = cs.elastic(rho)
    ko, _, _ = cs.thermal(rho)
    assert np.linalg.norm(CH[0].cpu().numpy() - CHo) <= 1e-9 * np.linalg.norm(CHo)
    assert np.linalg.norm(kap[0].cpu().numpy() - ko) <= 1e-9 * np.linalg.norm(ko)
    with pytest.raises(Exception):
        cb.set_smoother("jacobi")
        cb.set_chebyshev(0.1)   # the schedule needs the l1 bound of the spectrum
    cb.close()


@pytest.mark.parametrize("physics", ["elastic", "thermal"])
def test_vcycle_fp32_follows_fp64(physics):
    print(f"[vcycle fp32] {physics}", flush=True)
    r, nlev = 16, 3
    rho = _rho(r, 13, 1e-2)
    nd, nc = (3, 6) if physics == "elastic" else (1, 3)
    cb = _batch(r, nlev, "mixed")
    rt = cb.mg_setup(physics, _t(rho), "fp64")
    cb.mg_setup(physics, rt, "mixed")
    rng = np.random.default_rng(4)
    rv = rng.standard_normal((nc, nd, r ** 3))
    rv -= rv.mean(axis=2, keepdims=True)
    z64 = cb.mg_vcycle(physics, _t(rv), "fp64").cpu().numpy()
    z32 = cb.mg_vcycle(physics, _t(rv), "mixed").cpu().numpy()
    assert np.linalg.norm(z32 - z64) <= 2e-4 * np.linalg.norm(z64)
    cb.close()


@pytest.mark.parametrize("physics", ["elastic", "thermal"])
def test_coarsest_inverse(physics):
    """20^3 with 3 levels: a 5^3 coarsest grid, n = 375 (elastic) / 125 (thermal),
    i.e. 6 / 2 tiles of 64 with a partial last tile."""
    print(f"[coarsest inverse] {physics}", flush=True)
    r, nlev = 20, 3
    rho = _rho(r, 17, 5e-2)
    nd = 3 if physics == "elastic" else 1
    cb = _batch(r, nlev)
    rt = cb.mg_setup(physics, _t(rho))
    S = cb.mg_stencil(physics, nlev - 1).cpu().numpy()
    A = _stencil_matrix(S, (5, 5, 5))
    Ainv = cb.mg_coarsest_inverse(physics).cpu().numpy()
    n = A.shape[0]
    assert Ainv.shape == (n, n) and n == 125 * nd
    Z = omg.translation_basis(125, nd)
    Q = np.eye(n) - Z @ Z.T
    assert np.abs(Ainv @ A - Q).max() <= 1e-9
    assert np.abs(Ainv - Ainv.T).max() <= 1e-9 * np.abs(Ainv).max()
    # and the stencil it inverts is the oracle's Galerkin operator
    H = omg.Hierarchy(rho, _Ke(physics, r), (r, r, r), nlev)
    assert np.abs(A - H.A[-1].toarray()).max() <= 1e-13 * np.abs(A).max()
    del rt
    cb.close()


@pytest.mark.parametrize("physics", ["elastic", "thermal"])
@pytest.mark.parametrize("dims", [[64, 64, 64, 64], [64, 64, 64, 64, 48, 32]])
def test_prolong_node_form_matches_oracle(dims, physics):
    """Batches with at least 4096 level-0 blocks take the node form of the level-0
    prolongation (one thread per fine node for all cases, one branch on the x/y parity);
    in the 48^3 and 32^3 cells a wavefront spans two z-pencils and takes both branches."""
    import maple_amd
    from oracle import mg
    print(f"[prolong node form] {physics} {dims}", flush=True)
    nd, nc = (3, 6) if physics == "elastic" else (1, 3)
    cb = maple_amd.CellBatch(dims, nlevels=4, device=DEV)
    assert sum(r ** 3 for r in dims) // 256 >= 4096
    rng = np.random.default_rng(9)
    coarse = rng.standard_normal((nc, nd, cb.num_nodes(1)))
    fine = cb.prolong(physics, 1, torch.from_numpy(coarse).to(DEV)).cpu().numpy()
    o0 = o1 = 0
    for r in dims:
        n0, n1 = r ** 3, (r // 2) ** 3
        P = mg.prolongation((r, r, r), (r // 2,) * 3, 1)
        ref = np.stack([np.stack([P @ coarse[j, c, o1:o1 + n1] for c in range(nd)]) for j in range(nc)])
        np.testing.assert_allclose(fine[:, :, o0:o0 + n0], ref, rtol=1e-14, atol=1e-14)
        o0, o1 = o0 + n0, o1 + n1
    cb.close()
