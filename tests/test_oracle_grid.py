"""Pins oracle.grid: connectivity, RHS assembly and the element-by-element product."""
import numpy as np

from oracle import fe, grid


def test_connectivity_bruteforce():
    dims = (3, 4, 5)
    conn = grid.connectivity(dims)
    nx, ny, nz = dims
    e = 0
    for x in range(nx):
        for y in range(ny):
            for z in range(nz):
                for k, (ox, oy, oz) in enumerate(fe.HEX8):
                    n = (((x + ox) % nx) * ny + (y + oy) % ny) * nz + (z + oz) % nz
                    assert conn[e, k] == n
                e += 1


def test_matrix_free_equals_assembled():
    rng = np.random.default_rng(0)
    dims = (4, 4, 6)
    C = fe.isotropic_elastic_tensor(1.0, 0.3)
    Ke, Fe, X0 = fe.hexahedron(4, C)
    E = int(np.prod(dims))
    rho = rng.uniform(1e-3, 1.0, E)
    u = rng.standard_normal((E, 3))
    v = grid.ebe_matvec(u, rho, Ke, dims)
    K = grid.assemble_csr(Ke[None] * rho[:, None, None], grid.connectivity(dims), E, 3)
    assert np.abs(v.ravel() - K @ u.ravel()).max() < 1e-13 * np.abs(v).max()
    assert np.abs(K - K.T).max() < 1e-15
    # rigid translations are in the kernel
    for d in range(3):
        t = np.zeros((E, 3))
        t[:, d] = 1.0
        assert np.abs(grid.ebe_matvec(t, rho, Ke, dims)).max() < 1e-14


def test_rhs_assembly():
    rng = np.random.default_rng(1)
    dims = (4, 4, 4)
    C = fe.isotropic_elastic_tensor(1.0, 0.3)
    Ke, Fe, X0 = fe.hexahedron(4, C)
    rho = rng.uniform(0.1, 1.0, 64)
    F = grid.assemble_rhs(rho, Fe, dims)
    assert F.shape == (64, 3, 6)
    # every component of every case sums to zero (consistent singular system)
    assert np.abs(F.sum(axis=0)).max() < 1e-15
    # brute force scatter
    conn = grid.connectivity(dims)
    ref = np.zeros((64, 3, 6))
    for e in range(64):
        for k in range(8):
            ref[conn[e, k]] += rho[e] * Fe[3 * k:3 * k + 3]
    assert np.abs(F - ref).max() < 1e-16
    # homogeneous cell: F vanishes (the linear field is already in equilibrium)
    assert np.abs(grid.assemble_rhs(np.ones(64), Fe, dims)).max() < 1e-15
