"""GPU parity of the HIP library against the FP64 oracle (called through the C ABI).

Tolerances: fine-level operator and RHS 1e-13 (relative to the largest entry);
C_H / kappa_H Frobenius-relative 1e-9 when both sides solve to rtol 1e-10
(the FP64 form's bar), sensitivities 1e-8, filter 1e-14, OC 1e-12 on
identical inputs.
"""
import numpy as np
import pytest
import torch

from oracle import fe as ofe, grid as ogrid, homog as ohomog, objective as oobj, oc as ooc

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0) if torch.cuda.is_available() else None


def _cells(dims, nlevels):
    import maple_amd
    return maple_amd.CellBatch(dims, nlevels=nlevels, device=DEV)


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(DEV)


def _rho(r, seed, contrast=1e-9):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 1.0, r ** 3)
    return contrast + (1 - contrast) * x ** 3


@pytest.mark.parametrize("r,nlev", [(8, 3), (16, 3), (32, 3), (72, 4), (64, 4)])
@pytest.mark.parametrize("physics", ["elastic", "thermal"])
def test_operator_and_rhs(r, nlev, physics):
    print(f"[operator] r={r} {physics}", flush=True)
    cb = _cells([r], nlev)
    rho = _rho(r, r)
    nd, nc = (3, 6) if physics == "elastic" else (1, 3)
    rng = np.random.default_rng(1)
    u = rng.standard_normal((nc, nd, r ** 3))
    v = cb.apply_operator(physics, _t(rho), _t(u)).cpu().numpy()
    Ke = ofe.hexahedron(r, ofe.isotropic_elastic_tensor(1, 0.3))[0] if nd == 3 else ofe.hexahedron_scalar(r)[0]
    for j in range(nc):
        ref = ogrid.ebe_matvec(u[j].T.copy(), rho, Ke, (r, r, r)).T
        assert np.abs(v[j] - ref).max() <= 1e-13 * np.abs(ref).max()
    Fe = ofe.hexahedron(r, ofe.isotropic_elastic_tensor(1, 0.3))[1] if nd == 3 else ofe.hexahedron_scalar(r)[1]
    F = cb.rhs(physics, _t(rho)).cpu().numpy()                 # (nc, nd, N)
    Fo = ogrid.assemble_rhs(rho, Fe, (r, r, r))                 # (N, nd, nc)
    assert np.abs(F - Fo.transpose(2, 1, 0)).max() <= 1e-13 * np.abs(Fo).max()


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("r,nlev,ncase", [(16, 3, 6), (24, 3, 4), (72, 4, 5), (64, 4, 6)])
def test_operator_variants(r, nlev, ncase, variant):
    """Every level-0 elastic operator variant (element rows per pencil lane, staged;
    node blocks with six or three load cases per thread, one, 2x1 or 2x2 nodes
    per thread) against the oracle's EBE product, also for fewer load cases than six."""
    import maple_amd
    print(f"[operator variant] r={r} variant={variant} ncase={ncase}", flush=True)
    cb = _cells([r], nlev)
    cb.set_variant(maple_amd._lib.MAPLE_KERNEL_FINE_OPERATOR, variant)
    rho = _rho(r, r + variant)
    rng = np.random.default_rng(variant)
    u = rng.standard_normal((ncase, 3, r ** 3))
    v = cb.apply_operator("elastic", _t(rho), _t(u)).cpu().numpy()
    Ke = ofe.hexahedron(r, ofe.isotropic_elastic_tensor(1, 0.3))[0]
    for j in range(ncase):
        ref = ogrid.ebe_matvec(u[j].T.copy(), rho, Ke, (r, r, r)).T
        assert np.abs(v[j] - ref).max() <= 1e-13 * np.abs(ref).max()
    cb.close()


@pytest.mark.parametrize("variant", [2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_homogenize_node_variants(variant, precision):
    """The node-block operator inside the whole solve (CG matvec, residuals, and the
    FP32 Jacobi sweeps of the mixed form's preconditioner) against the oracle."""
    import maple_amd
    r, nlev = 16, 3
    print(f"[homogenize variant] variant={variant} {precision}", flush=True)
    cb = maple_amd.CellBatch([r], nlevels=nlev, device=DEV, precision=precision)
    cb.set_variant(maple_amd._lib.MAPLE_KERNEL_FINE_OPERATOR, variant)
    rho = _rho(r, 11)
    CH, kap, it = cb.homogenize(_t(rho), rtol=1e-10)
    cs = ohomog.CellSolver((r, r, r), nlevels=3, rtol=1e-12)
    CHo, _, _ = cs.elastic(rho)
    ko, _, _ = cs.thermal(rho)
    assert np.linalg.norm(CH[0].cpu().numpy() - CHo) <= 1e-9 * np.linalg.norm(CHo)
    assert np.linalg.norm(kap[0].cpu().numpy() - ko) <= 1e-9 * np.linalg.norm(ko)
    cb.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lane_rotation_variant_64(precision):
    """Variant 8 (z neighbours from the neighbouring lanes, nz = 64 only) inside the
    whole 64^3 solve agrees with variant 5 (checked against the oracle above): the
    same operator up to rounding, so the same C_H and CG counts."""
    import maple_amd
    r, nlev = 64, 4
    print(f"[variant 8] {precision}", flush=True)
    rho = _rho(r, 21)
    out = []
    for v in (5, 8):
        cb = maple_amd.CellBatch([r], nlevels=nlev, device=DEV, precision=precision)
        cb.set_variant(maple_amd._lib.MAPLE_KERNEL_FINE_OPERATOR, v)
        CH, kap, it = cb.homogenize(_t(rho), rtol=1e-8, maxit=2000)
        out.append((CH[0].cpu().numpy(), kap[0].cpu().numpy(), it[0].astype(int)))
        cb.close()
    assert np.linalg.norm(out[0][0] - out[1][0]) <= 1e-7 * np.linalg.norm(out[0][0])
    assert np.array_equal(out[0][1], out[1][1])   # thermal: the same kernel in both
    assert np.abs(out[0][2][:6] - out[1][2][:6]).max() <= 2


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lagged_coarsest_inverse(precision):
    """maple_set_coarse_refresh(3): the second and third densities in a row are solved
    with the first one's coarsest inverse (still SPD), and each still meets the FP64
    bar against the oracle."""
    import maple_amd
    r, nlev = 16, 3
    cb = maple_amd.CellBatch([r], nlevels=nlev, device=DEV, precision=precision, islands=True)
    cb.set_coarse_refresh(3)
    cs = ohomog.CellSolver((r, r, r), nlevels=3, rtol=1e-12)
    for k in range(3):
        rho = _rho(r, 30 + k)
        CH, kap, it = cb.homogenize(_t(rho), rtol=1e-10, maxit=3000)
        CHo, _, _ = cs.elastic(rho)
        ko, _, _ = cs.thermal(rho)
        assert np.linalg.norm(CH[0].cpu().numpy() - CHo) <= 1e-9 * np.linalg.norm(CHo)
        assert np.linalg.norm(kap[0].cpu().numpy() - ko) <= 1e-9 * np.linalg.norm(ko)
    cb.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_coarse_operator_variant(precision):
    """The node-per-thread coarse stencil product (all rows of a node in one thread)
    inside the whole solve against the oracle, and against the row-per-thread form."""
    import maple_amd
    r, nlev = 16, 4
    print(f"[coarse variant] {precision}", flush=True)
    rho = _rho(r, 12)
    out = []
    for v in (0, 1):
        cb = maple_amd.CellBatch([r], nlevels=nlev, device=DEV, precision=precision)
        cb.set_variant(maple_amd._lib.MAPLE_KERNEL_COARSE_OPERATOR, v)
        CH, kap, it = cb.homogenize(_t(rho), rtol=1e-10)
        out.append((CH[0].cpu().numpy(), kap[0].cpu().numpy(), it))
        cb.close()
    cs = ohomog.CellSolver((r, r, r), nlevels=3, rtol=1e-12)
    CHo, _, _ = cs.elastic(rho)
    ko, _, _ = cs.thermal(rho)
    for CH, kap, _ in out:
        assert np.linalg.norm(CH - CHo) <= 1e-9 * np.linalg.norm(CHo)
        assert np.linalg.norm(kap - ko) <= 1e-9 * np.linalg.norm(ko)
    # the same operator up to rounding: the same CG iteration counts
    assert np.abs(out[0][2].astype(int) - out[1][2].astype(int)).max() <= 1


@pytest.mark.parametrize("r,nlev,seed", [(8, 2, 0), (8, 3, 1), (16, 3, 2), (16, 4, 3)])
def test_homogenize_parity(r, nlev, seed):
    print(f"[homogenize] r={r} nlev={nlev}", flush=True)
    cb = _cells([r], nlev)
    rho = _rho(r, seed)
    CH, kap, it = cb.homogenize(_t(rho), rtol=1e-10)
    cs = ohomog.CellSolver((r, r, r), nlevels=3 if r >= 8 else 2, rtol=1e-12)
    CHo, S6, _ = cs.elastic(rho)
    ko, S3, _ = cs.thermal(rho)
    assert np.linalg.norm(CH[0].cpu().numpy() - CHo) <= 1e-9 * np.linalg.norm(CHo)
    assert np.linalg.norm(kap[0].cpu().numpy() - ko) <= 1e-9 * np.linalg.norm(ko)
    w = {"stiffness": 0.4, "shear": 0.3, "thermal": 0.3}
    J, G6, G3 = oobj.evaluate(w, CHo, ko)
    d = cb.sensitivity(_t(G6).reshape(1, 6, 6), _t(G3).reshape(1, 3, 3)).cpu().numpy()
    do = np.einsum("ij,eij->e", G6, S6) + np.einsum("ij,eij->e", G3, S3)
    assert np.abs(d - do).max() <= 1e-8 * np.abs(do).max()
    assert it.max() < 1000   # random per-voxel 1:1e9 contrast is the hardest case for point Jacobi


def test_golden_sphere_void_32(golden_dir):
    import os
    g = np.load(os.path.join(golden_dir, "sphere_void_32.npz"))
    r = 32
    cb = _cells([r], 3)
    rho = ooc.simp(g["x0"])
    CH, kap, it = cb.homogenize(_t(rho), rtol=1e-10)
    assert np.linalg.norm(CH[0].cpu().numpy() - g["CH"]) <= 1e-9 * np.linalg.norm(g["CH"])
    assert np.linalg.norm(kap[0].cpu().numpy() - g["kappa"]) <= 1e-9 * np.linalg.norm(g["kappa"])
    d = cb.sensitivity(_t(g["G6"]).reshape(1, 6, 6), _t(g["G3"]).reshape(1, 3, 3)).cpu().numpy()
    assert np.abs(d - g["dJdrho"]).max() <= 1e-8 * np.abs(g["dJdrho"]).max()
    # the plain rtol 1e-6 solve already sits far inside the 1e-5 mixed-form bar
    CH6, _, _ = cb.homogenize(_t(rho), rtol=1e-6)
    assert np.linalg.norm(CH6[0].cpu().numpy() - g["CH"]) <= 1e-9 * np.linalg.norm(g["CH"])


@pytest.mark.parametrize("k", [1, 3, 5])
def test_filter_parity(k):
    dims = [8, 16, 12]
    cb = _cells(dims, 2)
    rng = np.random.default_rng(k)
    v = rng.standard_normal(sum(d ** 3 for d in dims))
    out = cb.filter(_t(v), k).cpu().numpy()
    o = 0
    for d in dims:
        n = d ** 3
        ref = ooc.box_filter(v[o:o + n], (d, d, d), k)
        assert np.abs(out[o:o + n] - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max())
        o += n


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oc_parity(seed):
    dims = [16, 8]
    cb = _cells(dims, 2)
    rng = np.random.default_rng(seed)
    n = sum(d ** 3 for d in dims)
    x = np.clip(0.25 + rng.uniform(-0.2, 0.2, n), 1e-3, 1.0)
    dJ = rng.uniform(0.0, 1.0, n) ** 2 * 10.0 ** rng.uniform(-3, 1)
    vf = [0.25, 0.4]
    xn = cb.oc_update(_t(x), _t(dJ), vf, 0.2, 1e-3).cpu().numpy()
    o = 0
    for i, d in enumerate(dims):
        m = d ** 3
        ref, lam = ooc.oc_update(x[o:o + m], dJ[o:o + m], vf[i], 0.2, 1e-3)
        assert np.abs(xn[o:o + m] - ref).max() <= 1e-12
        o += m


def test_batched_mixed_sizes_equal_single():
    dims = [8, 16, 8]
    cb = _cells(dims, 3)
    rhos = [_rho(d, 10 + i) for i, d in enumerate(dims)]
    CH, kap, it = cb.homogenize(_t(np.concatenate(rhos)), rtol=1e-10)
    for i, d in enumerate(dims):
        single = _cells([d], 3)
        c1, k1, _ = single.homogenize(_t(rhos[i]), rtol=1e-10)
        assert np.linalg.norm(CH[i].cpu().numpy() - c1[0].cpu().numpy()) <= 1e-10 * np.linalg.norm(c1[0].cpu().numpy())
        assert np.linalg.norm(kap[i].cpu().numpy() - k1[0].cpu().numpy()) <= 1e-10 * np.linalg.norm(k1[0].cpu().numpy())


def test_closed_forms_on_gpu():
    r = 16
    cb = _cells([r], 3)
    C0 = ofe.isotropic_elastic_tensor(1, 0.3)
    CH, kap, it = cb.homogenize(_t(np.full(r ** 3, 0.7)), rtol=1e-10)
    assert np.abs(CH[0].cpu().numpy() - 0.7 * C0).max() < 1e-14
    assert np.abs(kap[0].cpu().numpy() - 0.7 * np.eye(3)).max() < 1e-14
    prof = np.random.default_rng(3).uniform(0.2, 1.0, r)
    prof[::4] = 1e-9
    CH, kap, it = cb.homogenize(_t(np.tile(prof, r * r)), rtol=1e-12)
    lam, mu = C0[0, 1], C0[3, 3]
    a = lam + 2 * mu
    h, m = 1 / np.mean(1 / prof), np.mean(prof)
    ref = np.zeros((6, 6))
    ref[0, 0] = ref[1, 1] = 4 * mu * (lam + mu) / a * m + lam ** 2 / a * h
    ref[0, 1] = ref[1, 0] = 2 * mu * lam / a * m + lam ** 2 / a * h
    ref[2, 2] = a * h
    ref[0, 2] = ref[2, 0] = ref[1, 2] = ref[2, 1] = lam * h
    ref[3, 3] = mu * m
    ref[4, 4] = ref[5, 5] = mu * h
    # Frobenius-relative: the 1:1e9 laminate's soft entries (~5e-9) sit below the
    # FP64 resolution of a tensor whose in-plane entries are ~0.3
    assert np.linalg.norm(CH[0].cpu().numpy() - ref) <= 1e-11 * np.linalg.norm(ref)
    kref = np.diag([m, m, h])
    assert np.linalg.norm(kap[0].cpu().numpy() - kref) <= 1e-11 * np.linalg.norm(kref)


def test_properties_64():
    """64^3 at the config-2 start: Euler identity, symmetry, translation invariance, determinism."""
    import maple_amd
    r = 64
    cb = _cells([r], 4)
    x = maple_amd.x0_uniform_noise(r, seed=0)
    rho = ooc.simp(x)
    CH, kap, it = cb.homogenize(_t(rho), rtol=1e-6)
    CHn = CH[0].cpu().numpy()
    assert np.abs(CHn - CHn.T).max() <= 1e-12 * np.abs(CHn).max()
    # Euler: sum_e rho_e S_e(i,j) = C_H(i,j)
    for (i, j) in [(0, 0), (0, 1), (3, 3), (2, 5)]:
        G = np.zeros((6, 6))
        G[i, j] = 1.0
        s = cb.sensitivity(_t(G).reshape(1, 6, 6)).cpu().numpy()
        assert abs(np.dot(rho, s) - CHn[i, j]) <= 1e-12 * np.abs(CHn).max()
    CH2, kap2, _ = cb.homogenize(_t(rho), rtol=1e-6)
    assert torch.equal(CH, CH2) and torch.equal(kap, kap2)          # bitwise run-to-run
    sh = np.roll(rho.reshape(r, r, r), (5, -3, 17), axis=(0, 1, 2)).ravel()
    CH3, kap3, _ = cb.homogenize(_t(sh), rtol=1e-8)
    CH4, kap4, _ = cb.homogenize(_t(rho), rtol=1e-8)
    assert np.linalg.norm((CH3 - CH4).cpu().numpy()) <= 1e-9 * np.linalg.norm(CH4.cpu().numpy())


def test_errors_are_exceptions():
    import maple_amd
    from maple_amd._lib import NotConverged
    r = 8
    cb = _cells([r], 3)
    rho = _rho(r, 5)
    with pytest.raises(NotConverged):
        cb.homogenize(_t(rho), rtol=1e-14, maxit=2)
    bad = rho.copy()
    bad[7] = np.nan
    with pytest.raises(maple_amd._lib.MapleError):
        cb.homogenize(_t(bad), rtol=1e-6)
    with pytest.raises(ValueError):
        cb.filter(_t(np.zeros(r ** 3)), 2)
    # the context still works after the failures
    CH, kap, it = cb.homogenize(_t(rho), rtol=1e-8)
    assert np.isfinite(CH.cpu().numpy()).all()


def test_ihd_dropin_autograd():
    """The reference-shaped IHD: J(C_H, kappa) through autograd gives the oracle's dJ/dx."""
    import maple_amd
    r = 8
    cfg = maple_amd.make_cfg(r, smooth_iter=[2, 2, 3])
    ihd = maple_amd.IHD(cfg, DEV, rtol=1e-11)
    x = torch.from_numpy(np.clip(0.3 + np.random.default_rng(2).uniform(-0.25, 0.25, r ** 3), 1e-3, 1)).to(DEV)
    batch = maple_amd.build_unitcell_batch(DEV, r, ooc.simp(x).reshape(r, r, r).float())
    xv = x.reshape(-1, 1).clone().requires_grad_(True)
    rho = 1e-9 + (1 - 1e-9) * xv.clamp(1e-3, 1.0) ** 3
    batch["rho"] = rho
    n = batch["feature"].shape[0]
    Ft = maple_amd.assemble_F_fast_scalar(batch["node_index"], ihd.Fe_1d, n, rho)
    kap = ihd.TO_homo_1d(batch, ihd.Ke_1d, Ft, ihd.X0_1d, batch_size=1)
    Fm = maple_amd.assemble_F_fast(batch["node_index"], ihd.Fe_3d, n, rho)
    CH = ihd.TO_homo_3d(batch, ihd.Ke_3d, Fm, ihd.X0_3d, batch_size=1)
    J = maple_amd.objective_from_spec(obj="stiffness:0.5,thermal:0.5", CH=CH, kappa=kap)
    (g,) = torch.autograd.grad(J, xv)
    xn = x.cpu().numpy()
    cs = ohomog.CellSolver((r, r, r), nlevels=3, rtol=1e-12)
    rr = ooc.simp(xn)
    CHo, S6, _ = cs.elastic(rr)
    ko, S3, _ = cs.thermal(rr)
    Jo, G6, G3 = oobj.evaluate({"stiffness": 0.5, "thermal": 0.5}, CHo, ko)
    go = (np.einsum("ij,eij->e", G6, S6) + np.einsum("ij,eij->e", G3, S3)) * ooc.simp_grad(xn)
    assert abs(J.item() - Jo) <= 1e-9 * abs(Jo)
    assert np.abs(g.cpu().numpy().ravel() - go).max() <= 1e-8 * np.abs(go).max()
    # module-level drop-ins on the same inputs
    f = maple_amd.avg_filter3d_x_periodic(torch.from_numpy(go).to(DEV).reshape(-1, 1), batch["elem_coords"], r, 3)
    assert np.abs(f.cpu().numpy().ravel() - ooc.box_filter(go, (r, r, r), 3)).max() <= 1e-14 * np.abs(go).max()


def test_ihd_dropin_batch_of_cells():
    """batch_size > 1: every cell is solved and, as the reference's TO_homo_3d /
    TO_homo_1d return homogenized_batch_3d(...)[0] (model.py:128, 158), the result is
    the first cell's C_H / kappa_H; its gradient is the single-cell one on the first
    cell's elements and zero on the others."""
    import maple_amd
    r, nb = 8, 3
    ihd = maple_amd.IHD(maple_amd.make_cfg(r, smooth_iter=[2, 2, 3]), DEV, rtol=1e-11)
    rng = np.random.default_rng(4)
    x = torch.from_numpy(np.clip(0.3 + rng.uniform(-0.25, 0.25, nb * r ** 3), 1e-3, 1)).to(DEV)
    outs = {}
    for m in (nb, 1):
        xv = x[:m * r ** 3].reshape(-1, 1).clone().requires_grad_(True)
        rho = 1e-9 + (1 - 1e-9) * xv ** 3
        batch = {"rho": rho}
        Fm = torch.zeros(m * r ** 3 * 18, dtype=torch.float64, device=DEV)
        Ft = torch.zeros(m * r ** 3 * 3, dtype=torch.float64, device=DEV)
        CH = ihd.TO_homo_3d(batch, ihd.Ke_3d, Fm, ihd.X0_3d, batch_size=m)
        kap = ihd.TO_homo_1d(batch, ihd.Ke_1d, Ft, ihd.X0_1d, batch_size=m)
        assert CH.shape == (6, 6) and kap.shape == (3, 3)
        J = maple_amd.objective_from_spec(obj="stiffness:0.5,thermal:0.5", CH=CH, kappa=kap)
        (g,) = torch.autograd.grad(J, xv)
        outs[m] = (CH.detach().cpu().numpy(), kap.detach().cpu().numpy(), g.cpu().numpy().ravel())
    CHb, kb, gb = outs[nb]
    CH1, k1, g1 = outs[1]
    assert np.linalg.norm(CHb - CH1) <= 1e-10 * np.linalg.norm(CH1)
    assert np.linalg.norm(kb - k1) <= 1e-10 * np.linalg.norm(k1)
    assert np.abs(gb[:r ** 3] - g1).max() <= 1e-9 * np.abs(g1).max()
    assert np.all(gb[r ** 3:] == 0.0)
    with pytest.raises(ValueError):
        ihd.TO_homo_3d({"rho": x.reshape(-1, 1)}, ihd.Ke_3d, None, ihd.X0_3d, batch_size=2)


def test_batched_oc_steps_equal_single_cells():
    """A mixed-size batch with per-cell objectives and volume fractions advances exactly like
    the same cells run one by one (convergence-aware block skipping included)."""
    import maple_amd
    spec = [(8, {"stiffness": 0.5, "thermal": 0.5}, 0.3), (16, {"shear": 1.0}, 0.2), (8, {"thermal": 1.0}, 0.4)]
    xs = [np.clip(vf + np.random.default_rng(i).uniform(-0.05, 0.05, r ** 3), 1e-3, 1) for i, (r, w, vf) in enumerate(spec)]
    cb = maple_amd.CellBatch([s[0] for s in spec], nlevels=3, device=DEV)
    drv = maple_amd.OCDriver(cb, [s[1] for s in spec], [s[2] for s in spec], rtol=1e-10, maxit=3000)
    x = _t(np.concatenate(xs))
    for _ in range(3):
        x, info = drv.step(x)
    o = 0
    for i, (r, w, vf) in enumerate(spec):
        c1 = maple_amd.CellBatch([r], nlevels=3, device=DEV)
        d1 = maple_amd.OCDriver(c1, w, vf, rtol=1e-10, maxit=3000)
        xi = _t(xs[i])
        for _ in range(3):
            xi, _ = d1.step(xi)
        assert np.abs(x[o:o + r ** 3].cpu().numpy() - xi.cpu().numpy()).max() <= 1e-9
        o += r ** 3


def test_concurrent_physics_equal_sequential():
    """Elastic and thermal solved at the same time (second stream and host thread) give
    bitwise the results of solving them one after the other."""
    r = 32
    rho = _t(_rho(r, 9))
    a = _cells([r, 16], 3)
    b = _cells([r, 16], 3)
    b.set_concurrency(False)
    rr = torch.cat([rho, _t(_rho(16, 10))])
    CHa, ka, ia = a.homogenize(rr, rtol=1e-8)
    CHb, kb, ib = b.homogenize(rr, rtol=1e-8)
    np.testing.assert_array_equal(ia, ib)
    assert torch.equal(CHa, CHb) and torch.equal(ka, kb)
    da = a.sensitivity(torch.ones(2, 6, 6, dtype=torch.float64, device=DEV), torch.ones(2, 3, 3, dtype=torch.float64, device=DEV))
    db = b.sensitivity(torch.ones(2, 6, 6, dtype=torch.float64, device=DEV), torch.ones(2, 3, 3, dtype=torch.float64, device=DEV))
    assert torch.equal(da, db)


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("r,ncase", [(64, 3), (64, 2), (32, 3)])
def test_thermal_operator_variants(r, ncase, variant):
    """The thermal level-0 operator: element rows (variants 0, 1) and the node form
    (variant 2; nz = 64 cells with a Q1-Laplacian element matrix, element rows
    otherwise) against the oracle's EBE product, also for fewer load cases than three."""
    import maple_amd
    print(f"[thermal variant] r={r} variant={variant} ncase={ncase}", flush=True)
    cb = _cells([r], 4)
    cb.set_variant(maple_amd._lib.MAPLE_KERNEL_THERMAL_OPERATOR, variant)
    rho = _rho(r, 7 + variant)
    u = np.random.default_rng(variant).standard_normal((ncase, 1, r ** 3))
    v = cb.apply_operator("thermal", _t(rho), _t(u)).cpu().numpy()
    Ke = ofe.hexahedron_scalar(r)[0]
    for j in range(ncase):
        ref = ogrid.ebe_matvec(u[j].T.copy(), rho, Ke, (r, r, r)).T
        assert np.abs(v[j] - ref).max() <= 1e-13 * np.abs(ref).max()
    cb.close()


def test_thermal_operator_general_element_matrix():
    """An element matrix without the Q1-Laplacian structure (one coupling of corners 0
    and 1 shifted, still symmetric with zero row sums) falls back to the element-row
    kernel and stays exact."""
    print("[thermal general Ke]", flush=True)
    r = 64
    cb = _cells([r], 4)
    Ke, Fe, X0 = ofe.hexahedron_scalar(r)
    Ka = Ke.copy()
    d = 0.01 * Ke[0, 0]
    Ka[0, 1] += d
    Ka[1, 0] += d
    Ka[0, 0] -= d
    Ka[1, 1] -= d
    cb.set_element_matrices(0, Ke8=Ka, Fe8=Fe, X0_8=X0)
    rho = _rho(r, 3)
    u = np.random.default_rng(5).standard_normal((3, 1, r ** 3))
    v = cb.apply_operator("thermal", _t(rho), _t(u)).cpu().numpy()
    for j in range(3):
        ref = ogrid.ebe_matvec(u[j].T.copy(), rho, Ka, (r, r, r)).T
        assert np.abs(v[j] - ref).max() <= 1e-13 * np.abs(ref).max()
    cb.close()


@pytest.mark.parametrize("variant", [2, 3])
def test_thermal_node_form_with_face_couplings(variant):
    """A conductivity matrix with the Hamming-distance structure whose distance-1 entry is
    not zero (the Q1 Laplacian plus delta * g[d], g = [3, -1, 0, 0]: symmetric, zero row
    sums) keeps the node form, now with the six face neighbours coupled, and stays exact."""
    from maple_amd import _lib as L
    print(f"[thermal node form, face couplings] variant {variant}", flush=True)
    r = 64
    cb = _cells([r], 4)
    Ke, Fe, X0 = ofe.hexahedron_scalar(r)
    corners = [tuple(int(v) for v in c) for c in ofe.HEX8]   # the element's corner order
    g = [3.0, -1.0, 0.0, 0.0]
    delta = 0.05 * Ke[0, 0]
    Ka = Ke.copy()
    for a in range(8):
        for b in range(8):
            d = sum(int(p != q) for p, q in zip(corners[a], corners[b]))
            Ka[a, b] += delta * g[d]
    cb.set_element_matrices(0, Ke8=Ka, Fe8=Fe, X0_8=X0)
    cb.set_variant(L.MAPLE_KERNEL_THERMAL_OPERATOR, variant)
    rho = _rho(r, 4)
    u = np.random.default_rng(6).standard_normal((3, 1, r ** 3))
    v = cb.apply_operator("thermal", _t(rho), _t(u)).cpu().numpy()
    for j in range(3):
        ref = ogrid.ebe_matvec(u[j].T.copy(), rho, Ka, (r, r, r)).T
        assert np.abs(v[j] - ref).max() <= 1e-13 * np.abs(ref).max()
    cb.close()
