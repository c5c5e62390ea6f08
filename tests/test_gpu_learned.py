"""GPU checks of the learned warm start: the BF16-MFMA network's forward pass and
parameter gradient against the FP64 restatement (with the same bfloat16
roundings emulated), training on the GPU, and the solver's start selection
(a start worse than zero is never used)."""
import os

import numpy as np
import pytest
import torch

from oracle import unet as ounet

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0) if torch.cuda.is_available() else None


def _inputs(n, B, seed):
    rng = np.random.default_rng(seed)
    feat = rng.uniform(0, 1, (B, n, n, n))
    target = ounet.bf16_round(rng.standard_normal((B, n, n, n, ounet.NOUT)))
    theta = ounet.bf16_round(ounet.init_params(seed)) * 1.0
    # random biases so that every bias gradient path is exercised
    th = ounet.unflatten(theta.copy())
    for k in th:
        th[k][1][:] = ounet.bf16_round(0.05 * rng.standard_normal(th[k][1].shape))
    return feat, target, ounet.flatten(th)


def _net(n, B, theta):
    from maple_amd import learned
    return learned.UNet(n=n, max_batch=B, device=DEV, theta=theta.astype(np.float32))


@pytest.mark.parametrize("n,B", [(16, 2), (32, 1)])
def test_forward_matches_oracle(n, B):
    print(f"[unet forward] n={n} B={B}", flush=True)
    feat, _, theta = _inputs(n, B, 1)
    net = _net(n, B, theta)
    out = net.forward(torch.from_numpy(feat).to(DEV)).cpu().numpy().reshape(B, n, n, n, ounet.NOUT)
    ref, _ = ounet.forward(theta, feat, q=ounet.bf16_round)
    err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    print(f"forward rel err {err:.2e}", flush=True)
    assert err < 5e-3
    # without the bf16 emulation the FP64 network is still close (bf16 noise only)
    ref64, _ = ounet.forward(theta, feat)
    assert np.linalg.norm(out - ref64) / np.linalg.norm(ref64) < 5e-2


@pytest.mark.parametrize("weighted", [False, True])
def test_gradient_matches_oracle(weighted):
    n, B = 16, 2
    print(f"[unet gradient] n={n} B={B} weighted={weighted}", flush=True)
    feat, target, theta = _inputs(n, B, 2)
    net = _net(n, B, theta)
    net.set_loss(weighted)
    tt = torch.from_numpy(target.astype(np.float32)).to(DEV).to(torch.bfloat16).reshape(B, n ** 3, ounet.NOUT)
    loss, g = net.loss_grad(torch.from_numpy(feat).to(DEV), tt)
    g = g.cpu().numpy().astype(np.float64)
    Lref, gref = ounet.backward(theta, feat, target, q=ounet.bf16_round, weighted=weighted)
    assert abs(loss - Lref) <= 1e-4 * Lref
    p = ounet.unflatten(g)
    pr = ounet.unflatten(gref)
    for name, _, _ in ounet.param_shapes():
        for k in (0, 1):
            a, b = p[name][k].ravel(), pr[name][k].ravel()
            err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300)
            print(f"{name}.{'Wb'[k]}: rel err {err:.2e}  cos {cos:.6f}", flush=True)
            assert err < 2e-2 and cos > 0.9995, name


def test_training_reduces_loss():
    """A target that is a fixed smooth function of the feature is learned."""
    n, B = 16, 4
    print(f"[unet training] n={n} B={B}", flush=True)
    from maple_amd import learned
    rng = np.random.default_rng(5)
    feats, tgts = [], []
    for _ in range(8):
        f = rng.uniform(0, 1, (n, n, n))
        f = (f + np.roll(f, 1, 0) + np.roll(f, 1, 1) + np.roll(f, 1, 2)) / 4
        a = (np.arange(ounet.NOUT) - 10) / 10
        t = a * (f[..., None] - 0.5) * 4 + (1 - a ** 2) * ((f[..., None] - 0.5) * 4) ** 2
        feats.append(f.reshape(-1))
        tgts.append(t.reshape(-1, ounet.NOUT))
    feat = torch.tensor(np.array(feats), dtype=torch.float32, device=DEV)
    tgt = torch.tensor(np.array(tgts), dtype=torch.float32, device=DEV).to(torch.bfloat16)
    net = learned.UNet(n=n, max_batch=B, device=DEV, seed=0)
    first = None
    for k in range(200):
        i = (k * B) % 8
        loss = net.train_step(feat[i:i + B], tgt[i:i + B], 2e-3)
        first = loss if first is None else first
    final = learned.evaluate(net, feat, tgt, B)
    print(f"loss {first:.4e} -> {final:.4e}", flush=True)
    assert final < 0.2 * first


def test_start_selection_never_worse_than_zero():
    """A supplied start is used only where its residual beats a zero start; a
    garbage start falls back to zero and the result is unchanged."""
    import maple_amd
    r = 16
    print(f"[start selection] r={r}", flush=True)
    rng = np.random.default_rng(3)
    rho = torch.from_numpy(1e-9 + (1 - 1e-9) * rng.uniform(0, 1, r ** 3) ** 3).to(DEV)
    cb = maple_amd.CellBatch([r], nlevels=3, device=DEV)
    CH0, k0, it0 = cb.homogenize(rho, rtol=1e-8)
    it0 = it0.copy()
    assert (cb.start_sources("elastic") == 0).all()
    ue, ut = cb.fields("elastic"), cb.fields("thermal")
    # garbage: zero is better everywhere
    cb.set_start("elastic", 1e3 * torch.randn_like(ue))
    cb.set_start("thermal", 1e3 * torch.randn_like(ut))
    CH1, k1, it1 = cb.homogenize(rho, rtol=1e-8)
    assert (cb.start_sources("elastic") == 0).all() and (cb.start_sources("thermal") == 0).all()
    np.testing.assert_array_equal(it1, it0)
    assert torch.equal(CH1, CH0) and torch.equal(k1, k0)
    # the converged fields: the start wins and CG stops at once
    cb.set_start("elastic", ue)
    cb.set_start("thermal", ut)
    CH2, k2, it2 = cb.homogenize(rho, rtol=1e-8)
    assert (cb.start_sources("elastic") == 2).all() and (cb.start_sources("thermal") == 2).all()
    assert it2.max() <= 1
    assert torch.linalg.norm(CH2 - CH0) <= 1e-9 * torch.linalg.norm(CH0)


def test_delta_start_is_fields_plus_predicted_change():
    """LearnedStart(mode="delta"): the first call supplies the prediction; later calls
    supply the previous solve's fields plus the predicted change between the two
    densities, and the solve still converges to the same C_H as a zero start."""
    import maple_amd
    from maple_amd import learned as ml
    from oracle import unet as ounet
    r = 32
    rng = np.random.default_rng(5)
    theta = ounet.bf16_round(ounet.init_params(2)).astype(np.float32)
    net = ml.UNet(n=16, max_batch=1, device=DEV, theta=theta)
    cb = maple_amd.CellBatch([r], nlevels=3, device=DEV)
    ls = ml.LearnedStart(cb, net, mode="delta")
    rho1 = torch.from_numpy(1e-9 + (1 - 1e-9) * rng.uniform(0.05, 1, r ** 3) ** 3).to(DEV)
    rho2 = (rho1 * torch.from_numpy(rng.uniform(0.9, 1.1, r ** 3)).to(DEV)).clamp(1e-9, 1.0)
    ls(rho1, ("elastic",))
    cb.homogenize(rho1, ("elastic",), rtol=1e-8)
    u1 = cb.fields("elastic")
    p1, _ = ls.predict(rho1)
    p2, _ = ls.predict(rho2)
    ls(rho2, ("elastic",))
    # the candidate the solve sees: the converged fields of rho1 plus p2 - p1
    cand = u1 + (p2 - p1)
    cb2 = maple_amd.CellBatch([r], nlevels=3, device=DEV)
    cb2.set_start("elastic", cand)
    CHa, _, _ = cb2.homogenize(rho2, ("elastic",), rtol=1e-8)
    CHb, _, _ = cb.homogenize(rho2, ("elastic",), rtol=1e-8)
    src_a, src_b = cb2.start_sources("elastic"), cb.start_sources("elastic")
    np.testing.assert_array_equal(src_a, src_b)
    assert torch.linalg.norm(CHa - CHb) <= 1e-9 * torch.linalg.norm(CHb)
    cz = maple_amd.CellBatch([r], nlevels=3, device=DEV)
    CHz, _, _ = cz.homogenize(rho2, ("elastic",), rtol=1e-10)
    assert torch.linalg.norm(CHb - CHz) <= 1e-6 * torch.linalg.norm(CHz)
    for c in (cb, cb2, cz):
        c.close()
    net.close()


def test_prolong_matches_oracle():
    import maple_amd
    from oracle import mg
    r = 16
    cb = maple_amd.CellBatch([r], nlevels=3, device=DEV)
    rng = np.random.default_rng(4)
    coarse = rng.standard_normal((6, 3, (r // 2) ** 3))
    fine = cb.prolong("elastic", 1, torch.from_numpy(coarse).to(DEV)).cpu().numpy()
    P = mg.prolongation((r, r, r), (r // 2,) * 3, 1)
    ref = np.stack([np.stack([P @ coarse[j, c] for c in range(3)]) for j in range(6)])
    np.testing.assert_allclose(fine, ref, rtol=1e-14, atol=1e-14)


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "maple_amd", "data",
                                                    "warmstart_unet32.npz")), reason="no trained weights")
def test_trained_network_saves_iterations():
    """The committed network, on designs of a held-out 64^3 OC trajectory (seed and
    objective unlike its training runs): how far below a zero start its prediction
    starts the solves, and what that saves in CG iterations (never a loss)."""
    from maple_amd import learned
    from maple_amd.train_warmstart import measure_gain
    print("[learned gain] 64^3 held-out trajectory", flush=True)
    net = learned.UNet.load(max_batch=4, device=DEV)
    g = measure_gain(net, DEV, iters=41, probe_every=10)
    print(f"cold {g['cold_total_iters']} learned {g['learned_total_iters']}", flush=True)
    print(f"median start residual / zero-start residual {g['median_start_residual_ratio']:.3f}", flush=True)
    # the prediction starts every solve below the zero start's residual, the more so
    # once the design has formed (the residual is dominated by the 64^3 scales that a
    # 32^3 prediction cannot carry, so the margin is modest) ...
    ratios = np.array([r["start_residual_ratio"] for r in g["probes"]])
    assert (ratios < 1.0).all()
    assert ratios[-1].max() < 0.95
    # ... and the start selection makes it never cost iterations
    for row in g["probes"]:
        assert sum(row["learned"]) <= sum(row["cold"]) + 1
    assert g["learned_total_iters"] <= g["cold_total_iters"]
