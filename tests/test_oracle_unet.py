"""Pins the FP64 warm-start network restatement: periodic convolution against
scipy.ndimage.correlate (mode "wrap"), pooling and upsampling against explicit
loops, every gradient against central finite differences, bf16 rounding against
the float32 bit pattern rounding rule."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import unet


def test_conv_matches_ndimage_correlate():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 6, 6, 6, 3))
    W = rng.standard_normal((4, 27, 3))
    b = rng.standard_normal(4)
    y = unet.conv(x, W, b)
    for bi in range(2):
        for co in range(4):
            ref = b[co] + sum(ndimage.correlate(x[bi, ..., ci], W[co, :, ci].reshape(3, 3, 3), mode="wrap")
                              for ci in range(3))
            np.testing.assert_allclose(y[bi, ..., co], ref, rtol=1e-12, atol=1e-12)
    # 1x1 convolution is a per-voxel matrix product
    W1 = rng.standard_normal((5, 1, 3))
    np.testing.assert_allclose(unet.conv(x, W1, np.zeros(5)), x @ W1[:, 0, :].T, rtol=1e-13)


def test_pool_and_up_against_loops():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 4, 4, 4, 2))
    p = unet.pool(x)
    for i in range(2):
        for j in range(2):
            for k in range(2):
                np.testing.assert_allclose(p[0, i, j, k], x[0, 2 * i:2 * i + 2, 2 * j:2 * j + 2, 2 * k:2 * k + 2].mean(axis=(0, 1, 2)))
    u = unet.up(p)
    for i in range(4):
        for j in range(4):
            for k in range(4):
                np.testing.assert_array_equal(u[0, i, j, k], p[0, i // 2, j // 2, k // 2])
    # adjoint identities <pool x, g> = <x, pool^T g>, <up x, g> = <x, up^T g>
    g = rng.standard_normal(p.shape)
    assert np.isclose((p * g).sum(), (x * unet.pool_backward(g)).sum())
    g2 = rng.standard_normal(x.shape)
    assert np.isclose((unet.up(p) * g2).sum(), (p * unet.up_backward(g2)).sum())


def test_conv_backward_finite_differences():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((1, 4, 4, 4, 2))
    W = rng.standard_normal((3, 27, 2))
    b = rng.standard_normal(3)
    g = rng.standard_normal((1, 4, 4, 4, 3))
    dx, dW, db = unet.conv_backward(x, W, g)
    f = lambda x_, W_, b_: float((unet.conv(x_, W_, b_) * g).sum())
    h = 1e-6
    for _ in range(6):
        i = tuple(rng.integers(0, s) for s in x.shape)
        e = np.zeros_like(x)
        e[i] = h
        assert np.isclose((f(x + e, W, b) - f(x - e, W, b)) / (2 * h), dx[i], rtol=1e-6, atol=1e-8)
        j = tuple(rng.integers(0, s) for s in W.shape)
        e = np.zeros_like(W)
        e[j] = h
        assert np.isclose((f(x, W + e, b) - f(x, W - e, b)) / (2 * h), dW[j], rtol=1e-6, atol=1e-8)
    assert np.allclose(db, g.sum(axis=(0, 1, 2, 3)))


@pytest.mark.parametrize("weighted", [False, True])
def test_network_gradient_finite_differences(weighted):
    theta = unet.init_params(3)
    assert theta.size == unet.num_params()
    rng = np.random.default_rng(4)
    feat = rng.uniform(0, 1, (2, 4, 4, 4))
    target = rng.standard_normal((2, 4, 4, 4, unet.NOUT))
    L, g = unet.backward(theta, feat, target, weighted=weighted)
    assert np.isclose(L, unet.loss(theta, feat, target, weighted=weighted))
    h = 1e-6
    idx = rng.choice(theta.size, 40, replace=False)
    for k in idx:
        e = np.zeros_like(theta)
        e[k] = h
        fd = (unet.loss(theta + e, feat, target, weighted=weighted)
              - unet.loss(theta - e, feat, target, weighted=weighted)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * max(1e-3, abs(g[k])) + 1e-9, (k, fd, g[k])


def test_bf16_round():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0, 1e-30, 3.0e38])
    r = unet.bf16_round(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1); 1 + 3*2^-8 rounds up to 1 + 2^-6
    np.testing.assert_array_equal(r[:4], [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6])
    assert r[4] == -2.5 and r[5] == 0.0
    rel = np.abs(r[6:] - vals[6:]) / vals[6:]
    assert (rel <= 2 ** -8).all()
    # every rounded value has its low 16 float32 bits clear
    x = np.random.default_rng(5).standard_normal(1000)
    bits = unet.bf16_round(x).astype(np.float32).view(np.uint32)
    assert (bits & 0xFFFF == 0).all()
    assert (np.abs(unet.bf16_round(x) - x) <= 2 ** -8 * np.abs(x)).all()


def test_parameter_count_and_layout():
    n = unet.num_params()
    assert 3.0e5 < n < 6.0e5
    theta = np.arange(n, dtype=np.float64)
    np.testing.assert_array_equal(unet.flatten(unet.unflatten(theta)), theta)
