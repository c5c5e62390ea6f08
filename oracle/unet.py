"""FP64 restatement of the learned warm start's network (test infrastructure only).

The network maps a coarse density feature (one channel on a periodic n^3 grid)
to the 21 characteristic-field channels on the same grid (6 elastic cases x 3
components, then 3 thermal cases), each divided by a fixed per-channel scale.
It is a 3-level U-Net with 16-32-64 channels, every convolution periodic
(circular padding), as the warm-start model asked for in the learned form of
the periodic inverse-homogenization loop (no counterpart in the reference,
whose warm-start hook `init_list` is unused, EBE_GMG.py:648-757, model.py:120).

Layout (shared with the HIP kernels): activations channels-last (B, n, n, n, C)
with voxel index (x*n + y)*n + z; a conv weight is W[co, o, ci] with
o = (dx+1)*9 + (dy+1)*3 + (dz+1) for a 3^3 kernel (o = 0 for 1^3), and
y[v, co] = b[co] + sum_{o,ci} W[co, o, ci] * x[(v + off_o) mod n, ci].

Passing ``q=bf16_round`` emulates where the GPU path rounds to bfloat16 (the
input, every stored activation, the weights and every pre-activation gradient
fed to a matrix core); ``q=None`` is the exact FP64 network.

Pinned by tests/test_oracle_unet.py: conv against scipy.ndimage.correlate
(mode "wrap"), pooling/upsampling against explicit loops, and every gradient
against central finite differences of the FP64 loss.
"""
from __future__ import annotations

import numpy as np

NOUT = 21
# (name, inputs, cin, cout, ksize, relu, level); an input "up:x" is x upsampled x2 (nearest)
LAYERS = [
    ("e1a", ["feat"], 1, 16, 3, True, 0),
    ("e1b", ["e1a"], 16, 16, 3, True, 0),
    ("e2a", ["pool:e1b"], 16, 32, 3, True, 1),
    ("e2b", ["e2a"], 32, 32, 3, True, 1),
    ("b1", ["pool:e2b"], 32, 64, 3, True, 2),
    ("b2", ["b1"], 64, 64, 3, True, 2),
    ("d2a", ["up:b2", "e2b"], 96, 32, 3, True, 1),
    ("d2b", ["d2a"], 32, 32, 3, True, 1),
    ("d1a", ["up:d2b", "e1b"], 48, 16, 3, True, 0),
    ("d1b", ["d1a"], 16, 16, 3, True, 0),
    ("head", ["d1b"], 16, NOUT, 1, False, 0),
]


def offsets(k: int):
    if k == 1:
        return [(0, 0, 0)]
    return [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]


def param_shapes():
    """[(name, W shape (cout, k^3, cin), b shape (cout,))] in storage order."""
    return [(n, (co, k ** 3, ci), (co,)) for n, _, ci, co, k, _, _ in LAYERS]


def num_params() -> int:
    return sum(int(np.prod(w)) + b[0] for _, w, b in param_shapes())


def unflatten(theta: np.ndarray):
    """Flat parameter vector (per layer: W then b) -> {name: (W, b)}."""
    out = {}
    i = 0
    for n, ws, bs in param_shapes():
        nw = int(np.prod(ws))
        out[n] = (theta[i:i + nw].reshape(ws), theta[i + nw:i + nw + bs[0]])
        i += nw + bs[0]
    return out


def flatten(p) -> np.ndarray:
    return np.concatenate([np.concatenate([p[n][0].ravel(), p[n][1].ravel()]) for n, _, _ in param_shapes()])


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights (std sqrt(2 / fan_in)), zero biases, numpy default_rng(seed)."""
    rng = np.random.default_rng(seed)
    parts = []
    for n, ws, bs in param_shapes():
        fan_in = ws[1] * ws[2]
        std = np.sqrt(2.0 / fan_in) if n != "head" else np.sqrt(1.0 / fan_in)
        parts.append(rng.standard_normal(int(np.prod(ws))) * std)
        parts.append(np.zeros(bs[0]))
    return np.concatenate(parts)


def bf16_round(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _id(a):
    return np.asarray(a, dtype=np.float64)


# ---- layers ------------------------------------------------------------------

def conv(x, W, b):
    """x (B, n, n, n, Cin), W (Cout, K, Cin), b (Cout,) -> (B, n, n, n, Cout)."""
    k = 3 if W.shape[1] == 27 else 1
    y = np.broadcast_to(b, x.shape[:4] + (W.shape[0],)).copy()
    for o, (dx, dy, dz) in enumerate(offsets(k)):
        xs = np.roll(x, (-dx, -dy, -dz), axis=(1, 2, 3))
        y += xs @ W[:, o, :].T
    return y


def conv_backward(x, W, g):
    """Gradients of sum(g * conv(x, W, b)): (dx, dW, db)."""
    k = 3 if W.shape[1] == 27 else 1
    dx = np.zeros_like(x)
    dW = np.zeros_like(W)
    for o, (dx_, dy_, dz_) in enumerate(offsets(k)):
        xs = np.roll(x, (-dx_, -dy_, -dz_), axis=(1, 2, 3))
        dW[:, o, :] = g.reshape(-1, g.shape[-1]).T @ xs.reshape(-1, xs.shape[-1])
        dx += np.roll(g @ W[:, o, :], (dx_, dy_, dz_), axis=(1, 2, 3))
    return dx, dW, g.sum(axis=(0, 1, 2, 3))


def pool(x):
    B, n = x.shape[0], x.shape[1]
    h = n // 2
    return x.reshape(B, h, 2, h, 2, h, 2, -1).mean(axis=(2, 4, 6))


def pool_backward(g):
    return up(g) / 8.0


def up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def up_backward(g):
    B, n = g.shape[0], g.shape[1]
    h = n // 2
    return g.reshape(B, h, 2, h, 2, h, 2, -1).sum(axis=(2, 4, 6))


# ---- network -------------------------------------------------------------------

def _source(acts, spec):
    if spec.startswith("pool:"):
        return pool(acts[spec[5:]])
    if spec.startswith("up:"):
        return up(acts[spec[3:]])
    return acts[spec]


def forward(theta, feat, q=None):
    """feat (B, n, n, n) -> (raw output (B, n, n, n, 21), activations)."""
    q = q or _id
    p = unflatten(theta)
    acts = {"feat": q(feat[..., None])}
    inputs = {}
    for name, srcs, _, _, _, relu, _ in LAYERS:
        x = np.concatenate([q(_source(acts, s)) for s in srcs], axis=-1)
        inputs[name] = x
        W, b = p[name]
        y = conv(x, q(W), b)
        acts[name] = q(np.maximum(y, 0.0)) if relu else y
    return acts["head"], (acts, inputs)


def voxel_weight(feat, q=None, weighted=False):
    """Per-voxel loss weight: 1, or f^3 + 1e-3 of the (rounded) input feature."""
    if not weighted:
        return np.ones(feat.shape + (1,))
    f = (q or _id)(feat)
    return (f ** 3 + 1e-3)[..., None]


def loss(theta, feat, target, q=None, weighted=False):
    y, _ = forward(theta, feat, q)
    return float(np.mean(voxel_weight(feat, q, weighted) * (y - target) ** 2))


def backward(theta, feat, target, q=None, weighted=False):
    """Loss mean(w (y - t)^2) and its gradient with respect to the flat parameters."""
    w = voxel_weight(feat, q, weighted)
    q = q or _id
    p = unflatten(theta)
    y, (acts, inputs) = forward(theta, feat, q)
    L = float(np.mean(w * (y - target) ** 2))
    dact = {"head": 2.0 * w * (y - target) / y.size}
    grads = {}
    for name, srcs, _, _, _, relu, _ in reversed(LAYERS):
        g = dact.pop(name)
        if relu:
            g = g * (acts[name] > 0)
        g = q(g)
        W, _ = p[name]
        dx, dW, db = conv_backward(inputs[name], q(W), g)
        grads[name] = (dW, db)
        c0 = 0
        for s in srcs:
            if s == "feat":
                break
            src = s.split(":")[-1]
            c = (acts[src].shape[-1])
            part = dx[..., c0:c0 + c]
            c0 += c
            if s.startswith("pool:"):
                part = pool_backward(part)
            elif s.startswith("up:"):
                part = up_backward(part)
            dact[src] = dact.get(src, 0.0) + part
    return L, flatten(grads)
