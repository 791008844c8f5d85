"""CPU restatement of the learned-threshold model (d -> 64 -> 64 -> 1, ReLU), in float64 numpy.

TEST INFRASTRUCTURE ONLY: the checker for the GPU forward/backward kernels in
costal_amd/csrc/mlp.hip (tests/ import it; the product path never does).  This part has no
counterpart in the reference (it is the admission-threshold model of the fused scan), so it is
pinned by mathematics instead: its gradients are checked against central finite differences on
small instances (tests/test_mlp_oracle.py), which a wrong transpose, sign or missing ReLU mask fails.
forward_bf16 (the BF16 inference's roundings) is pinned by round-to-nearest-even on halfway values
and by networks whose every value is exact in bfloat16, where it must equal forward() bit for bit.

Parameter layout (flat): W1[64][d], b1[64], W2[64][64], b2[64], w3[64], b3.
Features: x / |x| with the f64 sequential norm, rounded to float32 (zero rows -> zeros).
Loss: mean over rows of (y - target)^2.
"""

from __future__ import annotations

import numpy as np

H = 64


def param_count(d: int) -> int:
    return H * d + H + H * H + H + H + 1


def unpack(p: np.ndarray, d: int):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * d].reshape(H, d); o += H * d
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def features(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.float32)
    x64 = x.astype(np.float64)
    nrm = np.sqrt(np.cumsum(x64 * x64, axis=1)[:, -1]) if x.shape[1] else np.zeros(len(x))
    out = np.zeros_like(x64)
    nz = nrm > 0
    out[nz] = (x64[nz] / nrm[nz, None]).astype(np.float32)
    return out


def forward(p, x, d):
    W1, b1, W2, b2, w3, b3 = unpack(p, d)
    xn = features(x)
    z1 = xn @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ w3 + b3
    return y, (xn, z1, h1, z2, h2)


def bf16_round(a) -> np.ndarray:
    """Round-to-nearest-even to bfloat16, kept in float32 (the low 16 bits cleared)."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def forward_bf16(p, x, d):
    """forward() as the BF16 inference rounds it: features, W1, W2, w3 and both hidden layers'
    activations (after the ReLU) rounded to bfloat16; biases in float32; each layer's sum taken in
    float64 and rounded to float32 before the ReLU (the GPU accumulates in float32, in its own
    order, so the two agree to float32 rounding of the sums)."""
    W1, b1, W2, b2, w3, b3 = unpack(p, d)
    def r(a):
        return bf16_round(a).astype(np.float64)
    def f32(a):
        return np.asarray(a, dtype=np.float32).astype(np.float64)
    xn = r(features(x))
    h1 = r(np.maximum(f32(xn @ r(W1).T + b1), 0.0))
    h2 = r(np.maximum(f32(h1 @ r(W2).T + b2), 0.0))
    return f32(h2 @ r(w3) + b3)


def loss_and_grad(p, x, target, d):
    y, (xn, z1, h1, z2, h2) = forward(p, x, d)
    W1, b1, W2, b2, w3, b3 = unpack(p, d)
    n = len(y)
    err = y - np.asarray(target, dtype=np.float64)
    loss = float(np.mean(err * err))
    dy = 2.0 * err / n
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    dh2 = np.outer(dy, w3) * (z2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(axis=0)
    dh1 = (dh2 @ W2) * (z1 > 0)
    gW1 = dh1.T @ xn
    gb1 = dh1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, g


def init_params(d: int, seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases, as float32."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2.0 / d), size=(H, d))
    W2 = rng.normal(0, np.sqrt(2.0 / H), size=(H, H))
    w3 = rng.normal(0, np.sqrt(1.0 / H), size=H)
    return np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), w3, [0.0]]).astype(np.float32)
