"""The threshold model's CPU oracle, pinned by central finite differences (float64).

A wrong transpose, sign, missing ReLU mask or a loss-scale slip in the oracle's backward fails this;
the GPU kernels are then checked against the oracle (tests/test_gpu_learned.py)."""

import numpy as np

from oracle import mlp_oracle as M


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    d, n = 7, 9
    x = rng.normal(size=(n, d)).astype(np.float32)
    x[3] = 0.0  # a zero row
    p = M.init_params(d, seed=1).astype(np.float64)
    p[M.H * d:M.H * d + M.H] = rng.normal(0, 0.1, M.H)  # nonzero biases
    t = rng.normal(size=n)
    loss, g = M.loss_and_grad(p, x, t, d)
    assert g.shape == (M.param_count(d),)
    h = 1e-6
    idx = rng.choice(len(p), 60, replace=False)
    idx = np.concatenate([idx, [len(p) - 1]])  # b3
    for i in idx:
        q = p.copy(); q[i] += h
        lp, _ = M.loss_and_grad(q, x, t, d)
        q[i] -= 2 * h
        lm, _ = M.loss_and_grad(q, x, t, d)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_features_are_unit_rows_rounded_to_f32():
    x = np.array([[3.0, 4.0], [0.0, 0.0], [1e-3, 0.0]], np.float32)
    f = M.features(x)
    np.testing.assert_array_equal(f[0], np.array([0.6, 0.8], np.float32).astype(np.float64))
    assert (f[1] == 0).all() and f[2, 0] == 1.0
