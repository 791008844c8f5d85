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


def test_bf16_round_is_nearest_even():
    one = np.float32(1.0)
    cases = [
        (1.0, 1.0),
        (1.0 + 2.0 ** -8, 1.0),  # halfway between 1 and 1 + 2^-7: to the even one
        (1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -6),  # halfway between 1 + 2^-7 and 1 + 2^-6: to the even one
        (1.0 + 2.0 ** -8 + 2.0 ** -20, 1.0 + 2.0 ** -7),  # past halfway: up
        (-(1.0 + 3 * 2.0 ** -8), -(1.0 + 2.0 ** -6)),
        (0.0, 0.0),
        (3.0e38, 3.0040553e38),  # below the largest bfloat16 (3.3895e38): the nearest
        (float(np.finfo(np.float32).max), np.inf),  # past it by more than half an ulp: to infinity
    ]
    for v, want in cases:
        got = M.bf16_round(np.float32(v))
        assert got == np.float32(want), (v, got, want)
    assert M.bf16_round(one).dtype == np.float32


def _integer_net(d, seed):
    """Parameters and one-hot rows whose every intermediate value is a small integer (exact in
    bfloat16): W1, b1 in [-3, 3], W2 with four +-1 entries per row, b2, w3 in [-2, 2], b3 = 1."""
    rng = np.random.default_rng(seed)
    W1 = rng.integers(-3, 4, size=(M.H, d))
    b1 = rng.integers(-3, 4, size=M.H)
    W2 = np.zeros((M.H, M.H), dtype=np.int64)
    for j in range(M.H):
        W2[j, rng.choice(M.H, 4, replace=False)] = rng.choice([-1, 1], 4)
    b2 = rng.integers(-2, 3, size=M.H)
    w3 = rng.integers(-2, 3, size=M.H)
    p = np.concatenate([W1.ravel(), b1, W2.ravel(), b2, w3, [1]]).astype(np.float32)
    x = np.zeros((2 * d + 1, d), dtype=np.float32)
    for t in range(d):
        x[t, t] = 3.0  # unit rows after scaling
        x[d + t, t] = -0.5
    return p, x


def test_forward_bf16_equals_forward_where_everything_is_exact():
    for d, seed in [(5, 0), (15, 1), (40, 2)]:
        p, x = _integer_net(d, seed)
        want, _ = M.forward(p.astype(np.float64), x, d)
        got = M.forward_bf16(p, x, d)
        np.testing.assert_array_equal(got, want)
        assert np.abs(want).max() > 1.0  # the network is not trivially zero


def test_forward_bf16_rounds_each_stage():
    """One weight off bfloat16 by less than half an ulp: the BF16 forward sees the rounded weight
    (the f64 forward does not), so the output moves by the closed-form amount."""
    d = 5
    p, x = _integer_net(d, 3)
    p64 = p.astype(np.float64)
    base, (_, _, _, _, h2) = M.forward(p64, x, d)
    o = M.H * d + M.H + M.H * M.H + M.H  # w3
    # a unit active on some row with an integer output weight w (|w| <= 2, so its bfloat16 ulp is
    # at least 2^-6 and w + 2^-12 rounds back to w)
    active = np.flatnonzero((h2.max(axis=0) > 0) & (p[o:o + M.H] != 0))
    assert active.size
    g = int(active[0])
    q = p.copy()
    q[o + g] = np.float32(p[o + g] + 2.0 ** -12)
    np.testing.assert_array_equal(M.forward_bf16(q, x, d), M.forward_bf16(p, x, d))
    exact, _ = M.forward(q.astype(np.float64), x, d)
    np.testing.assert_allclose(exact - base, 2.0 ** -12 * h2[:, g], rtol=0, atol=1e-12)
    assert np.abs(exact - base).max() > 0
