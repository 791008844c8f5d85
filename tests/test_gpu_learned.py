"""GPU threshold model: kernels against the CPU oracle, exactness under learned thresholds, and what
the trained network gains over the scan without it."""

import numpy as np
import pytest

from costal_amd import _costal as ext
from costal_amd import learned, synth
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def test_gpu_gradient_matches_oracle():
    rng = np.random.default_rng(1)
    d, n = 40, 777
    x = synth.learned_draw(n, sample_seed=5)
    p = learned.init_params(d, seed=3)
    p[M.H * d:M.H * d + M.H] = rng.normal(0, 0.05, M.H).astype(np.float32)
    t = rng.normal(-4, 1, size=n).astype(np.float32)
    loss_gpu, g_gpu = learned.gpu_gradient(p, x, t)
    loss_cpu, g_cpu = M.loss_and_grad(p.astype(np.float64), x, t.astype(np.float64), d)
    # f32 kernels against the f64 oracle: 1e-4 of the gradient's scale
    scale = np.abs(g_cpu).max()
    np.testing.assert_allclose(g_gpu, g_cpu, rtol=1e-3, atol=1e-4 * scale)
    assert abs(loss_gpu - loss_cpu) <= 1e-5 * abs(loss_cpu)
    # deterministic: the same call gives the same bits
    loss2, g2 = learned.gpu_gradient(p, x, t)
    np.testing.assert_array_equal(g2, g_gpu)


def test_gpu_predict_matches_oracle_forward():
    x = synth.learned_draw(1000, sample_seed=6)
    p = learned.init_params(40, seed=4)
    y32 = learned.predict_raw(p, x, bf16=False)
    ycpu, _ = M.forward(p.astype(np.float64), x, 40)
    np.testing.assert_allclose(y32, ycpu, rtol=1e-5, atol=1e-5 * np.abs(ycpu).max())
    y16 = learned.predict_raw(p, x, bf16=True)
    assert np.abs(y16 - ycpu).max() <= 0.03 * np.abs(ycpu).max()


@pytest.mark.parametrize("d", [15, 40, 64])
def test_gpu_bf16_predict_matches_oracle_roundings(d):
    """The matrix-core BF16 inference (one to four 16-feature chunks in the first layer) against
    the oracle's restatement of the same roundings: the GPU sums in float32 in its own order, so a
    hidden value now and then rounds to the other bfloat16 neighbour; any slip in the operand
    layouts would be an error of the output's own size."""
    rng = np.random.default_rng(d)
    x = rng.normal(size=(3000, d)).astype(np.float32)
    x[7] = 0.0  # a zero row
    x[:1000] = np.abs(x[:1000])
    p = learned.init_params(d, seed=d)
    o = M.H * d  # nonzero biases, so that every bias reaches its unit
    p[o:o + M.H] = rng.normal(0, 0.2, M.H).astype(np.float32)
    p[o + M.H + M.H * M.H:o + 2 * M.H + M.H * M.H] = rng.normal(0, 0.2, M.H).astype(np.float32)
    y16 = learned.predict_raw(p, x, bf16=True)
    want = M.forward_bf16(p, x, d)
    err = np.abs(y16 - want)
    scale = np.abs(want).max()
    assert scale > 0.1
    assert np.median(err) <= 2e-6 * scale
    assert err.max() <= 2e-2 * scale
    assert (err <= 1e-5 * scale).mean() >= 0.9


def test_training_reduces_loss():
    x = synth.learned_draw(20000, sample_seed=7)
    rng = np.random.default_rng(0)
    target = (x[:, :3].sum(axis=1) * 0.3 - 4).astype(np.float32)
    p, losses = learned.fit(x, target, epochs=8, batch=2048, lr=3e-3, seed=0)
    assert losses[-1] < 0.25 * losses[0]


@pytest.fixture(scope="module")
def rare_model():
    # trained on other cells of the same mixture; serves 60,000 cells
    xt = synth.rare_396460x15(n=30000, sample_seed=11)
    return learned.train(xt, 10, n_target=60000, epochs=30, batch=512, seed=0)


@pytest.mark.parametrize("matmul", ["mfma-f32", "mfma-f16x3", "mfma-f16"])
def test_learned_thresholds_stay_exact(rare_model, matmul, oracle):
    x = synth.rare_396460x15(n=60000)
    got, st = ext._knn_exact(x, 10, matmul=matmul, threshold_model=rare_model)
    (widx, wcos, wnorms), _ = oracle.knn_blocked(x, 10, sort_by_norm=False)
    np.testing.assert_array_equal(got[0], widx)
    np.testing.assert_array_equal(got[1], wcos)
    np.testing.assert_array_equal(got[2], wnorms)
    # the model's calibration puts most rows above their threshold
    assert st["fallback_rows"] < 0.1 * len(x)


def test_learned_thresholds_cut_inserts(rare_model):
    """The gain of the trained network: far fewer candidate-list inserts, same result."""
    x = synth.rare_396460x15(n=60000)
    plain, st0 = ext._knn_exact(x, 10, matmul="mfma-f16x3")
    lrn, st1 = ext._knn_exact(x, 10, matmul="mfma-f16x3", threshold_model=rare_model)
    np.testing.assert_array_equal(plain[0], lrn[0])
    np.testing.assert_array_equal(plain[1], lrn[1])
    print("inserts plain", st0["inserts"], "learned", st1["inserts"], "fallback", st1["fallback_rows"])
    # (about 0.54 at this size with lists of k: the thresholds remove the list-filling inserts)
    assert st1["inserts"] < 0.6 * st0["inserts"]


def test_bad_model_still_exact(oracle):
    """A model whose thresholds are far too high sends every row to the plain scan: still exact."""
    x = synth.rare_396460x15(n=8000)
    p = learned.init_params(15, seed=0) * 0
    bad = learned.ThresholdModel(d=15, k=10, params=p, offset=-30.0)  # threshold ~ 1
    got, st = ext._knn_exact(x, 10, threshold_model=bad)
    want = oracle.knn_bruteforce(x, 10)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)
    assert st["fallback_rows"] == len(x) and st["short_rows"] == len(x)
    # the threshold-free rescan settles nearly every row (in the similarity-grouped working order a
    # row's nearest neighbors can share one candidate tile, hence one split's lists)
    assert st["rowscan_rows"] <= len(x) // 1000
