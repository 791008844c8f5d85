"""Batched samples: every sample's result equals its own single call and the oracle, the captured
graph replays to the same bits, fallbacks and errors are per sample."""

import os

import numpy as np
import pytest

from costal_amd import _costal as ext
from costal_amd import batched, synth

pytestmark = pytest.mark.gpu


def small_batch():
    xs = synth.batched_64(n_samples=7, n_lo=300, n_hi=5000)
    # a tie-heavy sample (fallbacks) and a tiny one (K' = n - 1)
    rng = np.random.default_rng(5)
    ties = np.floor(rng.uniform(0, 2, size=(2000, 40))).astype(np.float32)
    ties[:, 4:] = 0.0  # 16 distinct rows, ~125 copies each: tie groups longer than any list
    xs.append(ties)
    xs.append(synth.clustered(20, 33, 1))
    return xs


@pytest.mark.parametrize("matmul", ["mfma-f16x3", "mfma-f32"])
def test_batched_equals_single_calls(matmul, oracle):
    print("graph mode:", "off" if os.environ.get("COSTAL_NO_GRAPH") else "on", flush=True)
    xs = small_batch()
    k = 15
    b = batched.Batch(xs, k)
    st = b.run(matmul)
    outs = b.outputs()
    for x, (idx, cos, norms, w) in zip(xs, outs):
        want = oracle.knn_blocked(x, k, sort_by_norm=False)[0] if len(x) > 100 else oracle.knn_bruteforce(x, k)
        np.testing.assert_array_equal(idx, want[0])
        np.testing.assert_array_equal(cos, want[1])
        np.testing.assert_array_equal(norms, want[2])
        np.testing.assert_allclose(w, oracle.reweight(want[0], want[1], want[2], "tanimoto"), rtol=1e-12, atol=0)
    assert st.fallback_rows > 0  # the tie-heavy sample
    # replay of the captured graph: same bits
    b.run(matmul)
    for a, c in zip(outs, b.outputs()):
        for u, v in zip(a, c):
            np.testing.assert_array_equal(u, v)


def test_batched_per_sample_tuning_knobs_do_not_change_output():
    xs = small_batch()[:5]
    k = 15
    base = batched.Batch(xs, k)
    base.run()
    ref = base.outputs()
    tuned = batched.Batch(xs, k, margins=[5, 15, 30, 2, 40], col_splits=[1, 3, 2, 8, 1])
    tuned.run()
    for a, c in zip(ref, tuned.outputs()):
        np.testing.assert_array_equal(a[0], c[0])
        np.testing.assert_array_equal(a[1], c[1])


def test_batched_reports_non_finite_sample():
    xs = [synth.clustered(400, 33, 1), synth.clustered(500, 40, 2)]
    xs[1][77, 3] = np.inf
    b = batched.Batch(xs, 10)
    with pytest.raises(ValueError, match="sample 1: row 77 contains NaN or inf"):
        b.run()
