"""The five benchmark workloads at their full sizes on the GPU path, checked against the oracle.

Model: the reference's own exactness tests (ike12th/CosTaL tests/test_knn.py:38-61,
crates/costal-core/src/knn.rs:960-999): the exact path equals the brute force bit for bit.  Where
the full brute force would take minutes on the host, a contiguous sample of query rows is checked
against the oracle's blocked form (which restates knn.rs:403-512 and equals the brute force), and
size-independent properties are checked on every row: self never a neighbor, scores
non-increasing, ties ordered by index, indices in range, and different GPU forms (FP32 tiles, split
FP16 tiles, learned thresholds) equal on every row.
"""

import numpy as np
import pytest

import costal_amd as cp
from costal_amd import _costal as ext
from costal_amd import batched, learned, synth, tuner

pytestmark = pytest.mark.gpu


def assert_rows(idx, cos, n):
    """Properties that hold on every row of an exact kNN."""
    assert idx.min() >= 0 and idx.max() < n
    assert (idx != np.arange(n)[:, None]).all()
    assert (np.diff(cos, axis=1) <= 0).all()
    tie = np.diff(cos, axis=1) == 0
    assert (np.diff(idx, axis=1)[tie] > 0).all()


def assert_sample(idx, cos, norms, want, r0, r1):
    widx, wcos, wnorms = want
    np.testing.assert_array_equal(idx[r0:r1], widx[r0:r1])
    np.testing.assert_array_equal(cos[r0:r1], wcos[r0:r1])
    np.testing.assert_array_equal(norms, wnorms)


@pytest.mark.timeout(300)
def test_counts_3605x2000_full_vs_bruteforce(oracle):
    """C1: the float64 counts matrix through the public knn() (rounded to float32 once at the
    boundary, _api.py:89-100), every row against the oracle's brute force."""
    x64 = synth.counts_3605x2000()
    g = cp.knn(x64, 10)
    x32 = np.ascontiguousarray(x64, dtype=np.float32)
    widx, wcos, wnorms = oracle.knn_bruteforce(x32, 10)
    np.testing.assert_array_equal(g.idx, widx)
    np.testing.assert_array_equal(g.cos, wcos)
    np.testing.assert_array_equal(g.norms, wnorms)
    w = cp.reweight(g, "tanimoto")
    np.testing.assert_allclose(w, oracle.reweight(widx, wcos, wnorms, "tanimoto"), rtol=1e-12, atol=0)
    st = g.meta["stats"]
    print(f"C1 path={st['path']} fallback_rows={st['fallback_rows']} rowscan_rows={st['rowscan_rows']}")


@pytest.mark.timeout(600)
def test_batched_64_full_with_tuner_plan(oracle):
    """C4: all 64 samples (log-uniform sizes in [2000, 60000], d in 32..48, k = 15) in one batched
    call under the tuner's plan (tuned here on the GPU, per shape bucket), split FP16 and FP32
    tiles; every sample: properties on every row, the two forms equal, and up to 1,000 rows
    against the oracle."""
    xs = synth.batched_64()
    k = 15
    tn = tuner.Tuner(None)
    prec, margins, splits = tn.plan_batch([x.shape for x in xs], k)
    print(f"tuner: precision {prec}, {len(tn.table)} buckets", flush=True)
    b = batched.Batch(xs, k, margins=margins, col_splits=splits)
    st = b.run({"fp16x3": "mfma-f16x3", "fp32": "mfma-f32", "fp16": "mfma-f16"}[prec])
    outs = b.outputs()
    other = batched.Batch(xs, k)
    other.run("mfma-f32" if prec != "fp32" else "mfma-f16x3")
    outs2 = other.outputs()
    rng = np.random.default_rng(64)
    total = 0
    for s, (x, (idx, cos, norms, w), o2) in enumerate(zip(xs, outs, outs2)):
        n = len(x)
        assert_rows(idx, cos, n)
        np.testing.assert_array_equal(idx, o2[0])
        np.testing.assert_array_equal(cos, o2[1])
        m = min(1000, n)
        r0 = int(rng.integers(0, n - m + 1))
        want, _ = oracle.knn_blocked(x, k, sort_by_norm=False, rows=(r0, r0 + m))
        assert_sample(idx, cos, norms, want, r0, r0 + m)
        np.testing.assert_allclose(w[r0:r0 + m], oracle.reweight(want[0], want[1], want[2], "tanimoto")[r0:r0 + m],
                                   rtol=1e-12, atol=0)
        total += n
    print(f"C4: {total} cells, fallback_rows={st.fallback_rows}, rowscan_rows={st.rowscan_rows}, "
          f"fallback rate {st.fallback_rows / total:.5f}", flush=True)


@pytest.fixture(scope="module")
def c5_model():
    """The C5 threshold model: trained on 200,000 cells of the same mixture at another sampling
    seed, labels from the exact method itself."""
    xt = synth.learned_train_x40(n=200_000)
    m = learned.train(xt, 30, n_target=1_000_000, seed=0)
    print(f"C5 model: label rank {m.meta['k_label']}, held-out RMSE {m.meta['holdout_rmse']:.4f}, "
          f"fit {m.meta['fit_seconds']:.1f} s", flush=True)
    return m


@pytest.mark.timeout(900)
def test_learned_1000000x40_k30(c5_model, oracle):
    """C5: 1,000,000 x 40, k = 30, split FP16 tiles with learned admission thresholds: equal on every
    row to split FP16 without thresholds and to FP32 tiles, 1,500 rows equal to the oracle; the
    fallback rate (rows the thresholds sent to the plain scan) is reported and bounded."""
    x = synth.learned_1000000x40()
    n, k = len(x), 30
    (idx, cos, norms), st = ext._knn_exact(x, k, matmul="mfma-f16x3", threshold_model=c5_model, weights=True)
    assert_rows(idx, cos, n)
    rate = st["fallback_rows"] / n
    print(f"C5 learned: fallback_rows={st['fallback_rows']} ({rate:.4%}), short_rows={st['short_rows']}, "
          f"rowscan_rows={st['rowscan_rows']}, inserts={st['inserts']}, device {st['device_ms']:.1f} ms", flush=True)
    assert rate < 0.05
    (i2, c2, _), s2 = ext._knn_exact(x, k, matmul="mfma-f16x3")
    np.testing.assert_array_equal(idx, i2)
    np.testing.assert_array_equal(cos, c2)
    print(f"C5 plain split fp16: inserts={s2['inserts']}, device {s2['device_ms']:.1f} ms", flush=True)
    (i3, c3, _), _ = ext._knn_exact(x, k, matmul="mfma-f32")
    np.testing.assert_array_equal(idx, i3)
    np.testing.assert_array_equal(cos, c3)
    r0 = 123_456
    want, _ = oracle.knn_blocked(x, k, sort_by_norm=False, rows=(r0, r0 + 1500))
    assert_sample(idx, cos, norms, want, r0, r0 + 1500)
    w = st["weights"]
    np.testing.assert_allclose(w, oracle.reweight(idx, cos, norms, "tanimoto"), rtol=1e-12, atol=0)


def test_tuner_plan_of_one_bucket_is_exact(oracle):
    """The autotuner times its grid on the GPU for one shape bucket (precision, list margin,
    splits, then k-means center counts and the unpruned scan); the plan it keeps gives the oracle's
    result (tuning knobs never change the output)."""
    tn = tuner.Tuner(None)
    x = synth.cytometry_86864x39(n=6000)
    best = tn.get(*x.shape, 10)
    print("tuned:", best, flush=True)
    assert set(best) >= set(tuner.KNOBS)
    trials = tn.table[tuner.bucket(*x.shape, 10)]["trials"]
    assert any(not t["prune"] for t in trials) and any(t["prune_clusters"] for t in trials)
    got, _ = ext._knn_exact(x, 10, **tuner.knn_kwargs(best))
    want = oracle.knn_bruteforce(x, 10)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("matmul", ["mfma-f32", "mfma-f16x3"])
def test_nan_thresholds_admit_everything(matmul, oracle):
    """A NaN admission threshold is taken as no threshold; +inf admits nothing (the row rescans
    without one); the result stays exact."""
    x = synth.rare_396460x15(n=5000)
    want = oracle.knn_bruteforce(x, 10)
    th = np.full(len(x), np.nan, np.float32)
    th[::3] = np.inf
    th[1::7] = -np.inf
    got, st = ext._knn_exact(x, 10, matmul=matmul, row_threshold=th)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)
    assert st["short_rows"] >= int(np.isposinf(th).sum())


@pytest.mark.parametrize("matmul", ["mfma-f32", "mfma-f16x3"])
def test_large_caller_splits_are_capped(matmul, oracle):
    """col_splits far beyond the rescoring's capacity (2048 list entries per row): the plan caps
    it and the result is exact."""
    x = synth.rare_396460x15(n=200_000)
    got, st = ext._knn_exact(x, 10, matmul=matmul, col_splits=512)
    assert st["col_splits"] * 2 * st["candidates"] <= 2048
    (widx, wcos, wnorms), _ = oracle.knn_blocked(x, 10, sort_by_norm=False, rows=(0, 1000))
    np.testing.assert_array_equal(got[0][:1000], widx[:1000])
    np.testing.assert_array_equal(got[1][:1000], wcos[:1000])
    assert_rows(got[0], got[1], len(x))


def test_non_finite_row_is_a_value():
    """KnnError::NonFinite{row} (knn.rs:102-127): the row is returned as a value, not only as text."""
    from costal_amd import _lib

    x = synth.clustered(300, 6, 3)
    x[41, 2] = np.inf
    x[97, 0] = np.nan
    with pytest.raises(_lib.CostalError, match="row 41 contains NaN or inf") as e:
        ext._knn_exact(x, 5)
    assert e.value.code == _lib.E_NON_FINITE and e.value.row == 41
    xs = [synth.clustered(400, 33, 1), synth.clustered(500, 40, 2)]
    xs[1][77, 3] = np.nan
    b = batched.Batch(xs, 10)
    with pytest.raises(_lib.CostalError) as e:
        b.run()
    assert (e.value.row, e.value.sample) == (77, 1)
