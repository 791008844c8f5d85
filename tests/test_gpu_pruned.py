"""GPU parity of the pruned fused scan: rows in a locality order (spherical k-means), candidate
tiles visited in decreasing cosine-bound order and skipped once the bound is below every list's
threshold.  The result must stay the f64 brute force bit for bit (ike12th/CosTaL
crates/costal-core/src/knn.rs:884-924): pruning, like the reference's own norm-window pruning
(knn.rs:603-644), may only leave out candidates that cannot qualify, so the certificate of
finalize_row (knn.rs:704-714) holds unchanged.  Cases: Gaussian mixtures of the benchmark kinds,
zero rows, exact duplicates, integer ties, every precision form, admission thresholds, the number
of k-means centers, and pruning switched off.
"""

import numpy as np
import pytest

from costal_amd import _costal as ext
from costal_amd import synth

pytestmark = pytest.mark.gpu

BACKENDS = ["mfma-f32", "mfma-f16", "mfma-f16x3"]


def assert_same(got, want, what=""):
    for name, a, b in zip(("idx", "cos", "norms"), got, want):
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            raise AssertionError(f"{what}: {name} differs at {len(bad)} places, first {bad[:3].tolist()}")


def mixtures():
    return {
        "rare_20000x15": (synth.rare_396460x15(n=20_000), 10),
        "cytometry_12000x39": (synth.cytometry_86864x39(n=12_000), 10),
        "x40_15000_k30": (synth.learned_1000000x40(n=15_000), 30),
    }


@pytest.fixture(scope="module")
def cases(oracle):
    out = {}
    for name, (x, k) in mixtures().items():
        out[name] = (x, k, oracle.knn_bruteforce(x, k))
    return out


@pytest.mark.parametrize("matmul", BACKENDS)
@pytest.mark.parametrize("name", ["rare_20000x15", "cytometry_12000x39", "x40_15000_k30"])
def test_pruned_scan_equals_bruteforce(name, matmul, cases):
    x, k, want = cases[name]
    print(f"case {name} {matmul}", flush=True)
    got, st = ext._knn_exact(x, k, matmul=matmul)
    print(f"  skipped {st['blocks_skipped']} of {st['blocks_total']} tile pairs, fallback_rows={st['fallback_rows']}",
          flush=True)
    assert_same(got, want, f"{name} {matmul}")
    assert st["blocks_total"] > 0
    if name == "rare_20000x15":
        # a mixture of 23 well-separated components: some tile pairs cannot hold a neighbor
        assert st["blocks_skipped"] > 0


@pytest.mark.parametrize("matmul", BACKENDS)
def test_prune_off_equals_on(matmul, cases):
    x, k, want = cases["rare_20000x15"]
    got_off, st_off = ext._knn_exact(x, k, matmul=matmul, prune=False)
    assert st_off["blocks_skipped"] == 0
    assert_same(got_off, want, "prune off")


@pytest.mark.parametrize("clusters", [1, 7, 64, 1000])
def test_center_count_never_changes_output(clusters, cases):
    x, k, want = cases["rare_20000x15"]
    got, _ = ext._knn_exact(x, k, prune_clusters=clusters)
    assert_same(got, want, f"{clusters} centers")


@pytest.mark.parametrize("tile", [256, 512])
@pytest.mark.parametrize("name", ["rare_20000x15", "cytometry_12000x39"])
def test_candidate_tile_rows_never_change_output(name, tile, cases):
    """Candidate tiles of 512 rows: one LDS tile (J <= 2) or four (J > 4) per candidate tile."""
    x, k, want = cases[name]
    got, st = ext._knn_exact(x, k, prune_tile=tile)
    assert_same(got, want, f"{name} tile {tile}")
    assert st["blocks_total"] == -(-len(x) // 128) * -(-len(x) // tile)


def test_candidate_tile_rows_are_checked():
    from costal_amd import _lib

    with pytest.raises(_lib.CostalError, match="prune_tile=300: need 0, 256 or 512"):
        ext._knn_exact(synth.clustered(600, 8, 1), 5, prune_tile=300)


@pytest.mark.parametrize("matmul", BACKENDS)
def test_zero_rows_duplicates_and_integer_ties(matmul, oracle):
    """Zero rows (cosine 0 with everything; sorted after every center), exact duplicates of other
    rows (ties broken by original index) and integer-valued rows (many exact score ties)."""
    rng = np.random.default_rng(11)
    x = synth.rare_396460x15(n=7000).copy()
    x[rng.choice(len(x), 60, replace=False)] = 0.0
    src = rng.choice(len(x), 300, replace=False)
    dst = rng.choice(len(x), 300, replace=False)
    x[dst] = x[src]
    ints = rng.choice(len(x), 1500, replace=False)
    x[ints] = rng.integers(0, 3, size=(1500, x.shape[1])).astype(np.float32)
    want = oracle.knn_bruteforce(x, 12)
    got, st = ext._knn_exact(x, 12, matmul=matmul)
    print(f"  {matmul}: skipped {st['blocks_skipped']} of {st['blocks_total']}, fallback_rows={st['fallback_rows']}, "
          f"rowscan_rows={st['rowscan_rows']}", flush=True)
    assert_same(got, want, f"zeros/duplicates/ties {matmul}")


@pytest.mark.parametrize("matmul", ["mfma-f32", "mfma-f16x3"])
def test_signed_centered_mixture(matmul, oracle):
    """Signed, centered data: clusters in opposite directions give tile pairs with negative cosine
    bounds (the bound order's buckets run from 1 down to -1)."""
    rng = np.random.default_rng(17)
    centers = rng.normal(0.0, 1.0, size=(12, 16))
    lab = rng.integers(0, 12, size=16_000)
    x = np.ascontiguousarray(np.concatenate([centers, -centers])[np.concatenate([lab[:8000], lab[8000:] + 12])]
                             + rng.normal(0.0, 0.3, size=(16_000, 16)), dtype=np.float32)
    x -= x.mean(axis=0, dtype=np.float64).astype(np.float32)
    want = oracle.knn_bruteforce(x, 10)
    got, st = ext._knn_exact(x, 10, matmul=matmul)
    print(f"  skipped {st['blocks_skipped']} of {st['blocks_total']}", flush=True)
    assert_same(got, want, f"signed {matmul}")
    assert st["blocks_skipped"] > 0


@pytest.mark.parametrize("matmul", ["mfma-f32", "mfma-f16x3"])
def test_pruned_scan_with_admission_thresholds(matmul, cases):
    """Thresholds start the pruning before any list is full: good ones (just below each row's k-th
    cosine), and bad ones (above it on a third of the rows: those rows rescan)."""
    x, k, want = cases["rare_20000x15"]
    kth = want[1][:, k - 1]
    good = (kth - 1e-4).astype(np.float32)
    got, st = ext._knn_exact(x, k, matmul=matmul, row_threshold=good)
    assert_same(got, want, "good thresholds")
    bad = good.copy()
    bad[::3] = np.minimum(kth[::3] + 0.01, 1.5).astype(np.float32)
    got, st = ext._knn_exact(x, k, matmul=matmul, row_threshold=bad)
    assert st["short_rows"] > 0
    assert_same(got, want, "bad thresholds")
