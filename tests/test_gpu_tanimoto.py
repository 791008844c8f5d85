"""GPU parity of Tanimoto selection through the fused scan (select="tanimoto").

The reference ranks by T = a b c / (a^2 + b^2 - a b c) on raw rows (ike12th/CosTaL
crates/costal-core/src/knn.rs:197-284) and its exact path equals the brute force bit for bit
(knn.rs:1076-1147, tests/test_knn.py:118-148).  Here the scan scores normalized rows (T depends on
the cosine and the norm ratio only), keeps f32 Tanimoto scores in its lists, skips candidates
whose cosine is below the cut of the list's Tanimoto threshold for the tile's norm range, skips
tile pairs whose Tanimoto bound (cone cosine bound and norm ranges) is below every list's
threshold, and the f64 rescoring ranks by the f64 Tanimoto score.  Every case is checked against the oracle's f64
brute force (idx, cos, norms bitwise; weights to 1e-12 of the oracle's reweight, which the
Tanimoto score equals).  Cases follow the reference's tests: rows rescaled so that magnitude
matters, small signed integers with ties and zero rows, norms spanning 50x, plus the benchmark
mixtures, duplicates at other scales, and the row-scan fallback for norm spreads beyond f32.
"""

import numpy as np
import pytest

from costal_amd import _costal as ext
from costal_amd import _lib, synth

pytestmark = pytest.mark.gpu


def assert_same(got, want, what=""):
    for name, a, b in zip(("idx", "cos", "norms"), got, want):
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            raise AssertionError(f"{what}: {name} differs at {len(bad)} places, first {bad[:3].tolist()}")


def scaled(x, seed, lo=0.5, hi=3.0):
    """Rows rescaled so that magnitude matters and Tanimoto differs from cosine."""
    f = np.random.default_rng(seed).uniform(lo, hi, size=(len(x), 1)).astype(np.float32)
    return np.ascontiguousarray(x * f)


def run(x, k, **kw):
    got, st = ext._knn_exact(x, k, select="tanimoto", weights=True, **kw)
    print(f"  path={st['path']} fallback_rows={st['fallback_rows']} rowscan_rows={st['rowscan_rows']} "
          f"inserts={st['inserts']} coarse_err_max={st['coarse_err_max']:.3g} eps={st['coarse_eps']:.3g}", flush=True)
    return got, st


def check(x, k, oracle, what, **kw):
    (idx, cos, norms), st = run(x, k, **kw)
    want = oracle.knn_bruteforce(x, k, select="tanimoto")
    assert_same((idx, cos, norms), want, what)
    np.testing.assert_allclose(st["weights"], oracle.reweight(want[0], want[1], want[2], "tanimoto"), rtol=1e-12,
                               atol=0)
    # the measured list-score error stays within the bound the certificate uses
    assert st["coarse_err_max"] <= st["coarse_eps"]
    return st


@pytest.mark.parametrize("name", ["rare_20000x15", "cytometry_12000x39", "x40_15000_k30"])
def test_mixtures_scaled(name, oracle):
    x, k = {"rare_20000x15": (synth.rare_396460x15(n=20_000), 10),
            "cytometry_12000x39": (synth.cytometry_86864x39(n=12_000), 10),
            "x40_15000_k30": (synth.learned_1000000x40(n=15_000), 30)}[name]
    print(f"case {name}", flush=True)
    st = check(x, k, oracle, name)
    assert st["path"] == "fused"
    st = check(scaled(x, 5), k, oracle, name + " scaled")
    assert st["path"] == "fused"


def test_rescaled_clusters_differ_from_cosine(oracle):
    """tests/test_knn.py:118-125 and knn.rs:1077-1086: rescaled rows, Tanimoto != cosine."""
    x = scaled(synth.clustered(1500, 12, 14), 114)
    st = check(x, 10, oracle, "scaled clusters")
    assert st["select"] == "tanimoto" and st["path"] == "fused"
    assert not np.array_equal(ext._knn_exact(x, 10)[0][0], oracle.knn_bruteforce(x, 10, select="tanimoto")[0])


def test_signed_integers_ties_and_zero_rows(oracle):
    """knn.rs:1103-1117: small integers in [-1, 2) (negative scores, exact ties), zero rows."""
    rng = np.random.default_rng(7)
    x = np.floor(rng.random((400, 6)) * 3.0).astype(np.float32) - 1.0
    x[[0, 57, 399]] = 0.0
    check(x, 12, oracle, "signed integers")
    # the same at a size with many tiles
    x = np.floor(rng.random((9000, 6)) * 3.0).astype(np.float32) - 1.0
    x[rng.choice(9000, 40, replace=False)] = 0.0
    check(x, 12, oracle, "signed integers 9000")


def test_norms_spanning_50x(oracle):
    """knn.rs:1119-1137: norms spanning about 50x, so that whole tiles fall outside every row's
    norm window: the pruned scan skips tile pairs and stays exact."""
    x = synth.clustered(12000, 8, 9)
    x = np.ascontiguousarray(x * (1.0 + (np.arange(12000) % 50)).astype(np.float32)[:, None])
    skipped = {}
    for splits, clusters in ((0, 0), (1, 0), (1, 2)):
        st = check(x, 10, oracle, f"norms 50x, splits {splits}, centers {clusters}", col_splits=splits,
                   prune_clusters=clusters)
        skipped[splits, clusters] = st["blocks_skipped"] / st["blocks_total"]
        print(f"  splits {st['col_splits']}, centers {clusters}: skipped {st['blocks_skipped']} of "
              f"{st['blocks_total']} tile pairs", flush=True)
    # One candidate range per row (the reference's test has one; every split fills its own lists and
    # so prunes with lower thresholds).  The centers of this data lie in one orthant, so direction
    # groups barely separate rows: few groups leave long norm-sorted runs, the reference's layout.
    assert skipped[1, 0] > 0.10
    assert skipped[1, 2] > 0.25


def test_prune_off_equals_on(oracle):
    x = scaled(synth.rare_396460x15(n=20_000), 8)
    st = check(x, 10, oracle, "pruned")
    assert st["blocks_skipped"] > 0
    st = check(x, 10, oracle, "unpruned", prune=False)
    assert st["blocks_skipped"] == 0


@pytest.mark.parametrize("clusters", [1, 7, 64, 1000])
def test_center_count_never_changes_output(clusters, oracle):
    x = scaled(synth.cytometry_86864x39(n=10_000), 6)
    check(x, 10, oracle, f"{clusters} centers", prune_clusters=clusters)


@pytest.mark.parametrize("name", ["rare_20000x15", "cytometry_12000x39"])
def test_candidate_tiles_of_512_rows(name, oracle):
    x = {"rare_20000x15": synth.rare_396460x15(n=20_000), "cytometry_12000x39": synth.cytometry_86864x39(n=12_000)}[name]
    check(scaled(x, 12), 10, oracle, name + " tiles of 512", prune_tile=512)


def test_duplicates_at_other_scales(oracle):
    """Exact duplicates and copies scaled by 2 and 0.5 (cosine 1, Tanimoto 0.8 with the original)."""
    rng = np.random.default_rng(3)
    x = synth.cytometry_86864x39(n=6000).copy()
    rows = rng.permutation(len(x))
    src, dst = rows[:600], rows[600:1200]
    x[dst[:200]] = x[src[:200]]
    x[dst[200:400]] = x[src[200:400]] * np.float32(2.0)
    x[dst[400:]] = x[src[400:]] * np.float32(0.5)
    check(x, 10, oracle, "scaled duplicates")


@pytest.mark.parametrize("matmul", ["mfma-f32", "mfma-f16", "mfma-f16x3"])
def test_precision_knob_is_output_neutral(matmul, oracle):
    x = scaled(synth.rare_396460x15(n=8000), 9)
    check(x, 10, oracle, matmul, matmul=matmul)


@pytest.mark.parametrize("margin,splits", [(0, 1), (10, 4), (40, 0)])
def test_list_width_and_splits(margin, splits, oracle):
    x = scaled(synth.cytometry_86864x39(n=9000), 2)
    check(x, 10, oracle, f"margin {margin} splits {splits}", margin=margin, col_splits=splits)


def test_norm_spread_beyond_f32_takes_the_row_scan(oracle):
    """Norms more than 2^100 apart cannot share an f32 scale: the exact row scan answers."""
    x = scaled(synth.clustered(2000, 10, 3), 4)
    x[:5] *= np.float32(1e-30)
    x[5:10] *= np.float32(1e30)
    st = check(x, 10, oracle, "wide norm spread")
    assert st["path"] == "rowscan"


def test_thresholds_are_refused():
    x = synth.clustered(500, 8, 1)
    with pytest.raises(_lib.CostalError, match="Tanimoto selection takes none"):
        ext._knn_exact(x, 5, select="tanimoto", row_threshold=np.zeros(500, np.float32))
