"""GPU parity: the gfx950 library against the CPU oracle and the golden fixtures.

Indices, cosines and norms must be bit-identical to the f64 brute force (ike12th/CosTaL
crates/costal-core/src/knn.rs:884-924); Tanimoto weights within 1e-12 relative (bitwise expected).
Every case runs for both coarse precisions (plain FP32 MFMA tiles, FP16 MFMA candidates).
"""

import glob
import os

import numpy as np
import pytest

from costal_amd import _costal as ext
from costal_amd import synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BACKENDS = ["mfma-f32", "mfma-f16", "mfma-f16x3"]
W_RTOL = 1e-12


def assert_same(got, want, what=""):
    for name, a, b in zip(("idx", "cos", "norms"), got, want):
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            raise AssertionError(f"{what}: {name} differs at {len(bad)} places, first {bad[:3].tolist()}")


def golden_files():
    return sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


@pytest.mark.parametrize("matmul", BACKENDS)
@pytest.mark.parametrize("path", golden_files(), ids=lambda p: os.path.basename(p)[:-4])
def test_golden_fixtures(path, matmul):
    f = np.load(path)
    x, k = f["x"], int(f["k"])
    print(f"case {os.path.basename(path)} {matmul}", flush=True)
    (idx, cos, norms), st = ext._knn_exact(x, k, matmul=matmul, weights=True)
    assert_same((idx, cos, norms), (f["idx"], f["cos"], f["norms"]), os.path.basename(path))
    np.testing.assert_allclose(st["weights"], f["w"], rtol=W_RTOL, atol=0)


def test_bruteforce_kernel_matches_golden():
    for path in golden_files():
        f = np.load(path)
        got = ext._knn_bruteforce(f["x"], int(f["k"]))
        assert_same(got, (f["idx"], f["cos"], f["norms"]), os.path.basename(path))


@pytest.mark.parametrize("matmul", BACKENDS)
@pytest.mark.parametrize("seed", [0, 1])
def test_exact_matches_bruteforce_and_no_fallback(seed, matmul, oracle):
    # tests/test_knn.py:38-43.  Each row's candidates are split over two lists of k (one per lane
    # half), so a row whose k nearest mostly fall in one half fails the first certificate now and
    # then (about 1 row in 400) and is rescanned with longer lists; no row of well-separated data
    # needs the reference's whole-row fallback.
    x = synth.clustered(3000, 32, seed)
    got, stats = ext._knn_exact(x, 10, matmul=matmul)
    assert_same(got, oracle.knn_bruteforce(x, 10))
    assert stats["rowscan_rows"] == 0
    assert stats["fallback_rows"] <= 0.01 * len(x)


@pytest.mark.parametrize("matmul", BACKENDS)
def test_tie_heavy_integer_data_triggers_fallback_and_stays_exact(matmul, oracle):
    # tests/test_knn.py:87-94
    rng = np.random.default_rng(5)
    x = np.floor(rng.uniform(0, 3, size=(500, 6))).astype(np.float32)
    x[[0, 99, 499]] = 0.0
    got, stats = ext._knn_exact(x, 12, matmul=matmul)
    assert_same(got, oracle.knn_bruteforce(x, 12))
    assert stats["fallback_rows"] > 0


@pytest.mark.parametrize("matmul", BACKENDS)
def test_duplicates_break_ties_by_original_index(matmul):
    # tests/test_knn.py:78-84 and knn.rs:1001-1017
    rng = np.random.default_rng(4)
    protos = rng.normal(size=(10, 5)).astype(np.float32)
    x = np.ascontiguousarray(np.tile(protos, (30, 1)))
    (idx, cos, _), _ = ext._knn_exact(x, 10, matmul=matmul)
    np.testing.assert_array_equal(idx[0], np.arange(10, 101, 10))


@pytest.mark.parametrize("matmul", BACKENDS)
def test_k_close_to_n(matmul, oracle):
    # knn.rs:1034-1044
    x = synth.lcg_clustered(25, 4, 5)
    for k in (1, 23, 24):
        got, _ = ext._knn_exact(x, k, matmul=matmul)
        assert_same(got, oracle.knn_bruteforce(x, k), f"k={k}")


@pytest.mark.parametrize("matmul", BACKENDS)
def test_independent_numpy_oracle(matmul):
    # tests/test_knn.py:46-61
    x = synth.clustered(1500, 20, 2)
    (idx, cos, _), _ = ext._knn_exact(x, 12, matmul=matmul)
    x64 = x.astype(np.float64)
    xn = x64 / np.linalg.norm(x64, axis=1, keepdims=True)
    sim = xn @ xn.T
    np.fill_diagonal(sim, -np.inf)
    order = np.argsort(-sim, axis=1, kind="stable")[:, :13]
    want = np.take_along_axis(sim, order, axis=1)
    np.testing.assert_allclose(cos, want[:, :12], rtol=0, atol=1e-12)
    clear = (want[:, 11] - want[:, 12]) > 1e-9
    assert clear.mean() > 0.99
    for i in np.flatnonzero(clear):
        assert set(idx[i]) == set(order[i, :12])


@pytest.mark.parametrize("d", [15, 32])
def test_device_rows_at_any_float_offset(d, oracle):
    """The device entry point on the same rows starting on a 16-byte boundary and 4 bytes past one.
    The rescoring reads rows 16 bytes at a time: rows of 32 features straight from an aligned input,
    from a padded copy otherwise (15 features, or a misaligned start); the output is the host
    entry point's, bit for bit, which equals the brute force's."""
    import ctypes as C

    import torch

    from costal_amd import _lib

    x = np.ascontiguousarray(synth.cytometry_86864x39(n=4000)[:, :d])
    n, k = x.shape[0], 10
    want, _ = ext._knn_exact(x, k)
    assert_same(want, oracle.knn_bruteforce(x, k), f"d={d} host entry point")
    L = _lib.load()
    dev = torch.device("cuda:0")
    flat = torch.empty(n * d + 4, dtype=torch.float32, device=dev)
    for off in (0, 1):
        v = flat[off:off + n * d]
        v.copy_(torch.from_numpy(x.ravel()))
        assert (v.data_ptr() % 16 == 0) == (off == 0)
        idx = torch.empty((n, k), dtype=torch.int32, device=dev)
        cos = torch.empty((n, k), dtype=torch.float64, device=dev)
        norms = torch.empty(n, dtype=torch.float64, device=dev)
        p = _lib.KnnParams()
        p.k = k
        p.margin = -1
        p.precision = _lib.PRECISION["fp16x3"]
        st = _lib.KnnStats()
        torch.cuda.synchronize()
        _lib.check(L.costal_knn_exact_device(C.c_void_p(v.data_ptr()), n, d, C.byref(p), C.c_void_p(idx.data_ptr()),
                                             C.c_void_p(cos.data_ptr()), C.c_void_p(norms.data_ptr()), None,
                                             C.byref(st), C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        torch.cuda.synchronize()
        assert_same((idx.cpu().numpy().astype(np.uint32), cos.cpu().numpy(), norms.cpu().numpy()), want,
                    f"d={d} offset {4 * off} bytes")


@pytest.mark.parametrize("matmul", BACKENDS)
@pytest.mark.parametrize("splits", [0, 1, 3, 8])
def test_output_independent_of_splits(splits, matmul, oracle):
    # knn.rs:969-999, tests/test_knn.py:64-75: the candidate-range split is the GPU's blocking knob
    x = synth.clustered(1200, 17, 3)
    want = oracle.knn_bruteforce(x, 7)
    got, st = ext._knn_exact(x, 7, matmul=matmul, col_splits=splits)
    assert splits == 0 or st["col_splits"] <= splits
    assert_same(got, want)


@pytest.mark.parametrize("matmul", BACKENDS)
@pytest.mark.parametrize("margin", [0, 1, 5, 40])
def test_margin_never_changes_output(margin, matmul, oracle):
    x = synth.cytometry_86864x39(n=3000)
    got, st = ext._knn_exact(x, 10, matmul=matmul, margin=margin)
    assert_same(got, oracle.knn_bruteforce(x, 10), f"margin={margin}")


@pytest.mark.parametrize("matmul", BACKENDS)
def test_row_thresholds_keep_exactness(matmul, oracle):
    """Admission thresholds only skip inserts: too-high thresholds send rows to the fallback,
    never change the result."""
    x = synth.rare_396460x15(n=5000)
    want = oracle.knn_bruteforce(x, 10)
    rng = np.random.default_rng(0)
    kth = want[1][:, 9].astype(np.float32)
    for name, th in [("below", kth - 0.01), ("at", kth), ("above", kth + 0.01),
                     ("random", rng.uniform(-1, 1, size=len(x)).astype(np.float32))]:
        got, st = ext._knn_exact(x, 10, matmul=matmul, row_threshold=th)
        assert_same(got, want, f"thresholds {name}")
        if name == "above":
            assert st["short_rows"] == len(x)


@pytest.mark.parametrize("matmul", BACKENDS)
def test_wide_rows_counts_workload(matmul, oracle):
    # 3,605 x 2,000 counts-like workload, reduced to 600 rows for the brute force
    x = np.ascontiguousarray(synth.counts_3605x2000()[:600], dtype=np.float32)
    got, st = ext._knn_exact(x, 10, matmul=matmul)
    assert st["path"] == "wide"
    assert_same(got, oracle.knn_bruteforce(x, 10))


def test_wide_rows_two_score_blocks(oracle):
    """12,000 x 80: more query rows than one 512 MB block of stored scores holds, so the wide path
    runs two blocks; equal to the exact GPU brute force everywhere and to the oracle on a sample."""
    x = synth.clustered(12000, 80, 3)
    got, st = ext._knn_exact(x, 10)
    assert st["path"] == "wide"
    assert_same(got, ext._knn_bruteforce(x, 10))
    (widx, wcos, wnorms), _ = oracle.knn_blocked(x, 10, sort_by_norm=False, rows=(0, 300))
    np.testing.assert_array_equal(got[0][:300], widx[:300])
    np.testing.assert_array_equal(got[1][:300], wcos[:300])


@pytest.mark.parametrize("matmul", BACKENDS)
@pytest.mark.parametrize("k", [10, 30])
def test_rare_workload_sampled_rows(k, matmul, oracle):
    """396,460 x 15 with duplicates: every row of a 1,500-row sample equals the oracle's blocked
    form (sort_by_norm off, which is output-neutral)."""
    x = synth.rare_396460x15()
    (idx, cos, norms), st = ext._knn_exact(x, k, matmul=matmul, weights=True)
    rng = np.random.default_rng(k)
    r0 = int(rng.integers(0, len(x) - 1500))
    (widx, wcos, wnorms), _ = oracle.knn_blocked(x, k, sort_by_norm=False, rows=(r0, r0 + 1500))
    sl = slice(r0, r0 + 1500)
    assert_same((idx[sl], cos[sl], norms), (widx[sl], wcos[sl], wnorms))
    # size-independent properties on every row
    assert (idx != np.arange(len(x))[:, None]).all()
    assert (np.diff(cos, axis=1) <= 0).all()
    w = st["weights"]
    ww = oracle.reweight(idx, cos, norms, "tanimoto")
    np.testing.assert_allclose(w, ww, rtol=W_RTOL, atol=0)


@pytest.mark.parametrize("matmul", BACKENDS)
def test_cytometry_workload_sampled_rows(matmul, oracle):
    x = synth.cytometry_86864x39()
    (idx, cos, norms), st = ext._knn_exact(x, 10, matmul=matmul)
    (widx, wcos, wnorms), _ = oracle.knn_blocked(x, 10, sort_by_norm=False, rows=(0, 2000))
    assert_same((idx[:2000], cos[:2000], norms), (widx[:2000], wcos[:2000], wnorms))


@pytest.mark.parametrize("matmul,k", [("mfma-f32", 10), ("mfma-f16x3", 30)])
def test_balanced_launch_equals_direct_launch(matmul, k):
    """With tuned splits a large input takes the balanced launch for these forms (equal shares of
    (query block, tile) units per resident workgroup, so some blocks are scanned in two pieces by
    two workgroups); every row equals the one-split direct launch."""
    x = synth.cytometry_86864x39()
    a, sa = ext._knn_exact(x, k, matmul=matmul)
    b, sb = ext._knn_exact(x, k, matmul=matmul, col_splits=1)
    assert sa["col_splits"] == 2 and sb["col_splits"] == 1
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


def test_truncate_prefix_property():
    # tests/test_knn.py:97-108
    import costal_amd as cp

    x = synth.clustered(800, 10, 6)
    g = cp.knn(x, 15)
    assert (g.idx != np.arange(800)[:, None]).all()
    t = cp.truncate(g, 6)
    g6 = cp.knn(x, 6)
    np.testing.assert_array_equal(t.idx, g6.idx)
    np.testing.assert_array_equal(t.cos, g6.cos)


def test_reweight_matches_oracle_bitwise(oracle):
    import costal_amd as cp

    g = cp.knn(synth.clustered(600, 12, 8), 10)
    for weight in ("tanimoto", "cosine", "jaccard"):
        got = cp.reweight(g, weight)
        want = oracle.reweight(g.idx, g.cos, g.norms, weight)
        np.testing.assert_array_equal(got, want)


def test_fuzzy_union_matches_oracle(oracle):
    import costal_amd as cp

    x = synth.clustered(1000, 16, 10)
    g = cp.knn(x, 10)
    w = cp.reweight(g, "tanimoto")
    got = cp.build_graph(x, 10)
    want = oracle.fuzzy_union(g.idx, w)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)


def test_tanimoto_selection_matches_bruteforce(oracle):
    # tests/test_knn.py:118-125
    x = synth.clustered(1500, 12, 14)
    x *= np.random.default_rng(114).uniform(0.5, 3.0, size=(1500, 1)).astype(np.float32)
    x = np.ascontiguousarray(x)
    got, st = ext._knn_exact(x, 10, select="tanimoto")
    assert st["select"] == "tanimoto"
    assert_same(got, oracle.knn_bruteforce(x, 10, select="tanimoto"))


def test_non_finite_row_is_reported():
    x = synth.clustered(50, 4, 7)
    x[3, 1] = np.nan
    x[7, 0] = np.inf
    with pytest.raises(ValueError, match="row 3 contains NaN or inf"):
        ext._knn_exact(x, 5)


@pytest.mark.parametrize("case", ["nan", "inf", "-inf", "row of inf", "two rows"])
def test_non_finite_row_on_the_pruned_path(case, oracle):
    """At a size that takes the pruned scan: the first row with a NaN or inf is the call's error,
    as a value, and the next call, on clean rows, is exact and pruned."""
    from costal_amd import _lib

    x = synth.rare_396460x15(n=30000)
    y = x.copy()
    first = 20011
    if case == "nan":
        y[first, 4] = np.nan
    elif case == "inf":
        y[first, 0] = np.inf
    elif case == "-inf":
        y[first, 14] = -np.inf
    elif case == "row of inf":
        y[first] = np.inf
    else:
        first = 517
        y[first, 3] = np.inf
        y[29999, 2] = np.nan
    with pytest.raises(_lib.CostalError, match=f"row {first} contains NaN or inf") as e:
        ext._knn_exact(y, 10)
    assert e.value.code == _lib.E_NON_FINITE and e.value.row == first
    got, st = ext._knn_exact(x, 10)
    assert st["blocks_skipped"] > 0
    (widx, wcos, wnorms), _ = oracle.knn_blocked(x, 10, sort_by_norm=False, rows=(19500, 20500))
    np.testing.assert_array_equal(got[0][19500:20500], widx[19500:20500])
    np.testing.assert_array_equal(got[1][19500:20500], wcos[19500:20500])
    np.testing.assert_array_equal(got[2], wnorms)
