"""The CPU oracle, pinned by what the reference's own tests and the mathematics fix.

Every expected value below comes from the reference's tests (ike12th/CosTaL knn.rs:926-1160,
weight.rs:95-122, graph.rs:70-96, tests/test_knn.py), from an independent numpy computation, or
from an identity between the oracle's two independent code paths (brute force and blocked form).
"""

import glob
import itertools
import os

import numpy as np
import pytest

from costal_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))


def same(a, b):
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


# ------------------------------------------------------------------ closed forms from the reference


def test_heap_keeps_best_with_index_tiebreak(oracle):
    # knn.rs:1149-1159
    kept = oracle.heap_keep([(0.5, 7), (0.9, 3), (0.5, 2), (0.7, 9), (0.5, 1), (0.95, 0)], 3)
    assert sorted(kept) == [0, 3, 9]


def test_tanimoto_closed_forms(oracle):
    # weight.rs:99-109
    assert oracle.tani_from_cos(1.0, 2.0, 1.0) == 2.0 / 3.0
    c = 0.6
    assert abs(oracle.tani_from_cos(3.0, 3.0, c) - c / (2.0 - c)) < 1e-15
    assert oracle.tani_from_cos(0.0, 2.0, 0.9) == 0.0
    assert abs(oracle.tani_from_cos(1.0, 1.0, -1.0) + 1.0 / 3.0) < 1e-15


def test_jaccard_vector(oracle):
    # weight.rs:111-121
    idx = np.array([1, 2, 0, 2, 0, 1, 0, 1], np.uint32).reshape(4, 2)
    w = oracle.reweight(idx, np.zeros((4, 2)), np.ones(4), "jaccard").ravel()
    assert w[0] == 1.0 / 3.0 and w[6] == 1.0 / 3.0 and w[4] == 1.0 / 3.0


def test_fuzzy_union_vectors(oracle):
    # graph.rs:74-95
    idx = np.array([1, 2, 0, 2, 1, 0], np.uint32).reshape(3, 2)
    w = np.array([0.5, 0.3, 0.4, 0.2, 0.0, 0.6]).reshape(3, 2)
    rows, cols, ww = oracle.fuzzy_union(idx, w)
    assert rows.tolist() == [0, 0, 1] and cols.tolist() == [1, 2, 2]
    assert ww.tolist() == [(0.5 + 0.4) - 0.5 * 0.4, (0.3 + 0.6) - 0.3 * 0.6, 0.2]
    rows, cols, ww = oracle.fuzzy_union(np.array([[1], [2], [0]], np.uint32), np.array([[0.0], [0.7], [0.0]]))
    assert (rows.tolist(), cols.tolist(), ww.tolist()) == ([1], [2], [0.7])


def test_duplicates_break_ties_by_original_index(oracle):
    # knn.rs:1001-1017 (LCG data) and tests/test_knn.py:78-84 (numpy data): row 0 -> 10..100
    g = synth.Lcg(3)
    protos = (g.take(50) - np.float32(0.3)).reshape(10, 5)
    x = np.ascontiguousarray(np.tile(protos, (30, 1)))
    bf = oracle.knn_bruteforce(x, 10)
    bl, _ = oracle.knn_blocked(x, 10)
    same(bf, bl)
    assert bf[0][0].tolist() == list(range(10, 101, 10))
    rng = np.random.default_rng(4)
    x = np.ascontiguousarray(np.tile(rng.normal(size=(10, 5)).astype(np.float32), (30, 1)))
    assert oracle.knn_blocked(x, 10)[0][0][0].tolist() == list(range(10, 101, 10))


def test_zero_rows_and_integer_ties(oracle):
    # knn.rs:1019-1032
    g = synth.Lcg(4)
    x = np.floor(g.take(400 * 6) * np.float32(3.0)).reshape(400, 6)
    x[[0, 57, 399]] = 0.0
    x = np.ascontiguousarray(x)
    bf = oracle.knn_bruteforce(x, 12)
    bl, _ = oracle.knn_blocked(x, 12)
    same(bf, bl)
    assert (bf[1][0] == 0.0).all()


def test_k_close_to_n(oracle):
    # knn.rs:1034-1044
    x = synth.lcg_clustered(25, 4, 5)
    for k in (1, 23, 24):
        same(oracle.knn_blocked(x, k)[0], oracle.knn_bruteforce(x, k))


def test_rejects_bad_input(oracle):
    # knn.rs:1046-1067
    x = np.ones((4, 3), np.float32)
    for k in (4, 0):
        with pytest.raises(oracle.OracleError) as e:
            oracle.knn_blocked(x, k)
        assert e.value.kind == "bad_k" and f"invalid k={k}: need 1 <= k <= n-1 (n=4)" in str(e.value)
    y = x.copy()
    y.ravel()[7] = np.nan
    with pytest.raises(oracle.OracleError, match="row 2 contains NaN or inf"):
        oracle.knn_bruteforce(y, 1)


# ------------------------------------------------------------------ independent numerics


def test_matches_independent_numpy_oracle(oracle):
    # tests/test_knn.py:46-61: numpy f64 matmul (a different summation order) at atol 1e-12, and
    # identical neighbor sets on clear-gap rows
    x = synth.clustered(1500, 20, 2)
    idx, cos, _ = oracle.knn_bruteforce(x, 12)
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


def test_sequential_sums_bitwise_against_numpy_cumsum(oracle):
    """dot64 sums in index order (knn.rs:289-296); np.cumsum is a strictly sequential f64 sum, so
    norms and cosines must agree bit for bit; a reordered or contracted sum fails this."""
    rng = np.random.default_rng(7)
    x = (rng.normal(size=(60, 37)) * rng.uniform(0.01, 100, size=(60, 1))).astype(np.float32)
    idx, cos, norms = oracle.knn_bruteforce(x, 5)
    x64 = x.astype(np.float64)
    want_norms = np.sqrt(np.cumsum(x64 * x64, axis=1)[:, -1])
    np.testing.assert_array_equal(norms, want_norms)
    for i in range(60):
        for t in range(5):
            j = idx[i, t]
            dot = np.cumsum(x64[i] * x64[j])[-1]
            assert cos[i, t] == dot / (want_norms[i] * want_norms[j])
    # the ranking is the brute-force definition: (score desc, index asc)
    for i in range(60):
        dots = np.array([np.cumsum(x64[i] * x64[j])[-1] / (want_norms[i] * want_norms[j])
                         for j in range(60)])
        dots[i] = -np.inf
        order = sorted(range(60), key=lambda j: (-dots[j], j))[:5]
        assert idx[i].tolist() == order


def test_tanimoto_selection_against_numpy(oracle):
    # tests/test_knn.py:128-138
    x = synth.clustered(1000, 10, 15)
    x *= np.random.default_rng(115).uniform(0.5, 3.0, size=(1000, 1)).astype(np.float32)
    x = np.ascontiguousarray(x)
    idx, _, _ = oracle.knn_bruteforce(x, 8, select="tanimoto")
    x64 = x.astype(np.float64)
    dot = x64 @ x64.T
    sq = np.einsum("ij,ij->i", x64, x64)
    t = dot / (sq[:, None] + sq[None, :] - dot)
    np.fill_diagonal(t, -np.inf)
    want = -np.sort(-t, axis=1)[:, :8]
    got = np.take_along_axis(t, idx.astype(np.int64), axis=1)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def original_tani_from_cos(a, b, cos_val):
    d = a * b * cos_val
    if d == 0:
        return 0
    return d / (a * a + b * b - d)


def test_tanimoto_weights_bitwise_against_original_formula(oracle):
    # tests/test_knn.py:177-183, 195-205: Python floats are IEEE f64 without contraction
    idx, cos, norms = oracle.knn_bruteforce(synth.clustered(600, 12, 8), 10)
    w = oracle.reweight(idx, cos, norms, "tanimoto")
    want = np.array([[original_tani_from_cos(norms[i], norms[j], c) for j, c in zip(idx[i], cos[i])]
                     for i in range(600)], dtype=np.float64)
    np.testing.assert_array_equal(w, want)
    np.testing.assert_array_equal(oracle.reweight(idx, cos, norms, "cosine"), cos)


def test_jaccard_bitwise_against_original(oracle):
    # tests/test_knn.py:186-211
    idx, cos, norms = oracle.knn_bruteforce(synth.clustered(400, 8, 9), 10)
    want = []
    for i in range(len(idx)):
        s = np.fromiter((len(set(idx[i]).intersection(set(idx[j]))) for j in idx[i]), dtype=float)
        want.append(s / (2 * idx.shape[1] - s))
    np.testing.assert_array_equal(oracle.reweight(idx, cos, norms, "jaccard"), np.array(want))


def test_fuzzy_union_bitwise_against_scipy_pipeline(oracle):
    # tests/test_knn.py:217-237: the original scipy symmetrization
    sp = pytest.importorskip("scipy.sparse")
    idx, cos, norms = oracle.knn_bruteforce(synth.clustered(1000, 16, 10), 10)
    w = oracle.reweight(idx, cos, norms, "tanimoto")
    n, k = idx.shape
    csr = sp.csr_matrix((w.ravel(), (np.repeat(np.arange(n), k), idx.ravel())), shape=(n, n))
    csr = csr + csr.transpose() - csr.multiply(csr.transpose())
    csr.setdiag(0)
    csr = sp.triu(csr)
    csr.eliminate_zeros()
    src, dst = csr.nonzero()
    got = oracle.fuzzy_union(idx, w)
    np.testing.assert_array_equal(got[0], src)
    np.testing.assert_array_equal(got[1], dst)
    np.testing.assert_array_equal(got[2], csr.data)


# ------------------------------------------------------------------ the two code paths agree


@pytest.mark.parametrize("seed", [0, 1])
def test_blocked_matches_bruteforce_no_fallback(seed, oracle):
    # tests/test_knn.py:38-43
    x = synth.clustered(3000, 32, seed)
    got, st = oracle.knn_blocked(x, 10)
    same(got, oracle.knn_bruteforce(x, 10))
    assert st["fallback_rows"] == 0


def test_blocked_independent_of_blocks_layout_threads(oracle):
    # knn.rs:969-999
    x = synth.lcg_clustered(900, 17, 2)
    want = oracle.knn_bruteforce(x, 7)
    for (rb, cb), sort, threads in itertools.product([(1, 1), (8, 64), (33, 129), (512, 4096)],
                                                      [False, True], [1, 4]):
        got, _ = oracle.knn_blocked(x, 7, row_block=rb, col_block=cb, sort_by_norm=sort,
                                    n_threads=threads)
        same(got, want)


def test_tie_heavy_blocked_triggers_fallback(oracle):
    # tests/test_knn.py:87-94
    rng = np.random.default_rng(5)
    x = np.floor(rng.uniform(0, 3, size=(500, 6))).astype(np.float32)
    x[[0, 99, 499]] = 0.0
    got, st = oracle.knn_blocked(x, 12)
    same(got, oracle.knn_bruteforce(x, 12))
    assert st["fallback_rows"] > 0


def test_row_range_equals_full_call(oracle):
    x = synth.cytometry_86864x39(n=4000)
    full, _ = oracle.knn_blocked(x, 10, sort_by_norm=False)
    part, _ = oracle.knn_blocked(x, 10, sort_by_norm=False, rows=(1000, 1500))
    np.testing.assert_array_equal(part[0][1000:1500], full[0][1000:1500])
    np.testing.assert_array_equal(part[1][1000:1500], full[1][1000:1500])


def test_block_sizes_policy(oracle):
    # knn.rs:551-569 with a 1 MiB budget: side = 512 -> cb = 512; rb = min(budget/cb, n/(8T)) in [8, 512]
    import ctypes as C

    rb, cb = C.c_uint64(), C.c_uint64()
    oracle.lib().oracle_block_sizes(100_000, 0, 0, 1 << 20, 8, C.byref(rb), C.byref(cb))
    assert cb.value == 512 and rb.value == 512
    oracle.lib().oracle_block_sizes(1000, 0, 0, 1 << 20, 8, C.byref(rb), C.byref(cb))
    assert cb.value == 512 and rb.value == 16


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "golden", "*.npz"))),
                         ids=lambda p: os.path.basename(p)[:-4])
def test_golden_fixtures_reproduce(path, oracle):
    f = np.load(path)
    x, k = f["x"], int(f["k"])
    got, _ = oracle.knn_blocked(x, k)
    same(got, (f["idx"], f["cos"], f["norms"]))
    np.testing.assert_array_equal(oracle.reweight(f["idx"], f["cos"], f["norms"], "tanimoto"), f["w"])
