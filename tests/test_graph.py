"""The graph after the kNN: fuzzy union's edge list into the compressed rows the CPU Leiden consumes
(ike12th/CosTaL crates/costal-core/src/leiden.rs:52-100, Csr::from_edges).  Host code of the
library, so these run without a GPU; the oracle restates the same function, and both are pinned by
the reference's own test vectors (leiden.rs:1525-1553) and by scipy's sparse CSR of S + S^T."""

import numpy as np
import pytest

from costal_amd import _costal, synth


def both(n, edges):
    rows = [e[0] for e in edges]
    cols = [e[1] for e in edges]
    w = [e[2] for e in edges]
    return _costal.Csr.from_edges(n, rows, cols, w), rows, cols, w


def test_csr_holds_both_directions(oracle):
    # leiden.rs:1525-1533
    g, rows, cols, w = both(4, [(0, 1, 0.5), (0, 3, 0.25), (1, 2, 1.0)])
    assert g.indptr.tolist() == [0, 2, 4, 5, 6]
    assert g.nbr.tolist() == [1, 3, 0, 2, 1, 0]
    assert g.w.tolist() == [0.5, 0.25, 0.5, 1.0, 1.0, 0.25]
    assert g.strengths().tolist() == [0.75, 1.5, 1.0, 0.25]
    assert g.n_edges() == 3
    o = oracle.csr_from_edges(4, rows, cols, w)
    assert [a.tolist() for a in o] == [g.indptr.tolist(), g.nbr.tolist(), g.w.tolist()]


def test_csr_drops_zero_weights_and_rejects_bad_edges(oracle):
    # leiden.rs:1535-1553
    g, rows, cols, w = both(3, [(0, 1, 0.0), (1, 2, 0.5)])
    assert g.indptr.tolist() == [0, 0, 1, 2]
    assert oracle.csr_from_edges(3, rows, cols, w)[0].tolist() == [0, 0, 1, 2]
    cases = [((2, [0], [0], [1.0]), "self-loop"), ((2, [0], [2], [1.0]), "out of range"),
             ((2, [0], [1], [-1.0]), "weight"), ((2, [0], [1], [np.nan]), "weight")]
    for args, text in cases:
        with pytest.raises(ValueError, match=text):
            _costal.Csr.from_edges(*args)
        with pytest.raises(ValueError, match=text):
            oracle.csr_from_edges(*args)
    with pytest.raises(ValueError, match="length"):
        _costal.Csr.from_edges(2, [0, 1], [1], [1.0])


def test_error_texts_match_the_reference():
    with pytest.raises(ValueError, match=r"^edge 1 is a self-loop on node 2$"):
        _costal.Csr.from_edges(3, [0, 2], [1, 2], [1.0, 1.0])
    with pytest.raises(ValueError, match=r"^edge 0 \(0, 5\) is out of range \(n=3\)$"):
        _costal.Csr.from_edges(3, [0], [5], [1.0])
    with pytest.raises(ValueError, match=r"^edge 0 has weight -0.25; weights must be finite and non-negative$"):
        _costal.Csr.from_edges(3, [0], [1], [-0.25])
    with pytest.raises(ValueError, match=r"^edge 0 has weight inf; "):
        _costal.Csr.from_edges(3, [0], [1], [np.inf])
    with pytest.raises(ValueError, match=r"^edge arrays differ in length: 2 rows, 1 cols, 1 weights$"):
        _costal.Csr.from_edges(3, [0, 1], [1], [1.0])


def fuzzy_graph(oracle, n=1500, k=12, seed=3):
    x = synth.clustered(n, 10, seed)
    idx, cos, norms = oracle.knn_bruteforce(x, k)
    w = oracle.reweight(idx, cos, norms, "tanimoto")
    return oracle.fuzzy_union(idx, w)


def test_csr_of_the_fuzzy_union_equals_scipy(oracle):
    """On the sorted upper-triangular output of the fuzzy union, every row lists its neighbors in
    ascending order, so the CSR is scipy's CSR of S + S^T (sorted indices), bit for bit."""
    import scipy.sparse as sp

    rows, cols, w = fuzzy_graph(oracle)
    n = 1500
    g = _costal.Csr.from_edges(n, rows, cols, w)
    s = sp.coo_matrix((w, (rows, cols)), shape=(n, n))
    ref = (s + s.T).tocsr()
    ref.sort_indices()
    np.testing.assert_array_equal(g.indptr, ref.indptr)
    np.testing.assert_array_equal(g.nbr, ref.indices)
    np.testing.assert_array_equal(g.w, ref.data)
    # strengths: each row's weights summed in row order (numpy's cumsum is sequential)
    seq = [np.cumsum(ref.data[a:b])[-1] if b > a else 0.0 for a, b in zip(ref.indptr[:-1], ref.indptr[1:])]
    np.testing.assert_array_equal(g.strengths(), seq)
    o = oracle.csr_from_edges(n, rows, cols, w)
    for a, b in zip(o, (g.indptr, g.nbr, g.w)):
        np.testing.assert_array_equal(a, b)


def test_csr_keeps_edge_order_and_duplicates(oracle):
    """Unsorted input: rows list neighbors in edge order; a pair listed twice counts twice."""
    rows, cols, w = fuzzy_graph(oracle, n=600, k=8, seed=5)
    rng = np.random.default_rng(1)
    perm = rng.permutation(len(rows))
    flip = rng.random(len(rows)) < 0.5
    r = np.where(flip, cols, rows)[perm]
    c = np.where(flip, rows, cols)[perm]
    ww = w[perm]
    r = np.concatenate([r, r[:5]])
    c = np.concatenate([c, c[:5]])
    ww = np.concatenate([ww, ww[:5]])
    ww[7] = 0.0
    g = _costal.Csr.from_edges(600, r, c, ww)
    o = oracle.csr_from_edges(600, r, c, ww)
    for a, b in zip(o, (g.indptr, g.nbr, g.w)):
        np.testing.assert_array_equal(a, b)
    assert g.n_edges() == len(r) - 1
    # row v in edge order: the other endpoints of the kept edges that touch v, as listed
    v = int(r[0])
    want = [int(c[e]) if r[e] == v else int(r[e]) for e in range(len(r)) if (r[e] == v or c[e] == v) and ww[e] > 0]
    assert g.nbr[g.indptr[v]:g.indptr[v + 1]].tolist() == want


def test_empty_graph():
    g = _costal.Csr.from_edges(0, [], [], [])
    assert g.indptr.tolist() == [0] and g.n_edges() == 0
    g = _costal.Csr.from_edges(5, [], [], [])
    assert g.indptr.tolist() == [0] * 6
