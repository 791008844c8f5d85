"""The C-ABI library and its Python mirror, without a GPU: exports, argument checks, error text.

Argument checks that the reference performs before any work (ike12th/CosTaL
crates/costal-py/src/lib.rs:104-214, crates/costal-core/src/knn.rs:325-342, python/costal/_api.py)
must behave the same here; they run on the host before the device is touched.
"""

import os
import re
import sys

import numpy as np
import pytest

from costal_amd import _api, _costal, _lib, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "costal_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|uint64_t|const char\*)\s+(costal_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 9
    L = _lib.load()
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(_lib.EXPORTS)


def test_version_string():
    assert "gfx950" in _lib.version()


def _has_gpu():
    try:
        _lib.device_info()
        return True
    except _lib.CostalError:
        return False


def test_no_device_is_a_clean_error():
    if _has_gpu():
        pytest.skip("a GPU is present")
    x = synth.clustered(50, 4, 7)
    with pytest.raises(_lib.CostalError) as e:
        _costal._knn_exact(x, 5)
    assert e.value.code == _lib.E_DEVICE


@pytest.mark.parametrize("k", [0, 50, 51])
def test_bad_k_message_matches_reference(k):
    x = synth.clustered(50, 4, 7)
    with pytest.raises(ValueError, match=re.escape(f"invalid k={k}: need 1 <= k <= n-1 (n=50)")):
        _costal._knn_exact(x, k)


def test_input_type_checks():
    x = synth.clustered(50, 4, 7)
    with pytest.raises(TypeError):
        _costal._knn_exact(x.astype(np.float64), 5)
    with pytest.raises(ValueError, match="C-contiguous"):
        _costal._knn_exact(np.asfortranarray(x), 5)
    with pytest.raises(ValueError, match='unknown select "jaccard"'):
        _costal._knn_exact(x, 5, select="jaccard")
    with pytest.raises(ValueError, match='unknown matmul backend "blas"'):
        _costal._knn_exact(x, 5, matmul="blas")
    with pytest.raises(ValueError, match="accelerate is not available"):
        _costal._knn_exact(x, 5, matmul="accelerate")


def test_reweight_argument_checks():
    idx = np.array([[1], [5]], np.int32)
    with pytest.raises(ValueError, match="neighbor index 5 out of range"):
        _costal._reweight(idx, np.zeros((2, 1)), np.ones(2), "tanimoto")
    with pytest.raises(ValueError, match='unknown edge weight "l2"'):
        _costal._reweight(np.array([[1], [0]], np.int32), np.zeros((2, 1)), np.ones(2), "l2")
    with pytest.raises(ValueError, match="shapes"):
        _costal._reweight(np.array([[1], [0]], np.int32), np.zeros((2, 2)), np.ones(2), "cosine")


def test_jaccard_and_fuzzy_union_run_on_host(oracle):
    # weight.rs:111-121 and graph.rs:74-95 through the library's C ABI
    idx = np.array([1, 2, 0, 2, 0, 1, 0, 1], np.int32).reshape(4, 2)
    w = _costal._reweight(idx, np.zeros((4, 2)), np.ones(4), "jaccard").ravel()
    assert w[0] == 1 / 3 and w[4] == 1 / 3 and w[6] == 1 / 3
    idx = np.array([1, 2, 0, 2, 1, 0], np.int32).reshape(3, 2)
    ww = np.array([0.5, 0.3, 0.4, 0.2, 0.0, 0.6]).reshape(3, 2)
    rows, cols, wo = _costal._fuzzy_union(idx, ww)
    assert rows.tolist() == [0, 0, 1] and cols.tolist() == [1, 2, 2]
    assert wo.tolist() == [(0.5 + 0.4) - 0.5 * 0.4, (0.3 + 0.6) - 0.3 * 0.6, 0.2]
    # a larger random graph against the oracle's union
    rng = np.random.default_rng(0)
    n, k = 500, 7
    idx = np.stack([rng.choice(np.delete(np.arange(n), i), k, replace=False) for i in range(n)]).astype(np.int32)
    w = rng.uniform(-0.3, 1, size=(n, k))
    w[rng.random((n, k)) < 0.05] = 0.0
    got = _costal._fuzzy_union(idx, w)
    want = oracle.fuzzy_union(idx, w)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)


def test_api_choices_match_reference():
    # python/costal/_api.py:66-100
    x = synth.clustered(50, 4, 7)
    with pytest.raises(ValueError):
        _api.knn(x, 5, select="jaccard")
    with pytest.raises(ValueError):
        _api.knn(x, 5, compat="nope")
    with pytest.raises(NotImplementedError):
        _api.knn(x, 5, backend="hnsw")
    sp = pytest.importorskip("scipy.sparse")
    with pytest.raises(TypeError):
        _api.knn(sp.csr_matrix(x), 5)
    m = _api._as_matrix(np.asfortranarray(x.astype(np.float64)))
    assert m.dtype == np.float32 and m.flags.c_contiguous
    np.testing.assert_array_equal(m, x)
    with pytest.raises(ValueError, match="k must be within"):
        _api.truncate(_api.KnnGraph(np.zeros((3, 2), np.int32), np.zeros((3, 2)), np.ones(3)), 3)


def test_diagnostics_flags():
    # python/costal/_api.py:112-144 and tests/test_knn.py:243-261
    x = synth.clustered(300, 8, 11)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    norms = np.linalg.norm(x.astype(np.float64), axis=1)
    assert _api._diagnose(x, norms)["row_norm_cv"] < 1e-4
    y = synth.clustered(300, 8, 12)
    y -= y.mean(axis=0)
    assert _api._diagnose(y, np.linalg.norm(y.astype(np.float64), axis=1))["columns_centered"]
    z = synth.clustered(300, 8, 13)
    dz = _api._diagnose(z, np.linalg.norm(z.astype(np.float64), axis=1))
    assert not dz["columns_centered"] and dz["row_norm_cv"] > 1e-3


def test_synthetic_workloads_are_deterministic_and_shaped():
    a = synth.cytometry_86864x39(n=5000)
    b = synth.cytometry_86864x39(n=5000)
    np.testing.assert_array_equal(a, b)
    assert a.dtype == np.float32 and a.shape == (5000, 39) and (a < 0).any()
    c = synth.counts_3605x2000(n=400)
    assert c.shape == (400, 2000) and abs((c == 0).mean() - 0.70) < 0.03
    r = synth.rare_396460x15(n=20000)
    assert len(np.unique(r, axis=0)) == 20000 - 100
    s = synth.batched_64(n_samples=6)
    assert all(2000 <= len(v) <= 60000 and 32 <= v.shape[1] <= 48 for v in s)


def test_lcg_matches_rust_generator():
    # knn.rs:931-940: first draws of Lcg(1)
    g = synth.Lcg(1)
    s = 1
    for _ in range(5):
        s = (s * 6364136223846793005 + 1442695040888963407) % 2**64
        assert g.next_f32() == np.float32(s >> 40) / np.float32(1 << 24)


def test_one_hip_runtime_whichever_loads_first():
    """The library loaded before PyTorch must not bring a second HIP runtime into the process
    (PyTorch would then find no GPU): a fresh process loads the library, then imports torch."""
    import subprocess

    code = ("from costal_amd import _lib; _lib.load(); import torch; "
            "m = open('/proc/self/maps').read().splitlines(); "
            "print(sorted({l.split()[-1] for l in m if 'libamdhip64' in l}))")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    import ast

    paths = ast.literal_eval(out.stdout.strip().splitlines()[-1])
    assert len(paths) == 1, paths


def test_tuner_choice_is_stable_under_timing_noise():
    """Configurations within 3% of the fastest are ties, settled by the shortest lists, then split
    FP16, then tuned splits, so one noisy timing run does not flip the plan."""
    from costal_amd import tuner

    res = [
        {"precision": "fp16x3", "margin": 15, "col_splits": 0, "ms": 1.00},
        {"precision": "fp16x3", "margin": 7, "col_splits": 4, "ms": 1.02},
        {"precision": "fp16x3", "margin": 7, "col_splits": 0, "ms": 1.025},
        {"precision": "fp32", "margin": 7, "col_splits": 0, "ms": 1.01},
        {"precision": "fp16x3", "margin": 0, "col_splits": 0, "ms": 1.20},
    ]
    assert tuner.choose(res) == res[2]
    # a clear winner is kept whatever its knobs
    res[0]["ms"] = 0.90
    assert tuner.choose(res) == res[0]


def test_tuner_grid_covers_the_order_knobs(tmp_path):
    """The second stage times the best first-stage configuration under other k-means center counts
    and unpruned; ties keep the pruned scan with the default center count; cached entries from
    before the order knobs read as the defaults."""
    from costal_amd import tuner

    first = list(tuner.candidates(10))
    assert all(c["prune"] and c["prune_clusters"] == 0 for c in first)
    assert {c["precision"] for c in first} == {"fp16x3", "fp32"}
    best = {"precision": "fp16x3", "margin": 5, "col_splits": 0, "prune_clusters": 0, "prune_tile": 0, "prune": True}
    second = list(tuner.order_candidates(best))
    assert {(c["prune"], c["prune_clusters"], c["prune_tile"]) for c in second} == {
        (True, 32, 0), (True, 128, 0), (True, 0, 512), (False, 0, 0)}
    assert all(c["margin"] == 5 and c["precision"] == "fp16x3" for c in second)
    res = [{**best, "ms": 1.0}] + [{**c, "ms": 0.99} for c in second]
    assert tuner.choose(res) == res[0]
    res[3]["ms"] = 0.5
    assert tuner.choose(res) == res[3]
    tn = tuner.Tuner(str(tmp_path / "tune.json"))
    tn.table[tuner.bucket(5000, 40, 10)] = {"best": {"precision": "fp32", "margin": 10, "col_splits": 1}}
    got = tn.get(5000, 40, 10)
    assert got["prune"] and got["prune_clusters"] == 0 and got["prune_tile"] == 0
    assert tuner.knn_kwargs(got) == {"matmul": "mfma-f32", "margin": 10, "col_splits": 1, "prune_clusters": 0,
                                     "prune_tile": 0, "prune": True}


def test_batch_plan_drops_splits_when_samples_fill_the_gpu(tmp_path):
    from costal_amd import tuner

    tn = tuner.Tuner(str(tmp_path / "tune.json"))
    for n, d in [(300, 32), (200000, 40), (100000, 48)]:
        tn.table[tuner.bucket(n, d, 15)] = {"best": {"precision": "fp16x3", "margin": 7, "col_splits": 4}}
    prec, margins, splits = tn.plan_batch([(300, 32)], 15)
    assert (prec, margins, splits) == ("fp16x3", [7], [4])
    _, _, splits = tn.plan_batch([(300, 32), (200000, 40), (100000, 48)], 15)  # 2,348 query blocks
    assert splits == [1, 1, 1]


def test_ctypes_mirrors_match_the_header_layout(tmp_path):
    """The ctypes structures of the Python mirror have the C structures' sizes and field offsets
    (a C program built against include/costal_hip.h prints them)."""
    import ctypes as C
    import subprocess

    structs = {"costal_knn_params": _lib.KnnParams, "costal_knn_stats": _lib.KnnStats,
               "costal_sample": _lib.Sample}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "costal_hip.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(line.split()[:2]): int(line.split()[2]) for line in out if line}
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def _plan(n, d, k, **kw):
    import ctypes as C

    p = _lib.KnnParams()
    p.k = k
    p.margin = kw.get("margin", -1)
    p.precision = _lib.PRECISION[kw.get("precision", "fp16x3")]
    p.select = _lib.SELECT[kw.get("select", "cosine")]
    p.col_splits = kw.get("col_splits", 0)
    st = _lib.KnnStats()
    _lib.check(_lib.load().costal_knn_plan(n, d, C.byref(p), C.byref(st)))
    return st


@pytest.mark.parametrize("k,splits", [(10, 128), (10, 512), (30, 512), (64, 100)])
def test_plan_caps_caller_splits_at_the_rescoring_capacity(k, splits):
    """A caller's col_splits is speed only: the plan never hands the rescoring more than 2048 list
    entries per row (2 lists per split, K' entries each), whatever was asked for."""
    st = _plan(200_000, 15, k, col_splits=splits)
    assert st.path == 0
    assert 1 <= st.col_splits <= splits
    assert st.col_splits * 2 * st.candidates <= 2048


def test_plan_paths_and_bounds():
    st = _plan(396_460, 15, 10)
    assert (st.path, st.candidates, st.row_block) == (0, 10, 128)
    eps16x3 = st.coarse_eps
    st32 = _plan(396_460, 15, 10, precision="fp32")
    assert st32.coarse_eps == (15 + 4) * 2.0 ** -23  # the reference's bound, knn.rs:185-187
    assert 0 < eps16x3 < 2 * st32.coarse_eps
    assert _plan(3605, 2000, 10, precision="fp32").path == 2  # wide rows
    tani = _plan(5000, 15, 10, select="tanimoto")  # the fused scan with Tanimoto lists
    assert tani.path == 0
    # the list error of a Tanimoto score: twice the split-FP16 cosine bound and then some
    assert 2 * eps16x3 < tani.coarse_eps < 2 * eps16x3 * 1.01 + 2.0 ** -19
    assert _plan(5000, 15, 100, select="tanimoto").path == 1  # lists beyond 64: exact row scan
    assert _plan(5000, 80, 10, select="tanimoto").path == 1  # rows beyond 64 features


def test_plan_rejects_what_the_reference_rejects():
    with pytest.raises(ValueError, match=re.escape("invalid k=0: need 1 <= k <= n-1 (n=50)")):
        _plan(50, 4, 0)


def test_diagnose_merges_blocks_exactly():
    """The block-merged column moments equal the direct ones: a matrix centred as a whole but not
    within any one block is still reported centred, and the fractions count every entry."""
    n, d = 200_000, 12  # about 2.3 blocks of rows
    rng = np.random.default_rng(21)
    x = rng.normal(0, 1, size=(n, d)).astype(np.float32)
    x[: n // 2] += 3.0
    x[n // 2:] -= 3.0
    x -= x.astype(np.float64).mean(axis=0).astype(np.float32)
    norms = np.linalg.norm(x.astype(np.float64), axis=1)
    got = _api._diagnose(x, norms)
    assert got["columns_centered"]
    assert got["negative_fraction"] == np.count_nonzero(x < 0) / x.size
    assert got["nonzero_fraction"] == np.count_nonzero(x) / x.size
    assert np.isclose(got["row_norm_cv"], norms.std() / norms.mean(), rtol=1e-12)
    y = x + np.float32(0.2)  # |mean| = 0.2 > 0.05 std (std ~ 3.2)
    assert not _api._diagnose(y, np.linalg.norm(y.astype(np.float64), axis=1))["columns_centered"]
