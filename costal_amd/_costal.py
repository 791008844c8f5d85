"""The low-level entry points of CosTaL's ``costal._costal`` module, served by the gfx950 library.

Same names, signatures, return values and error behaviour as the reference's PyO3 bindings
(ike12th/CosTaL crates/costal-py/src/lib.rs):

- :func:`_knn_exact` -- lib.rs:113-167
- :func:`_knn_bruteforce` -- lib.rs:171-191
- :func:`_reweight` -- lib.rs:194-214
- :func:`_fuzzy_union` -- lib.rs:218-239
- :func:`default_matmul` -- the backend name reported by ``costal info``

Every error is a ``ValueError`` carrying the library's message (lib.rs:31-33); wrong dtypes raise
``TypeError`` as PyO3's array extraction does.

The ``matmul`` argument names the coarse-tile backend. On this build it is an MFMA precision:
``"mfma-f32"`` (plain FP32 tiles), ``"mfma-f16"`` (FP16 tiles whose candidates are re-scored in
f32) or ``"mfma-f16x3"`` (FP16 hi/lo split tiles, three FP16 MFMAs, an FP32-class error bound; the
default); ``"gemm"`` is accepted as the reference's portable name and means ``"mfma-f32"``;
``"accelerate"`` is unavailable here, as on the reference's Linux builds. The output never depends
on the backend. ``sort_by_norm`` and ``n_threads`` are accepted for compatibility and cannot
change the output (the GPU path has its own working order and no host threads); ``prune`` switches
the fused scan's exact tile pruning (on by default, as the reference's knob), which cannot change
the output either.
"""

from __future__ import annotations

import ctypes as C
import json

import numpy as np

from . import _lib

__version__ = "2.1.0+mi355x.0.1.0"

NOT_CONTIGUOUS = "input must be a C-contiguous array; call np.ascontiguousarray first"

_BACKENDS = {"mfma-f32": "fp32", "mfma-f16": "fp16", "mfma-f16x3": "fp16x3", "gemm": "fp32"}
_DEFAULT_MATMUL = "mfma-f16x3"


def default_matmul() -> str:
    return _DEFAULT_MATMUL


def _q(name) -> str:
    # Rust's {:?} on a &str: double-quoted
    return json.dumps(str(name))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _matrix_f32(x, what: str = "x") -> np.ndarray:
    if not isinstance(x, np.ndarray) or x.ndim != 2:
        raise TypeError(f"argument '{what}': expected a 2-dimensional numpy array")
    if x.dtype != np.float32:
        raise TypeError(f"argument '{what}': expected float32, got {x.dtype}")
    if not x.flags.c_contiguous:
        raise ValueError(NOT_CONTIGUOUS)
    return x


def _parse_select(select) -> str:
    if select is None:
        return "cosine"
    if select not in _lib.SELECT:
        raise ValueError(f"unknown select {_q(select)}")
    return select


def _parse_backend(matmul) -> tuple[str, str]:
    name = _DEFAULT_MATMUL if matmul is None else matmul
    if name == "accelerate":
        raise ValueError("matmul backend accelerate is not available on this platform")
    if name not in _BACKENDS:
        raise ValueError(f"unknown matmul backend {_q(name)}")
    return name, _BACKENDS[name]


def _call(fn, *args, stats=None):
    rc = fn(*args)
    if rc != _lib.OK:
        row = sample = None
        if rc == _lib.E_NON_FINITE and stats is not None and stats.nonfinite_row != _lib.NO_ROW:
            row = int(stats.nonfinite_row)
            sample = None if stats.nonfinite_sample == _lib.NO_ROW else int(stats.nonfinite_sample)
        raise _lib.CostalError(rc, _lib.load().costal_last_error().decode(), row, sample)


def _knn_exact(x, k, *, select=None, margin=None, row_block=None, col_block=None, matmul=None,
               sort_by_norm=True, prune=True, n_threads=None, row_threshold=None,
               weights=False, col_splits=None, threshold_model=None, threshold_bf16=True,
               prune_clusters=None, prune_tile=None):
    """Exact kNN: ``((idx int32[n,k], cos f64[n,k], norms f64[n]), stats)``.

    Extensions beyond the reference signature (keyword-only, default off):
    ``row_threshold`` -- per-row admission thresholds (float32[n]); the result stays exact.
    ``weights=True`` -- also return the fused Tanimoto weights as ``stats["weights"]``.
    ``col_splits`` -- number of candidate-range splits (speed only; 0/None = tuned).
    ``threshold_model`` -- a :class:`costal_amd.learned.ThresholdModel`; its thresholds are computed
    on the GPU inside the call (BF16 unless ``threshold_bf16=False``); the result stays exact.
    ``prune_clusters`` -- k-means centers of the locality order of the pruned scan (0/None = about
    n/3072); ``prune_tile`` -- rows of a candidate tile of that order, 256 or 512 (0/None = 256, or
    512 past 4096 tiles).  ``prune`` keeps the reference's meaning of "skip what provably cannot qualify": here
    candidate tiles whose cosine bound is below every list's threshold (speed only).
    """
    x = _matrix_f32(x)
    n, d = x.shape
    if n > np.iinfo(np.int32).max:
        raise ValueError(f"n={n} exceeds the int32 index range")
    sel = _parse_select(select)
    backend, prec = _parse_backend(matmul)
    k = int(k)
    if k < 0:
        raise OverflowError("can't convert negative int to unsigned")
    m = k if margin is None else int(margin)
    p = _lib.KnnParams()
    p.k = k
    p.margin = m
    p.select = _lib.SELECT[sel]
    p.precision = _lib.PRECISION[prec]
    p.row_block = int(row_block or 0)
    p.col_block = int(col_block or 0)
    p.col_splits = int(col_splits or 0)
    p.no_prune = 0 if prune else 1
    p.prune_clusters = int(prune_clusters or 0)
    p.prune_tile = int(prune_tile or 0)
    th = None
    if row_threshold is not None:
        th = np.ascontiguousarray(row_threshold, dtype=np.float32)
        if th.shape != (n,):
            raise ValueError("row_threshold must have one entry per row")
        p.row_threshold = th.ctypes.data
    mp = None
    if threshold_model is not None and row_threshold is None:
        if threshold_model.d != d:
            raise ValueError(f"threshold model takes {threshold_model.d} features, x has {d}")
        mp = np.ascontiguousarray(threshold_model.params, dtype=np.float32)
        p.mlp_params = mp.ctypes.data
        p.mlp_offset = float(threshold_model.offset)
        if getattr(threshold_model, "offset2", None) is not None:
            p.mlp_offset2 = float(threshold_model.offset2)
        p.mlp_bf16 = int(bool(threshold_bf16))
    idx = np.empty((n, max(k, 0)), np.uint32)
    cos = np.empty((n, max(k, 0)), np.float64)
    norms = np.empty(n, np.float64)
    w = np.empty((n, max(k, 0)), np.float64) if weights else None
    st = _lib.KnnStats()
    L = _lib.load()
    _call(L.costal_knn_exact, _ptr(x), n, d, C.byref(p), _ptr(idx), _ptr(cos), _ptr(norms),
          _ptr(w) if w is not None else None, C.byref(st), stats=st)
    stats = {
        "fallback_rows": int(st.fallback_rows),
        "row_block": int(st.row_block),
        "col_block": int(st.col_block),
        "matmul": backend,
        "margin": m,
        "select": sel,
        "blocks_skipped": int(st.blocks_skipped),
        "blocks_total": int(st.blocks_total),
        # GPU-side extras
        "precision": prec,
        "candidates": int(st.candidates),
        "col_splits": int(st.col_splits),
        "short_rows": int(st.short_rows),
        "path": {0: "fused", 1: "rowscan", 2: "wide"}.get(int(st.path), str(st.path)),
        "inserts": int(st.inserts),
        "rowscan_rows": int(st.rowscan_rows),
        "device_ms": float(st.device_ms),
        "coarse_err_max": float(st.coarse_err_max),
        "coarse_eps": float(st.coarse_eps),
        "kernel_ms": {"prep": st.prep_ms, "order": st.order_ms, "coarse": st.coarse_ms, "rescore": st.rescore_ms,
                      "fallback": st.fallback_ms},
    }
    if w is not None:
        stats["weights"] = w
    return (idx.view(np.int32), cos, norms), stats


def _knn_bruteforce(x, k, *, select=None, n_threads=None):
    """Brute force by the definition (f64 everywhere, no blocking), on the GPU."""
    x = _matrix_f32(x)
    n, d = x.shape
    sel = _parse_select(select)
    k = int(k)
    if k < 0:
        raise OverflowError("can't convert negative int to unsigned")
    idx = np.empty((n, k), np.uint32)
    cos = np.empty((n, k), np.float64)
    norms = np.empty(n, np.float64)
    _call(_lib.load().costal_knn_bruteforce, _ptr(x), n, d, k, _lib.SELECT[sel], _ptr(idx),
          _ptr(cos), _ptr(norms))
    return idx.view(np.int32), cos, norms


def _neighbours(idx) -> tuple[int, int, np.ndarray]:
    if not isinstance(idx, np.ndarray) or idx.ndim != 2 or idx.dtype != np.int32:
        raise TypeError("argument 'idx': expected a 2-dimensional int32 numpy array")
    if not idx.flags.c_contiguous:
        raise ValueError(NOT_CONTIGUOUS)
    n, k = idx.shape
    flat = idx.ravel()
    bad = np.flatnonzero((flat < 0) | (flat >= n))
    if bad.size:
        raise ValueError(f"neighbor index {int(flat[bad[0]])} out of range (n={n})")
    return n, k, flat.view(np.uint32)


def _reweight(idx, cos, norms, weight):
    """Edge weights on a kNN ("tanimoto" / "cosine" / "jaccard"), float64[n, k]."""
    if weight not in _lib.WEIGHT:
        raise ValueError(f"unknown edge weight {_q(weight)}")
    n, k, u = _neighbours(idx)
    if not isinstance(cos, np.ndarray) or cos.dtype != np.float64:
        raise TypeError("argument 'cos': expected a float64 numpy array")
    if not isinstance(norms, np.ndarray) or norms.dtype != np.float64:
        raise TypeError("argument 'norms': expected a float64 numpy array")
    if cos.shape != (n, k) or norms.shape != (n,):
        raise ValueError("shapes of idx, cos and norms do not match")
    if not (cos.flags.c_contiguous and norms.flags.c_contiguous):
        raise ValueError(NOT_CONTIGUOUS)
    out = np.empty((n, k), np.float64)
    _call(_lib.load().costal_reweight, n, k, _ptr(u), _ptr(cos), _ptr(norms), _lib.WEIGHT[weight],
          _ptr(out))
    return out


def _fuzzy_union(idx, w):
    """Fuzzy union of a weighted kNN: upper-triangular ``(rows int32, cols int32, w float64)``."""
    n, k, u = _neighbours(idx)
    if not isinstance(w, np.ndarray) or w.dtype != np.float64:
        raise TypeError("argument 'w': expected a float64 numpy array")
    if w.shape != (n, k):
        raise ValueError("shapes of idx and w do not match")
    w = np.ascontiguousarray(w)
    rows = np.empty(n * k, np.uint32)
    cols = np.empty(n * k, np.uint32)
    wo = np.empty(n * k, np.float64)
    m = C.c_uint64(0)
    _call(_lib.load().costal_fuzzy_union, n, k, _ptr(u), _ptr(w), _ptr(rows), _ptr(cols), _ptr(wo),
          C.byref(m))
    m = m.value
    return rows[:m].view(np.int32).copy(), cols[:m].view(np.int32).copy(), wo[:m].copy()


def core_version() -> str:
    return _lib.version()


class Csr:
    """The undirected weighted graph the CPU Leiden consumes, in compressed rows
    (ike12th/CosTaL crates/costal-core/src/leiden.rs:25-100): ``indptr`` uint64[n+1], ``nbr``
    uint32 and ``w`` float64 holding both directions of every edge with a positive weight."""

    def __init__(self, indptr, nbr, w):
        self.indptr, self.nbr, self.w = indptr, nbr, w

    @property
    def n(self) -> int:
        return len(self.indptr) - 1

    def n_edges(self) -> int:
        return len(self.nbr) // 2

    def strengths(self) -> np.ndarray:
        """k_i = the sum of row i's weights, in row order."""
        out = np.zeros(self.n)
        for v in range(self.n):
            s = 0.0
            for x in self.w[int(self.indptr[v]):int(self.indptr[v + 1])]:
                s += float(x)
            out[v] = s
        return out

    @classmethod
    def from_edges(cls, n, rows, cols, w) -> "Csr":
        """Csr::from_edges (leiden.rs:52-100): each undirected edge once (such as _fuzzy_union's
        output), rows listing neighbors in edge order, exact-zero weights dropped; ValueError with
        the reference's message for mismatched lengths, out-of-range nodes, self-loops and negative
        or non-finite weights."""
        rows = np.asarray(rows)
        cols = np.asarray(cols)
        w = np.asarray(w, dtype=np.float64)
        if not (len(rows) == len(cols) == len(w)):
            raise ValueError(f"edge arrays differ in length: {len(rows)} rows, {len(cols)} cols, {len(w)} weights")
        n = int(n)
        m = len(rows)
        r = np.ascontiguousarray(rows, dtype=np.int64)
        c = np.ascontiguousarray(cols, dtype=np.int64)
        # indices that do not fit u32 are out of range whatever n is (n < 2^32 is checked below)
        big = np.flatnonzero((r < 0) | (c < 0) | (r > 0xFFFFFFFF) | (c > 0xFFFFFFFF))
        if big.size and n <= 0xFFFFFFFF:
            e = int(big[0])
            raise ValueError(f"edge {e} ({int(r[e])}, {int(c[e])}) is out of range (n={n})")
        r32 = np.ascontiguousarray(r, dtype=np.uint32)
        c32 = np.ascontiguousarray(c, dtype=np.uint32)
        w = np.ascontiguousarray(w)
        indptr = np.empty(n + 1, np.uint64)
        nbr = np.empty(max(2 * m, 1), np.uint32)
        wo = np.empty(max(2 * m, 1), np.float64)
        nnz = C.c_uint64(0)
        _call(_lib.load().costal_csr_from_edges, n, m, _ptr(r32), _ptr(c32), _ptr(w), _ptr(indptr), _ptr(nbr),
              _ptr(wo), C.byref(nnz))
        k = nnz.value
        return cls(indptr, nbr[:k].copy(), wo[:k].copy())
