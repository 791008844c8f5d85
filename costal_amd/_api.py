"""CosTaL's public graph-construction API on the gfx950 library.

Same functions, signatures, defaults, warnings and errors as ike12th/CosTaL
``python/costal/_api.py``:

- :func:`knn` (_api.py:163-218), :class:`KnnGraph` (:37-60), input conversion ``_as_matrix``
  (:89-100), knob validation ``_check_choices`` (:66-86), input diagnostics ``_diagnose``
  (:112-144);
- :func:`reweight` (:242-267), :func:`truncate` (:270-281), :func:`build_graph` (:284-321);
- :func:`find_communities` (:350-490) with Leiden delegated to ``leidenalg``/``igraph`` when they
  are installed (Leiden itself is CPU work outside this library).
"""

from __future__ import annotations

import warnings
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from . import _costal as _ext

__all__ = ["CostalWarning", "KnnGraph", "build_graph", "find_communities", "knn", "reweight",
           "truncate"]


class CostalWarning(UserWarning):
    """An input-diagnostics note: the result is correct, but a knob may not mean what was meant."""


@dataclass(frozen=True)
class KnnGraph:
    """An exact kNN: ``idx`` (n, k) int32 original indices ranked by (similarity desc, index asc),
    ``cos`` (n, k) float64 rescored cosines, ``norms`` (n,) float64 raw row norms, ``meta``."""

    idx: np.ndarray
    cos: np.ndarray
    norms: np.ndarray
    meta: dict[str, Any] = field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.idx.shape[0]

    @property
    def k(self) -> int:
        return self.idx.shape[1]


_WEIGHTS = ("tanimoto", "cosine", "jaccard")


def _check_choices(select, normalize, backend, compat):
    if compat not in (None, "paper"):
        raise ValueError(f"compat must be None or 'paper', got {compat!r}")
    if select not in ("cosine", "tanimoto"):
        raise ValueError(f"select must be 'cosine' or 'tanimoto', got {select!r}")
    if normalize is None:
        normalize = "l2" if select == "cosine" else "none"
    if normalize not in ("l2", "none"):
        raise ValueError(f"normalize must be 'l2' or 'none', got {normalize!r}")
    if backend in ("approx", "hnsw"):
        raise NotImplementedError("approximate backends never enter the exact path")
    if backend == "sparse":
        raise NotImplementedError("the sparse backend is not implemented")
    if backend != "exact":
        raise ValueError(f"backend must be 'exact', got {backend!r}")
    return normalize


def _as_matrix(X) -> np.ndarray:
    if hasattr(X, "tocsr") or hasattr(X, "toarray"):
        raise TypeError("sparse input is not supported by the exact dense backend; use .toarray() "
                        "or reduce dimensionality first")
    X = np.asarray(X)
    if X.ndim != 2:
        raise ValueError(f"X must be a 2-D matrix, got {X.ndim} dimensions")
    if not (np.issubdtype(X.dtype, np.floating) or np.issubdtype(X.dtype, np.integer)):
        raise TypeError(f"X must be a real-valued matrix, got dtype={X.dtype}")
    return np.ascontiguousarray(X, dtype=np.float32)


_BLOCK_VALUES = 1 << 20  # entries of X converted to float64 at a time by _diagnose


def _diagnose(X: np.ndarray, norms: np.ndarray) -> dict[str, Any]:
    """Record-only input statistics; the data is never modified.

    ``row_norm_cv``: standard deviation over mean of the exact f64 row norms (0 for a zero mean).
    ``columns_centered``: some column varies, and every varying column has |mean| <= 0.05 std.
    ``negative_fraction`` / ``nonzero_fraction``: shares of negative and of non-zero entries.

    The column moments are merged block by block with the pairwise update of Chan, Golub and
    LeVeque (mean and sum of squared deviations per block, then combined), so the extra memory
    is one block of rows whatever the size of X.  The semantics (keys, thresholds) are those of
    the reference's diagnostics, ike12th/CosTaL python/costal/_api.py:112-144."""
    n, d = X.shape
    block = max(1, _BLOCK_VALUES // max(d, 1))
    seen = 0
    mean = np.zeros(d)
    m2 = np.zeros(d)
    negative = nonzero = 0
    for a in range(0, n, block):
        rows = X[a:a + block]
        negative += int(np.count_nonzero(rows < 0))
        nonzero += int(np.count_nonzero(rows))
        b = rows.astype(np.float64)
        m = len(b)
        b_mean = b.mean(axis=0)
        b_m2 = np.square(b - b_mean).sum(axis=0)
        delta = b_mean - mean
        total = seen + m
        mean = mean + delta * (m / total)
        m2 = m2 + b_m2 + np.square(delta) * (seen * m / total)
        seen = total
    std = np.sqrt(m2 / n) if n else np.zeros(d)
    varying = std > 0
    mu = norms.mean() if n else 0.0
    entries = max(n * d, 1)
    return {
        "row_norm_cv": float(norms.std() / mu) if mu > 0 else 0.0,
        "columns_centered": bool(varying.any() and np.all(np.abs(mean[varying]) <= 0.05 * std[varying])),
        "negative_fraction": negative / entries,
        "nonzero_fraction": nonzero / entries,
    }


def knn(X, k: int = 10, *, select: str = "cosine", normalize: str | None = None,
        backend: str = "exact", compat: str | None = None, n_threads: int | None = None,
        matmul: str | None = None) -> KnnGraph:
    """Exact kNN of the rows of ``X`` (k excludes the row itself), by cosine or, with
    ``select="tanimoto"``, by Tanimoto on the raw rows. ``matmul`` picks the coarse-tile precision
    (``"mfma-f16x3"`` split FP16, the default; ``"mfma-f16"``; ``"mfma-f32"``); the result is
    identical either way."""
    normalize = _check_choices(select, normalize, backend, compat)
    effective = select
    if select == "tanimoto" and normalize == "l2":
        warnings.warn("select='tanimoto' with normalize='l2' is equivalent to cosine selection "
                      "(on unit rows T = c/(2-c) is monotone in c); running cosine",
                      CostalWarning, stacklevel=2)
        effective = "cosine"
    X = _as_matrix(X)
    (idx, cos, norms), stats = _ext._knn_exact(X, int(k), select=effective, matmul=matmul,
                                               n_threads=n_threads)
    meta = {
        "k": int(k),
        "select": select,
        "select_effective": effective,
        "normalize": normalize,
        "backend": backend,
        "compat": compat,
        "version": _ext.__version__,
        "stats": dict(stats),
        "diagnostics": _diagnose(X, norms),
    }
    if effective == "tanimoto":
        _warn_about_magnitude(meta)
    return KnnGraph(idx=idx, cos=cos, norms=norms, meta=meta)


def _warn_about_magnitude(meta) -> None:
    diag = meta.get("diagnostics")
    if not diag:
        return
    if diag["row_norm_cv"] < 1e-4:
        warnings.warn("row norms are nearly constant (rows already unit-normalized?): Tanimoto "
                      "edge weights reduce to the monotone cosine transform c/(2-c)",
                      CostalWarning, stacklevel=3)
    if diag["columns_centered"]:
        warnings.warn("columns are centered (PCA / z-score / Pearson residuals?): Tanimoto edge "
                      "weights change meaning and can be negative", CostalWarning, stacklevel=3)


def reweight(knn: KnnGraph, weight: str = "tanimoto") -> np.ndarray:
    """Edge weights on the same kNN as (n, k) float64: ``"tanimoto"`` (default), ``"cosine"`` or
    ``"jaccard"``. Only the numbers on edges change, never which edges exist."""
    if weight not in _WEIGHTS:
        raise ValueError(f"weight must be one of {_WEIGHTS}, got {weight!r}")
    if weight == "tanimoto":
        _warn_about_magnitude(knn.meta)
    w = _ext._reweight(knn.idx, knn.cos, knn.norms, weight)
    neg = int(np.count_nonzero(w < 0))
    if neg:
        warnings.warn(f"{neg} edges have negative weights (Tanimoto's lower bound on signed input "
                      "is -1/3); they are kept as is", CostalWarning, stacklevel=2)
    return w


def truncate(knn: KnnGraph, k: int) -> KnnGraph:
    """The first k neighbors of every row -- the exact kNN for the smaller k, since rows are
    totally ordered by (cosine desc, index asc)."""
    if not 1 <= k <= knn.k:
        raise ValueError(f"k must be within 1..{knn.k}, got {k}")
    return KnnGraph(idx=np.ascontiguousarray(knn.idx[:, :k]),
                    cos=np.ascontiguousarray(knn.cos[:, :k]), norms=knn.norms,
                    meta={**knn.meta, "k": int(k), "truncated_from": knn.k})


def build_graph(X, k: int = 10, *, select: str = "cosine", weight: str = "tanimoto",
                normalize: str | None = None, backend: str = "exact",
                symmetrize: str = "fuzzy_union", compat: str | None = None,
                n_threads: int | None = None, matmul: str | None = None):
    """kNN -> edge weights -> fuzzy union ``S + S^T - S.S^T``: the upper-triangular edge list
    ``(rows int32, cols int32, weights float64)`` in ascending (row, col) order."""
    if symmetrize != "fuzzy_union":
        raise ValueError(f"symmetrize currently supports only 'fuzzy_union', got {symmetrize!r}")
    g = knn(X, k, select=select, normalize=normalize, backend=backend, compat=compat,
            n_threads=n_threads, matmul=matmul)
    return _ext._fuzzy_union(g.idx, reweight(g, weight))


_LEIDEN_ITERS = {"paper": 10, "igraph": 2}


def find_communities(X, k: int = 10, *, select: str = "cosine", weight: str = "tanimoto",
                     normalize: str | None = None, resolution: float = 0.8, leiden: str = "paper",
                     n_iterations: int | None = None, seed: int | None = None,
                     compat: str | None = None, n_threads: int | None = None, **legacy) -> np.ndarray:
    """kNN -> Tanimoto -> fuzzy union on the GPU, then Leiden on the CPU through leidenalg
    (``leiden="paper"``, the published method) or igraph (``leiden="igraph"``)."""
    if legacy:
        raise TypeError("the CosTaL 1.0 keyword interface is not part of this build")
    if leiden == "fast":
        raise NotImplementedError("the reference's native Rust Leiden is not part of this build; "
                                  "use leiden='paper' or 'igraph'")
    if leiden not in _LEIDEN_ITERS:
        raise ValueError(f"leiden must be one of ('paper', 'fast', 'igraph'), got {leiden!r}")
    iterations = _LEIDEN_ITERS[leiden] if n_iterations is None else int(n_iterations)
    if compat == "paper" and (leiden != "paper" or resolution != 0.8 or iterations != 10
                              or select != "cosine" or weight != "tanimoto"):
        raise ValueError("compat='paper' pins cosine selection, Tanimoto weights and leidenalg "
                         "with resolution 0.8 and 10 iterations")
    try:
        import igraph as ig
    except ImportError as exc:
        raise ImportError(f"leiden={leiden!r} needs igraph: pip install igraph") from exc
    rows, cols, w = build_graph(X, k, select=select, weight=weight, normalize=normalize,
                                compat=compat, n_threads=n_threads)
    n = int(np.asarray(X).shape[0])
    graph = ig.Graph(n=n)
    graph.add_edges(np.column_stack([rows, cols]))
    graph.es["weight"] = w.tolist()
    if leiden == "paper":
        import leidenalg as la

        part = la.find_partition(graph, la.RBConfigurationVertexPartition, weights="weight",
                                 n_iterations=iterations, resolution_parameter=resolution, seed=seed)
        return np.asarray(part.membership, dtype=np.int64)
    if seed is not None:
        import random

        random.seed(seed)
    part = graph.community_leiden(objective_function="modularity", weights="weight",
                                  resolution=resolution, n_iterations=iterations)
    return np.asarray(part.membership, dtype=np.int64)
