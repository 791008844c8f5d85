"""Autotuner for the fused scan: coarse-tile precision, candidate-list width, candidate splits and
the locality order of the pruned scan (its k-means center count, or no pruning).

For a shape bucket (feature count rounded up to 8, row count rounded to a power of two, k) it times
the whole call (ordering, scan, rescoring, fallback) on synthetic data of that shape (an arcsinh
Gaussian mixture like cytometry data) over a small grid of configurations and keeps the fastest:
first precision x list margin x splits under the default order, then, for the best of those, other
k-means center counts and the unpruned scan.  Results are cached in a JSON file
(``$COSTAL_TUNE_CACHE``, default ``~/.cache/costal_amd/tune.json``).  None of these knobs can change
the output; they only trade candidate-list inserts, rescoring work, skipped tiles and parallelism.
"""

from __future__ import annotations

import ctypes as C
import json
import math
import os

import numpy as np

from . import _lib

DEFAULT_CACHE = os.environ.get("COSTAL_TUNE_CACHE",
                               os.path.join(os.path.expanduser("~"), ".cache", "costal_amd", "tune.json"))


def bucket(n: int, d: int, k: int) -> str:
    nb = 1 << max(0, int(round(math.log2(max(n, 1)))))
    db = -(-d // 8) * 8
    return f"d{db}:n{nb}:k{k}"


def _synthetic(n: int, d: int, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng([seed, n, d])
    comps = int(rng.integers(10, 31))
    means = rng.normal(0.0, 20.0, size=(comps, d))
    w = rng.dirichlet(np.ones(comps))
    c = rng.choice(comps, size=n, p=w)
    x = means[c] + rng.normal(0.0, 5.0, size=(n, d))
    return np.ascontiguousarray(np.arcsinh(x / 5.0), dtype=np.float32)


KNOBS = ("precision", "margin", "col_splits", "prune_clusters", "prune_tile", "prune")


def candidates(k: int):
    margins = sorted({max(0, k // 2), k, min(2 * k, 64 - k)})
    margins = [m for m in margins if k + m <= 64]
    for prec in ("fp16x3", "fp32"):
        for m in margins:
            for ns in (0, 1, 4):
                yield {"precision": prec, "margin": m, "col_splits": ns, "prune_clusters": 0, "prune_tile": 0,
                       "prune": True}


def order_candidates(best: dict):
    """The order knobs for the best configuration of the first stage: k-means centers (0 = the
    library's default, about n/3072), candidate tiles of 512 rows instead of 256 (the tile shape
    of the pruned scan), and the unpruned scan."""
    for c in (32, 128):
        yield {**best, "prune_clusters": c, "prune_tile": 0, "prune": True}
    yield {**best, "prune_clusters": 0, "prune_tile": 512, "prune": True}
    yield {**best, "prune_clusters": 0, "prune_tile": 0, "prune": False}


def time_config(x_dev, n: int, d: int, k: int, cfg: dict, outs, reps: int = 3) -> float:
    L = _lib.load()
    p = _lib.KnnParams()
    p.k = k
    p.margin = cfg["margin"]
    p.precision = _lib.PRECISION[cfg["precision"]]
    p.col_splits = cfg["col_splits"]
    p.prune_clusters = cfg.get("prune_clusters", 0)
    p.prune_tile = cfg.get("prune_tile", 0)
    p.no_prune = 0 if cfg.get("prune", True) else 1
    st = _lib.KnnStats()
    times = []
    for r in range(reps + 1):
        _lib.check(L.costal_knn_exact_device(C.c_void_p(x_dev.data_ptr()), n, d, C.byref(p),
                                             *[C.c_void_p(o.data_ptr()) for o in outs], C.byref(st), None))
        if r:
            times.append(st.device_ms)
    return float(np.median(times))


def choose(results: list, tie: float = 0.03) -> dict:
    """The configuration to keep.  Those within ``tie`` (timing noise) of the fastest are ties:
    take the shortest lists (the fewest inserts, which is what a batch of many samples pays for),
    then split FP16, then tuned splits, then fewer splits, so the choice does not flip with the
    noise of one timing run."""
    fastest = min(r["ms"] for r in results)
    near = [r for r in results if r["ms"] <= (1.0 + tie) * fastest]
    return min(near, key=lambda r: (r["margin"], r["precision"] != "fp16x3", r["col_splits"] != 0, r["col_splits"],
                                    not r.get("prune", True), r.get("prune_clusters", 0) != 0,
                                    r.get("prune_tile", 0) != 0))


def tune_shape(n: int, d: int, k: int, *, reps: int = 3, seed: int = 0) -> dict:
    import torch

    x = _synthetic(n, d, seed)
    dev = torch.device("cuda:0")
    xd = torch.from_numpy(x).to(dev)
    outs = [torch.empty((n, k), dtype=torch.int32, device=dev), torch.empty((n, k), dtype=torch.float64, device=dev),
            torch.empty(n, dtype=torch.float64, device=dev), torch.empty((n, k), dtype=torch.float64, device=dev)]
    results = []
    for cfg in candidates(k):
        ms = time_config(xd, n, d, k, cfg, outs, reps)
        results.append({**cfg, "ms": ms})
    best = choose(results)
    for cfg in order_candidates({kk: best[kk] for kk in KNOBS}):
        ms = time_config(xd, n, d, k, cfg, outs, reps)
        results.append({**cfg, "ms": ms})
    best = choose(results)
    return {"best": {kk: best[kk] for kk in KNOBS}, "ms": best["ms"], "trials": results, "shape": [n, d, k]}


class Tuner:
    """Cached per-bucket choices; ``get`` tunes a bucket the first time it is asked for."""

    def __init__(self, cache_path: str | None = DEFAULT_CACHE):
        self.path = cache_path
        self.table = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.table = json.load(f)

    def save(self) -> None:
        if not self.path:
            return
        os.makedirs(os.path.dirname(self.path) or ".", exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.table, f, indent=1)
        os.replace(tmp, self.path)

    def get(self, n: int, d: int, k: int) -> dict:
        """The kept configuration of the bucket of (n, d, k), every knob present (entries cached
        before the order knobs existed get the defaults: pruned, default center count)."""
        key = bucket(n, d, k)
        if key not in self.table:
            nb = int(key.split(":")[1][1:])
            db = int(key.split(":")[0][1:])
            self.table[key] = tune_shape(max(nb, k + 2), db, k)
            self.save()
        return {"prune_clusters": 0, "prune_tile": 0, "prune": True, **self.table[key]["best"]}

    def plan_batch(self, shapes, k: int, fill_blocks: int = 2048):
        """Per-sample (margin, col_splits) and the batch precision: the precision chosen for the
        buckets holding most of the scan work (sum of n^2 d).  A bucket is tuned as a single call,
        where candidate splits add parallelism; a batch whose samples hold at least
        ``fill_blocks`` query blocks (128 rows) fills the GPU without them, and every split would
        fill its candidate lists again, so such a batch scans each sample whole."""
        plans = [self.get(n, d, k) for n, d in shapes]
        work = {}
        for (n, d), pl in zip(shapes, plans):
            work[pl["precision"]] = work.get(pl["precision"], 0.0) + float(n) * n * d
        prec = max(work, key=work.get)
        splits = [pl["col_splits"] for pl in plans]
        if sum(-(-n // 128) for n, _ in shapes) >= fill_blocks:
            splits = [1] * len(shapes)
        return prec, [pl["margin"] for pl in plans], splits


def knn_kwargs(best: dict) -> dict:
    """Keyword arguments of ``_costal._knn_exact`` for a kept configuration."""
    matmul = {"fp16x3": "mfma-f16x3", "fp32": "mfma-f32", "fp16": "mfma-f16"}[best["precision"]]
    return {"matmul": matmul, "margin": best["margin"], "col_splits": best["col_splits"],
            "prune_clusters": best.get("prune_clusters", 0), "prune_tile": best.get("prune_tile", 0),
            "prune": best.get("prune", True)}
