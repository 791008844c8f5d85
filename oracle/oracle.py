"""ctypes front end of the CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py``, as the checker and as the timed CPU baseline. The product path
(``costal_amd``) never imports it.

Each function mirrors the reference function it restates (paths in ike12th/CosTaL):

- :func:`knn_bruteforce` -- ``knn_bruteforce_select`` (crates/costal-core/src/knn.rs:884-924)
- :func:`knn_blocked` -- ``run_blocked`` (knn.rs:403-512) in its plain f32 form
- :func:`reweight` -- ``weight::reweight`` (crates/costal-core/src/weight.rs:41-93)
- :func:`tani_from_cos` -- weight.rs:28-37
- :func:`fuzzy_union` -- ``graph::fuzzy_union`` (crates/costal-core/src/graph.rs:24-68)
- :func:`heap_keep` -- ``BlockHeaps::push`` (knn.rs:791-875)
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_ERR = {1: "shape", 2: "bad_k", 3: "too_many_rows", 4: "non_finite", 8: "invalid", 9: "oom"}

_P = C.c_void_p
_U64 = C.c_uint64


def build() -> None:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_dot64.restype = C.c_double
        L.oracle_dot64.argtypes = [_P, _P, _U64]
        L.oracle_tani_from_cos.restype = C.c_double
        L.oracle_tani_from_cos.argtypes = [C.c_double] * 3
        L.oracle_row_norms64.argtypes = [_P, _U64, _U64, _P, C.c_int]
        L.oracle_knn_bruteforce.argtypes = [_P, _U64, _U64, _U64, C.c_int, _P, _P, _P, C.c_int]
        L.oracle_block_sizes.argtypes = [_U64, _U64, _U64, _U64, C.c_int, _P, _P]
        L.oracle_block_sizes.restype = None
        L.oracle_knn_blocked.argtypes = [_P, _U64, _U64, _U64, _U64, _U64, _U64, _U64, C.c_int,
                                         C.c_int, _U64, _U64, _P, _P, _P, _P]
        L.oracle_heap_keep.restype = _U64
        L.oracle_heap_keep.argtypes = [_P, _P, _U64, _U64, _P]
        L.oracle_reweight.argtypes = [_U64, _U64, _P, _P, _P, C.c_int, _P]
        L.oracle_fuzzy_union.argtypes = [_U64, _U64, _P, _P, _P, _P, _P, _P]
        L.oracle_csr_from_edges.argtypes = [_U64, _U64, _P, _P, _P, _P, _P, _P]
        _lib = L
    return _lib


class OracleError(ValueError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code
        self.kind = _ERR.get(code, "unknown")


def _check(rc: int) -> None:
    if rc:
        raise OracleError(rc, lib().oracle_last_error().decode())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _f32(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.float32)


def row_norms64(x, n_threads: int = 0) -> np.ndarray:
    x = _f32(x)
    n, d = x.shape
    out = np.empty(n, np.float64)
    _check(lib().oracle_row_norms64(_ptr(x), n, d, _ptr(out), n_threads))
    return out


def knn_bruteforce(x, k: int, *, select: str = "cosine", n_threads: int = 0):
    """(idx int32[n,k], cos f64[n,k], norms f64[n]) by the brute-force definition."""
    x = _f32(x)
    n, d = x.shape
    idx = np.empty((n, k), np.uint32)
    cos = np.empty((n, k), np.float64)
    norms = np.empty(n, np.float64)
    _check(lib().oracle_knn_bruteforce(_ptr(x), n, d, k, int(select == "tanimoto"), _ptr(idx),
                                       _ptr(cos), _ptr(norms), n_threads))
    return idx.astype(np.int32), cos, norms


def knn_blocked(x, k: int, *, margin: int | None = None, row_block: int = 0, col_block: int = 0,
                budget_bytes: int = 1 << 20, sort_by_norm: bool = True, n_threads: int = 0,
                rows: tuple[int, int] | None = None):
    """The blocked form. Returns ((idx, cos, norms), stats). ``rows`` restricts the computation to
    a range of query rows (needs ``sort_by_norm=False``); other rows of idx/cos are left at 0."""
    x = _f32(x)
    n, d = x.shape
    r0, r1 = rows if rows is not None else (0, n)
    idx = np.zeros((n, k), np.uint32)
    cos = np.zeros((n, k), np.float64)
    norms = np.empty(n, np.float64)
    st = np.zeros(4, np.uint64)
    _check(lib().oracle_knn_blocked(_ptr(x), n, d, k, k if margin is None else margin, row_block,
                                    col_block, budget_bytes, int(sort_by_norm), n_threads, r0, r1,
                                    _ptr(idx), _ptr(cos), _ptr(norms), _ptr(st)))
    stats = {"fallback_rows": int(st[0]), "row_block": int(st[1]), "col_block": int(st[2]),
             "blocks_total": int(st[3])}
    return (idx.astype(np.int32), cos, norms), stats


def heap_keep(pairs, cap: int) -> list[int]:
    s = np.array([p[0] for p in pairs], np.float32)
    i = np.array([p[1] for p in pairs], np.uint32)
    kept = np.zeros(cap, np.uint32)
    m = lib().oracle_heap_keep(_ptr(s), _ptr(i), len(pairs), cap, _ptr(kept))
    return kept[:m].tolist()


def tani_from_cos(a: float, b: float, c: float) -> float:
    return lib().oracle_tani_from_cos(a, b, c)


_WEIGHT = {"tanimoto": 0, "cosine": 1, "jaccard": 2}


def reweight(idx, cos, norms, weight: str = "tanimoto") -> np.ndarray:
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    cos = np.ascontiguousarray(cos, dtype=np.float64)
    norms = np.ascontiguousarray(norms, dtype=np.float64)
    n, k = idx.shape
    out = np.empty((n, k), np.float64)
    _check(lib().oracle_reweight(n, k, _ptr(idx), _ptr(cos), _ptr(norms), _WEIGHT[weight], _ptr(out)))
    return out


def fuzzy_union(idx, w):
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    w = np.ascontiguousarray(w, dtype=np.float64)
    n, k = idx.shape
    rows = np.empty(n * k, np.uint32)
    cols = np.empty(n * k, np.uint32)
    wo = np.empty(n * k, np.float64)
    m = C.c_uint64(0)
    _check(lib().oracle_fuzzy_union(n, k, _ptr(idx), _ptr(w), _ptr(rows), _ptr(cols), _ptr(wo),
                                    C.byref(m)))
    m = m.value
    return rows[:m].astype(np.int32), cols[:m].astype(np.int32), wo[:m].copy()


def csr_from_edges(n: int, rows, cols, w):
    """leiden.rs:52-100: (indptr uint64[n+1], nbr uint32[nnz], w float64[nnz])."""
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    cols = np.ascontiguousarray(cols, dtype=np.uint32)
    w = np.ascontiguousarray(w, dtype=np.float64)
    m = len(rows)
    indptr = np.empty(n + 1, np.uint64)
    nbr = np.empty(max(2 * m, 1), np.uint32)
    wo = np.empty(max(2 * m, 1), np.float64)
    _check(lib().oracle_csr_from_edges(n, m, _ptr(rows), _ptr(cols), _ptr(w), _ptr(indptr), _ptr(nbr), _ptr(wo)))
    nnz = int(indptr[n])
    return indptr, nbr[:nnz].copy(), wo[:nnz].copy()


def dot64(a, b) -> float:
    a = _f32(a)
    b = _f32(b)
    return lib().oracle_dot64(_ptr(a), _ptr(b), a.size)
