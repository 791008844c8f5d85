"""Many independent samples in one batched call (costal_knn_batched_device).

Each sample's result is exactly what ``_knn_exact`` gives for it alone.  The library runs the batch
as one persistent scan launch over a (sample, query block, candidate split) work queue, captured
once as a hipGraph and replayed while the batch layout is unchanged.  Device buffers are torch
tensors (plumbing only).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

_PREC = {"mfma-f32": "fp32", "mfma-f16": "fp16", "mfma-f16x3": "fp16x3", "gemm": "fp32"}


class Batch:
    """Samples resident on the GPU, with their output buffers, for repeated batched calls."""

    def __init__(self, xs: list[np.ndarray], k: int, *, margins=None, col_splits=None, weights=True):
        import torch

        self.torch = torch
        self.k = int(k)
        self.xs = [np.ascontiguousarray(x, dtype=np.float32) for x in xs]
        dev = torch.device("cuda:0")
        self.bufs = []
        arr = (_lib.Sample * len(xs))()
        for i, x in enumerate(self.xs):
            n, d = x.shape
            b = {
                "x": torch.from_numpy(x).to(dev),
                "idx": torch.empty((n, self.k), dtype=torch.int32, device=dev),
                "cos": torch.empty((n, self.k), dtype=torch.float64, device=dev),
                "norms": torch.empty(n, dtype=torch.float64, device=dev),
                "w": torch.empty((n, self.k), dtype=torch.float64, device=dev) if weights else None,
            }
            self.bufs.append(b)
            a = arr[i]
            a.x = b["x"].data_ptr()
            a.n, a.d = n, d
            a.idx = b["idx"].data_ptr()
            a.cos = b["cos"].data_ptr()
            a.norms = b["norms"].data_ptr()
            a.weights = b["w"].data_ptr() if weights else None
            a.margin = -1 if margins is None else int(margins[i])
            a.col_splits = 0 if col_splits is None else int(col_splits[i])
        self.samples = arr
        self.stream = torch.cuda.current_stream().cuda_stream
        torch.cuda.synchronize()

    def run(self, matmul: str = "mfma-f16x3"):
        p = _lib.KnnParams()
        p.k = self.k
        p.margin = -1
        p.precision = _lib.PRECISION[_PREC[matmul]]
        st = _lib.KnnStats()
        rc = _lib.load().costal_knn_batched_device(self.samples, len(self.xs), C.byref(p), C.byref(st),
                                                   C.c_void_p(self.stream))
        if rc != _lib.OK:
            row = sample = None
            if rc == _lib.E_NON_FINITE and st.nonfinite_row != _lib.NO_ROW:
                row, sample = int(st.nonfinite_row), int(st.nonfinite_sample)
            raise _lib.CostalError(rc, _lib.load().costal_last_error().decode(), row, sample)
        return st

    def outputs(self):
        out = []
        for b in self.bufs:
            out.append((b["idx"].cpu().numpy(), b["cos"].cpu().numpy(), b["norms"].cpu().numpy(),
                        None if b["w"] is None else b["w"].cpu().numpy()))
        return out


def knn_batched(xs: list[np.ndarray], k: int, *, matmul: str = "mfma-f16x3", margins=None, col_splits=None):
    """Exact kNN of every sample: a list of (idx int32, cos f64, norms f64, tanimoto f64) and the
    batch statistics."""
    b = Batch(xs, k, margins=margins, col_splits=col_splits)
    st = b.run(matmul)
    return b.outputs(), {"fallback_rows": int(st.fallback_rows), "inserts": int(st.inserts),
                         "col_splits": int(st.col_splits), "candidates": int(st.candidates),
                         "device_ms": float(st.device_ms), "coarse_err_max": float(st.coarse_err_max),
                         "coarse_eps": float(st.coarse_eps)}
