"""costal_knn_exact_device_noalloc: the device entry point without any allocation, against
costal_knn_exact_device on the same inputs (bit-exact idx / cos / norms, weights within 1e-12),
two calls in a row with the second one smaller, and a call too large for the static scratch
refused with a clear error."""

import ctypes as C

import numpy as np
import pytest

from costal_amd import _lib, synth

pytestmark = pytest.mark.gpu


def device_call(fn_name, x, k, matmul="fp16x3", threshold=None, select="cosine"):
    import torch

    L = _lib.load()
    n, d = x.shape
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(x).to(dev)
    idx = torch.full((n, k), -1, dtype=torch.int32, device=dev)
    cos = torch.full((n, k), np.nan, dtype=torch.float64, device=dev)
    norms = torch.full((n,), np.nan, dtype=torch.float64, device=dev)
    w = torch.full((n, k), np.nan, dtype=torch.float64, device=dev)
    p = _lib.KnnParams()
    p.k = k
    p.margin = -1
    p.precision = _lib.PRECISION[matmul]
    p.select = _lib.SELECT[select]
    th = None
    if threshold is not None:
        th = torch.from_numpy(threshold).to(dev)
        p.row_threshold = th.data_ptr()
    st = _lib.KnnStats()
    stream = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    rc = getattr(L, fn_name)(C.c_void_p(dx.data_ptr()), n, d, C.byref(p), C.c_void_p(idx.data_ptr()),
                             C.c_void_p(cos.data_ptr()), C.c_void_p(norms.data_ptr()), C.c_void_p(w.data_ptr()),
                             C.byref(st), C.c_void_p(stream))
    _lib.check(rc)
    torch.cuda.synchronize()
    return idx.cpu().numpy(), cos.cpu().numpy(), norms.cpu().numpy(), w.cpu().numpy(), st


@pytest.mark.parametrize("matmul", ["fp16x3", "fp32"])
def test_noalloc_equals_device_entry_point_twice(matmul):
    rng = np.random.default_rng(3)
    for n in (60_000, 5_000):  # the second call smaller: the scratch keeps the first call's data
        x = synth.rare_396460x15(n=n, sample_seed=5)
        th = None
        if n == 5_000:
            th = rng.uniform(0.5, 1.0, size=n).astype(np.float32)
        print(f"noalloc {matmul} n={n}", flush=True)
        a = device_call("costal_knn_exact_device", x, 10, matmul, th)
        b = device_call("costal_knn_exact_device_noalloc", x, 10, matmul, th)
        np.testing.assert_array_equal(b[0], a[0])
        np.testing.assert_array_equal(b[1], a[1])
        np.testing.assert_array_equal(b[2], a[2])
        np.testing.assert_allclose(b[3], a[3], rtol=1e-12, atol=0)
        assert b[4].fallback_rows == a[4].fallback_rows and b[4].candidates == a[4].candidates


def test_noalloc_tanimoto_selection_twice():
    """Tanimoto selection (its norm tables also live in the static scratch), two calls, the second
    smaller."""
    for n in (40_000, 3_000):
        x = synth.cytometry_86864x39(n=n)
        x = np.ascontiguousarray(x * np.random.default_rng(n).uniform(0.5, 3.0, size=(n, 1)).astype(np.float32))
        print(f"noalloc tanimoto n={n}", flush=True)
        a = device_call("costal_knn_exact_device", x, 10, select="tanimoto")
        b = device_call("costal_knn_exact_device_noalloc", x, 10, select="tanimoto")
        for u, v in zip(a[:3], b[:3]):
            np.testing.assert_array_equal(v, u)
        np.testing.assert_allclose(b[3], a[3], rtol=1e-12, atol=0)
        assert b[4].path == a[4].path == 0


def test_noalloc_refuses_what_does_not_fit():
    x = synth.rare_396460x15(n=2_000_000, sample_seed=6)
    with pytest.raises(_lib.CostalError, match="bytes of device scratch") as e:
        device_call("costal_knn_exact_device_noalloc", x, 30)
    assert e.value.code == _lib.E_UNSUPPORTED


def test_noalloc_largest_fit_reads_rows_unpadded():
    """The largest row count the static scratch accepts (found by bisection; a refused call
    launches nothing).  There the call has given up, in this order, the padded copy of the rows
    that the rescoring reads 16 bytes at a time (it then reads the input one float at a time),
    the pruned scan (no tile pair skipped) and candidate splits; its output still equals the
    allocating entry point's bit for bit."""
    x_all = synth.rare_396460x15(n=2_000_000, sample_seed=7)
    lo, hi = 60_000, 2_000_000  # lo fits, hi does not (test_noalloc_refuses_what_does_not_fit)
    while hi - lo > 4096:
        mid = (lo + hi) // 2
        try:
            device_call("costal_knn_exact_device_noalloc", np.ascontiguousarray(x_all[:mid]), 10)
            lo = mid
        except _lib.CostalError as e:
            assert e.code == _lib.E_UNSUPPORTED
            hi = mid
    x = np.ascontiguousarray(x_all[:lo])
    print(f"noalloc largest fit n={lo}", flush=True)
    a = device_call("costal_knn_exact_device", x, 10)
    b = device_call("costal_knn_exact_device_noalloc", x, 10)
    assert b[4].blocks_skipped == 0 and a[4].blocks_skipped > 0
    for u, v in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(v, u)
    np.testing.assert_allclose(b[3], a[3], rtol=1e-12, atol=0)
