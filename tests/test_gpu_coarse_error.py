"""The measured candidate-score error against the bound the exactness certificate relies on.

Every admitted candidate's list score is compared, inside the rescoring kernels, with its f64
cosine; the call reports the largest difference (``coarse_err_max``) next to the bound it used
(``coarse_eps``: (d+4) 2^-23 for FP32 tiles, the reference's own (ike12th/CosTaL
crates/costal-core/src/knn.rs:183-187); the derived split-FP16 bound for hi/lo tiles).  The bound
carries a factor-2 margin, so the measured error must stay below half of it at every shape.
"""

import numpy as np
import pytest

from costal_amd import _costal as ext
from costal_amd import batched, synth

pytestmark = pytest.mark.gpu

FORMS = ["mfma-f32", "mfma-f16", "mfma-f16x3"]


def check(st, what):
    err, eps = st["coarse_err_max"], st["coarse_eps"]
    print(f"{what}: max |list score - f64 cos| = {err:.3e}, bound {eps:.3e}, ratio {err / eps:.4f}", flush=True)
    assert eps > 0
    assert 0 < err <= 0.5 * eps, what


@pytest.mark.parametrize("matmul", FORMS)
def test_cytometry_shape(matmul):
    x = synth.cytometry_86864x39()
    _, st = ext._knn_exact(x, 10, matmul=matmul)
    check(st, f"86864x39 {matmul}")


@pytest.mark.parametrize("matmul", FORMS)
@pytest.mark.parametrize("k", [10, 30])
def test_rare_shape(matmul, k):
    x = synth.rare_396460x15()
    _, st = ext._knn_exact(x, k, matmul=matmul)
    check(st, f"396460x15 k={k} {matmul}")


@pytest.mark.timeout(600)
@pytest.mark.parametrize("matmul", ["mfma-f32", "mfma-f16x3"])
def test_million_row_shape(matmul):
    x = synth.learned_1000000x40()
    _, st = ext._knn_exact(x, 30, matmul=matmul)
    check(st, f"1000000x40 k=30 {matmul}")


def test_wide_rows_shape():
    x = np.ascontiguousarray(synth.counts_3605x2000(), dtype=np.float32)
    _, st = ext._knn_exact(x, 10)
    assert st["path"] == "wide"
    check(st, "3605x2000 wide rows (fp32 tiles)")


def test_batched_shape():
    xs = synth.batched_64(n_samples=12)
    b = batched.Batch(xs, 15)
    st = b.run("mfma-f16x3")
    err, eps = st.coarse_err_max, st.coarse_eps
    print(f"batched: {err:.3e} of {eps:.3e}", flush=True)
    assert 0 < err <= 0.5 * eps


def subnormal_lo_rows(n=20000, d=64, seed=0):
    """Unit rows whose components all have magnitude near 1/8: hi = fp16(v) lies in [2^-4, 2^-2),
    so lo = fp16(v - hi) is below the smallest normal FP16 (2^-14) for nearly every component."""
    rng = np.random.default_rng(seed)
    protos = rng.choice([-1.0, 1.0], size=(50, d))
    signs = protos[rng.integers(0, 50, size=n)]
    flip = rng.random((n, d)) < 0.05
    signs[flip] *= -1
    x = signs * rng.uniform(0.9, 1.1, size=(n, d)) / np.sqrt(d)
    return np.ascontiguousarray(x, dtype=np.float32)


def split16(x):
    x64 = x.astype(np.float64)
    nrm = np.sqrt(np.sum(x64 * x64, axis=1, keepdims=True))
    v = (x64 / nrm).astype(np.float32)
    hi = v.astype(np.float16)
    lo = (v - hi.astype(np.float32)).astype(np.float16)
    return hi.astype(np.float64), lo.astype(np.float64)


def test_subnormal_lo_halves_are_kept():
    """The split-FP16 bound assumes the matrix core keeps FP16 subnormal inputs.  On rows whose lo
    halves are nearly all subnormal the measured error stays below half the bound, while the same
    neighbor pairs scored with those lo halves flushed to zero (computed here on the CPU) would be
    off by more than half the bound: a flushing matrix core would fail this test."""
    x = subnormal_lo_rows()
    hi, lo = split16(x)
    nz = lo != 0
    sub = nz & (np.abs(lo) < 2.0 ** -14)
    assert sub.sum() > 0.9 * nz.sum()
    (idx, cos, _), st = ext._knn_exact(x, 10, matmul="mfma-f16x3")
    check(st, "20000x64 subnormal lo halves, split fp16")
    # the flushed model on the returned neighbor pairs
    lof = np.where(np.abs(lo) < 2.0 ** -14, 0.0, lo)
    rows = np.repeat(np.arange(len(x)), idx.shape[1])
    cols = idx.ravel()
    flushed = (np.sum(hi[rows] * hi[cols], axis=1) + np.sum(hi[rows] * lof[cols], axis=1)
               + np.sum(lof[rows] * hi[cols], axis=1))
    flush_err = np.abs(flushed - cos.ravel()).max()
    print(f"error if subnormal lo halves were flushed: {flush_err:.3e}", flush=True)
    assert flush_err > 0.5 * st["coarse_eps"]


def test_matrix_core_rounds_to_nearest_and_keeps_subnormals():
    """The split-FP16 bound takes the matrix core's FP32 accumulation as rounded to nearest
    (gamma_m = m 2^-24 / (1 - m 2^-24)) and FP16 subnormal inputs as kept; costal_mfma_probe
    runs MFMAs whose results differ under truncation or flushing."""
    import ctypes as C

    from costal_amd import _lib

    out = (C.c_float * 8)()
    _lib.check(_lib.load().costal_mfma_probe(out))
    r = [float(v) for v in out]
    print("mfma probe:", [v.hex() for v in r], flush=True)
    one = 1.0
    assert r[0] == one + 2.0 ** -23  # fp16 tile: 1 + (2^-24 + 2^-30) rounds up
    assert r[1] in (one + 2.0 ** -23, one)  # three quarter-ulps: rounded once, or per addition
    assert r[2] == one - 2.0 ** -24  # 1 - (2^-24 + 2^-30) rounds to the nearer neighbour
    assert r[3] == r[4] == r[7] == 2.0 ** -24  # subnormal FP16 inputs kept
    assert r[5] == one + 2.0 ** -23 and r[6] == one - 2.0 ** -24  # fp32 tiles alike
