"""Deterministic synthetic inputs for the five benchmark workloads and the reference's test data.

Every workload is fully determined by its seed; the same bytes feed the GPU path, the CPU
baseline and the parity tests.

Workloads (cells x features):

- ``counts_3605x2000``: 12-component Gaussian mixture, component means N(0, 2^2) per feature,
  per-cell noise sd 1, then ``max(0, x - 1.1726)``, float64. The offset puts ~70% of entries at
  zero like log-normalized counts (without it the literal recipe gives ~50%). k = 10.
- ``cytometry_86864x39``: 24 components, weights Dirichlet(1), raw means N(0, 20^2), within sd 5,
  ``arcsinh(x/5)``, float32. k = 10.
- ``rare_396460x15``: 20 components (Dirichlet(1)) plus 3 rare components of 0.05% each, arcsinh,
  float32, then 0.5% of rows overwritten by exact copies of other rows. k = 10 and 30.
- ``batched_64``: 64 samples, N ~ floor(exp(U(ln 2000, ln 60000))), d ~ U{32..48}, 10-30
  components each, arcsinh, float32. k = 15.
- ``learned_1000000x40``: 40 components, arcsinh, float32, k = 30; the threshold model trains on
  cells drawn from the same mixture with another sampling seed (``learned_train_x40``).

``clustered`` reproduces the data generator of the reference's Python tests
(ike12th/CosTaL tests/test_knn.py:20-24) and ``lcg_clustered`` / ``lcg_uniform`` the Rust unit-test
generator (crates/costal-core/src/knn.rs:931-954), so the reference's own test cases can be rerun
bit for bit.
"""

from __future__ import annotations

import numpy as np

COUNTS_OFFSET = 1.1726  # Phi(o / sqrt(5)) = 0.70


def _mixture(rng, n, d, n_comp, *, mean_sd=20.0, within_sd=5.0, weights=None, counts=None,
             sample_rng=None):
    """Mixture parameters from ``rng``; cells from ``sample_rng`` (default: the same stream)."""
    means = rng.normal(0.0, mean_sd, size=(n_comp, d))
    srng = rng if sample_rng is None else sample_rng
    if counts is None:
        if weights is None:
            weights = rng.dirichlet(np.ones(n_comp))
        comp = srng.choice(n_comp, size=n, p=weights)
    else:
        comp = np.repeat(np.arange(n_comp), counts)
        srng.shuffle(comp)
    x = means[comp] + srng.normal(0.0, within_sd, size=(n, d))
    return x, comp


def arcsinh(x, cofactor=5.0):
    return np.ascontiguousarray(np.arcsinh(x / cofactor), dtype=np.float32)


def counts_3605x2000(seed: int = 0, n: int = 3605, d: int = 2000) -> np.ndarray:
    rng = np.random.default_rng(seed)
    means = rng.normal(0.0, 2.0, size=(12, d))
    comp = rng.integers(0, 12, size=n)
    x = means[comp] + rng.normal(0.0, 1.0, size=(n, d))
    return np.maximum(0.0, x - COUNTS_OFFSET)


def cytometry_86864x39(seed: int = 1, n: int = 86_864, d: int = 39) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x, _ = _mixture(rng, n, d, 24)
    return arcsinh(x)


def rare_396460x15(seed: int = 2, n: int = 396_460, d: int = 15, dup_frac: float = 0.005,
                   sample_seed: int = 0) -> np.ndarray:
    """The mixture (weights, means) comes from ``seed``; the cells, duplicate rows included, from
    ``(seed, sample_seed)``, so another ``sample_seed`` draws other cells from the same mixture
    (the threshold model's training data)."""
    rng = np.random.default_rng(seed)
    srng = np.random.default_rng([seed, sample_seed])
    n_rare = int(round(0.0005 * n))
    n_main = n - 3 * n_rare
    w = rng.dirichlet(np.ones(20))
    main = srng.multinomial(n_main, w)
    counts = np.concatenate([main, [n_rare] * 3])
    x, _ = _mixture(rng, n, d, 23, counts=counts, sample_rng=srng)
    x = arcsinh(x)
    n_dup = int(round(dup_frac * n))
    perm = srng.permutation(n)
    dst, src = perm[:n_dup], perm[n_dup:2 * n_dup]
    x[dst] = x[src]
    return x


def batched_64(seed: int = 3, n_samples: int = 64, n_lo: int = 2000, n_hi: int = 60_000,
               d_lo: int = 32, d_hi: int = 48) -> list[np.ndarray]:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_samples):
        n = int(np.floor(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
        d = int(rng.integers(d_lo, d_hi + 1))
        c = int(rng.integers(10, 31))
        x, _ = _mixture(rng, n, d, c)
        out.append(arcsinh(x))
    return out


def learned_mixture(seed: int = 4, d: int = 40, n_comp: int = 40):
    """Mixture parameters (means, weights) of the 10^6 x 40 workload."""
    rng = np.random.default_rng(seed)
    means = rng.normal(0.0, 20.0, size=(n_comp, d))
    weights = rng.dirichlet(np.ones(n_comp))
    return means, weights


def learned_draw(n: int, sample_seed: int, mix_seed: int = 4, d: int = 40) -> np.ndarray:
    means, weights = learned_mixture(mix_seed, d)
    rng = np.random.default_rng(sample_seed)
    comp = rng.choice(len(weights), size=n, p=weights)
    x = means[comp] + rng.normal(0.0, 5.0, size=(n, d))
    return arcsinh(x)


def learned_1000000x40(n: int = 1_000_000) -> np.ndarray:
    return learned_draw(n, sample_seed=40)


def learned_train_x40(n: int = 1_000_000) -> np.ndarray:
    return learned_draw(n, sample_seed=41)


# ----------------------------------------------------------------- the reference's test generators


def clustered(n, d, seed, n_centers=8, noise=0.5) -> np.ndarray:
    """tests/test_knn.py:20-24 of the reference."""
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0, 4, size=(n_centers, d))
    x = centers[rng.integers(0, n_centers, size=n)] + rng.normal(0, noise, size=(n, d))
    return np.ascontiguousarray(x, dtype=np.float32)


class Lcg:
    """crates/costal-core/src/knn.rs:931-940: 64-bit LCG, 24 high bits -> f32 in [0, 1)."""

    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next_f32(self) -> np.float32:
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        return np.float32(self.s >> 40) / np.float32(1 << 24)

    def take(self, m: int) -> np.ndarray:
        return np.array([self.next_f32() for _ in range(m)], dtype=np.float32)


def lcg_clustered(n: int, d: int, seed: int) -> np.ndarray:
    """crates/costal-core/src/knn.rs:943-954 (CyTOF-like, signed)."""
    g = Lcg(seed)
    centers = g.take(8 * d) * np.float32(4.0)
    x = np.empty((n, d), np.float32)
    for i in range(n):
        c = i % 8
        x[i] = centers[c * d:(c + 1) * d] + g.take(d) - np.float32(0.55)
    return x


WORKLOADS = {
    "counts_3605x2000": (counts_3605x2000, 10),
    "cytometry_86864x39": (cytometry_86864x39, 10),
    "rare_396460x15": (rare_396460x15, 10),
    "learned_1000000x40": (learned_1000000x40, 30),
}
