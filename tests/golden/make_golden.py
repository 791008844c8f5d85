"""Writes the golden kNN fixtures in this directory from the CPU oracle's brute force.

Inputs are the reference's own test generators (ike12th/CosTaL crates/costal-core/src/knn.rs:931-954
and tests/test_knn.py:20-24) at the sizes its tests use; outputs are the brute-force definition
(knn.rs:884-924): idx int32, cos float64, norms float64, plus Tanimoto weights (weight.rs:28-37).
Run from the repository root:  python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from costal_amd import synth  # noqa: E402
from oracle import oracle  # noqa: E402


def cases():
    # knn.rs:960-967 matches_bruteforce_exactly
    yield "lcg_clustered_1500x32_k10", synth.lcg_clustered(1500, 32, 1), 10
    # knn.rs:969-999 output_is_independent_of_blocks...
    yield "lcg_clustered_900x17_k7", synth.lcg_clustered(900, 17, 2), 7
    # knn.rs:1001-1017 exact duplicates (10 prototypes x 30)
    g = synth.Lcg(3)
    protos = g.take(10 * 5) - np.float32(0.3)
    yield "lcg_dups_300x5_k10", np.ascontiguousarray(np.tile(protos.reshape(10, 5), (30, 1))), 10
    # knn.rs:1019-1032 zero rows and integer ties
    g = synth.Lcg(4)
    x = np.floor(g.take(400 * 6) * np.float32(3.0)).reshape(400, 6)
    x[[0, 57, 399]] = 0.0
    yield "lcg_ties_400x6_k12", np.ascontiguousarray(x), 12
    # tests/test_knn.py:38-43 (seed 0) and :87-94
    yield "clustered_3000x32_k10", synth.clustered(3000, 32, 0), 10
    rng = np.random.default_rng(5)
    x = np.floor(rng.uniform(0, 3, size=(500, 6))).astype(np.float32)
    x[[0, 99, 499]] = 0.0
    yield "ties_500x6_k12", x, 12


def main():
    for name, x, k in cases():
        idx, cos, norms = oracle.knn_bruteforce(x, k)
        w = oracle.reweight(idx, cos, norms, "tanimoto")
        np.savez_compressed(os.path.join(HERE, name + ".npz"), x=x, k=k, idx=idx, cos=cos,
                            norms=norms, w=w)
        print(name, x.shape, k)


if __name__ == "__main__":
    main()
