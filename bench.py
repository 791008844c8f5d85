"""Benchmark: exact cosine kNN + Tanimoto weights on one MI355X, in cells/s.

Default workload (the headline): 396,460 cells x 15 features (arcsinh Gaussian mixture, 20
components plus 3 rare ones, 0.5% exact duplicate rows), k = 10; synthetic, seed-pinned
(costal_amd.synth.rare_396460x15).  One step = one complete drop-in call on inputs already
resident in HBM: norms, fused MFMA tiles with per-row candidate lists, f64 rescoring with the
exactness certificate, fallbacks and fused Tanimoto weights, ending with idx/cos/norms/weights in
HBM.

Forms timed side by side: plain FP32 MFMA tiles; split-FP16 MFMA tiles (hi.hi + hi.lo + lo.hi);
split-FP16 tiles with learned BF16 admission thresholds (the headline ``value``; the threshold
model is trained before timing on as many other cells of the same mixture as it serves).  The line reports that
all forms give identical neighbors and weights.  ``roofline`` is for the dominant kernel (the fused
tile + scan kernel), its time measured live with HIP events on the stream it runs on.
``cpu_baseline`` times the CPU oracle's blocked form (a C restatement of the reference's run_blocked,
OpenMP) on a bounded sample of query rows of the same input, on this host's cores.

Other workloads (--workload): cytometry_86864x39, learned_1000000x40 (k=30), batched_64 (64 uneven
samples in one batched hipGraph call, k=15), counts_3605x2000 (wide rows: the exact row-scan path).

    python bench.py [--gpus 1] [--steps 10] [--warmup 3] [--workload NAME] [--no-cpu] [--no-tanimoto]
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_TFLOPS = {"fp32": 157.3, "fp16": 2500.0, "fp16x3": 2500.0}  # MI355X dense MFMA peaks
METRIC = "cells/s exact k=10 cosine kNN+Tanimoto; mixed-prec vs FP32 GPU at 100% match"
FORM_DESC = {
    "fp32": "fp32 MFMA candidates + f64 rescoring",
    "fp16": "fp16 MFMA candidates + f32 re-score + f64 rescoring",
    "fp16x3": "split-fp16 MFMA candidates (hi.hi + hi.lo + lo.hi) + f64 rescoring",
    "fp16x3+learned": "split-fp16 MFMA candidates + learned BF16 admission thresholds + f64 rescoring",
    "tanimoto-select": "neighbors selected by Tanimoto: split-fp16 MFMA candidates + f64 rescoring",
}
PRECISION_OF = {"fp32": "fp32", "fp16": "fp16", "fp16x3": "fp16x3", "fp16x3+learned": "fp16x3",
                "tanimoto-select": "fp16x3"}
# HBM bytes per launch of the dominant kernel: FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (KiB), the scale
# factors from a 1 GiB device copy under the same counters; rocprofv3 --pmc runs of this script,
# mean over the first passes of the run (profiles/, one summary per kernel and workload).
MEASURED_TRAFFIC = {
    ("rare_396460x15", "fp16x3+learned"): 8461.6e6,
    ("learned_1000000x40", "fp16x3+learned"): 95774.3e6,
    ("cytometry_86864x39", "fp16x3"): 1847.9e6,
    ("counts_3605x2000", "fp32"): 1015.0e6 + 57.3e6,
    ("batched_64", "fp16x3"): 2839.3e6,
    ("rare_396460x15", "tanimoto-select"): 10571.8e6,
    ("cytometry_86864x39", "tanimoto-select"): 1859.4e6,
}
WORKLOADS = {
    # name: (k, forms, train the threshold model on this many other cells)
    "rare_396460x15": (10, "fp32,fp16x3,fp16x3+learned", 396_460),
    "cytometry_86864x39": (10, "fp32,fp16,fp16x3", 0),
    "learned_1000000x40": (30, "fp32,fp16x3,fp16x3+learned", 200_000),
    "batched_64": (15, "fp32,fp16x3", 0),
    "counts_3605x2000": (10, "fp32", 0),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="rare_396460x15", choices=sorted(WORKLOADS))
    ap.add_argument("--k", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tanimoto", action="store_true", help="skip the Tanimoto-selection timing")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline duration")
    ap.add_argument("--forms", default=None, help="forms to time (headline = last)")
    ap.add_argument("--train-cells", type=int, default=None, help="threshold-model training cells")
    ap.add_argument("--quantile", type=float, default=0.995,
                    help="held-out rows whose learned threshold must stay below their k-th similarity")
    ap.add_argument("--model-cache", default=None,
                    help="threshold model file: loaded when it exists, else trained and written there")
    return ap.parse_args()


def load_workload(name: str, train_cells: int):
    from costal_amd import synth

    if name == "rare_396460x15":
        return synth.rare_396460x15(), (lambda: synth.rare_396460x15(n=train_cells, sample_seed=1))
    if name == "cytometry_86864x39":
        return synth.cytometry_86864x39(), None
    if name == "learned_1000000x40":
        return synth.learned_1000000x40(), (lambda: synth.learned_train_x40(n=train_cells))
    if name == "counts_3605x2000":
        return np.ascontiguousarray(synth.counts_3605x2000(), dtype=np.float32), None
    if name == "batched_64":
        return synth.batched_64(), None
    raise ValueError(name)


class DeviceRun:
    """Device buffers (torch allocations) and one call of costal_knn_exact_device."""

    def __init__(self, x: np.ndarray, k: int):
        import torch

        from costal_amd import _lib

        self.torch = torch
        self.L = _lib.load()
        self.lib = _lib
        self.n, self.d = x.shape
        self.k = k
        dev = torch.device("cuda:0")
        self.x = torch.from_numpy(x).to(dev)
        self.idx = torch.empty((self.n, k), dtype=torch.int32, device=dev)
        self.cos = torch.empty((self.n, k), dtype=torch.float64, device=dev)
        self.norms = torch.empty(self.n, dtype=torch.float64, device=dev)
        self.w = torch.empty((self.n, k), dtype=torch.float64, device=dev)
        self.stream = torch.cuda.current_stream().cuda_stream
        self.model = None
        torch.cuda.synchronize()

    def set_model(self, model):
        self.model = model
        self.model_params = self.torch.from_numpy(model.params).cuda()

    def step(self, form: str, threshold=None, prune=True):
        p = self.lib.KnnParams()
        p.k = self.k
        p.margin = -1
        p.select = 1 if form == "tanimoto-select" else 0
        p.no_prune = 0 if prune else 1
        p.precision = self.lib.PRECISION[PRECISION_OF[form]]
        if threshold is not None:
            p.row_threshold = threshold.data_ptr()
        if form.endswith("+learned"):
            p.mlp_params = self.model_params.data_ptr()
            p.mlp_offset = float(self.model.offset)
            if self.model.offset2 is not None:
                p.mlp_offset2 = float(self.model.offset2)
            p.mlp_bf16 = 1
        st = self.lib.KnnStats()
        rc = self.L.costal_knn_exact_device(
            C.c_void_p(self.x.data_ptr()), self.n, self.d, C.byref(p), C.c_void_p(self.idx.data_ptr()),
            C.c_void_p(self.cos.data_ptr()), C.c_void_p(self.norms.data_ptr()),
            C.c_void_p(self.w.data_ptr()), C.byref(st), C.c_void_p(self.stream))
        self.lib.check(rc)
        return st

    def outputs(self):
        return (self.idx.cpu().numpy(), self.cos.cpu().numpy(), self.norms.cpu().numpy(),
                self.w.cpu().numpy())


class BatchRun:
    """The 64-sample batch: one costal_knn_batched_device call per step (a hipGraph replay)."""

    def __init__(self, xs, k: int, tuner_plan=None):
        from costal_amd import batched

        margins = col_splits = None
        if tuner_plan:
            _, margins, col_splits = tuner_plan
        self.b = batched.Batch(xs, k, margins=margins, col_splits=col_splits)
        self.torch = self.b.torch

    def step(self, form: str, threshold=None):
        return self.b.run({"fp32": "mfma-f32", "fp16": "mfma-f16", "fp16x3": "mfma-f16x3"}[form])

    def outputs(self):
        outs = self.b.outputs()
        return tuple(np.concatenate([o[c].ravel() for o in outs]) for c in range(4))


def time_form(run, form: str, steps: int, warmup: int):
    torch = run.torch
    for _ in range(warmup):
        run.step(form)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = [run.step(form) for _ in range(steps)]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1000.0 / steps
    kern = {
        "prep": float(np.mean([s.prep_ms for s in stats])),
        "order": float(np.mean([s.order_ms for s in stats])),
        "coarse": float(np.mean([s.coarse_ms for s in stats])),
        "rescore": float(np.mean([s.rescore_ms for s in stats])),
        "fallback": float(np.mean([s.fallback_ms for s in stats])),
    }
    return ms, kern, stats[-1]


def host_to_host(x: np.ndarray, k: int, form: str, model, reps: int = 3):
    """The drop-in call as the reference's callers make it (bench/stage_a.py:157-166): a host
    float32 array in, host idx/cos/norms and the Tanimoto weights out, through costal_knn_exact
    (PCIe transfers and output allocation included); warm, best of ``reps``."""
    from costal_amd import _costal

    kw = {"matmul": {"fp32": "mfma-f32", "fp16": "mfma-f16"}.get(PRECISION_OF[form], "mfma-f16x3"), "weights": True}
    if form == "tanimoto-select":
        kw["select"] = "tanimoto"
    if form.endswith("+learned"):
        kw["threshold_model"] = model
    _costal._knn_exact(x, k, **kw)
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        _costal._knn_exact(x, k, **kw)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return {"form": form, "ms": best * 1000.0, "cells_per_s": len(x) / best,
            "bytes_in": int(x.nbytes), "bytes_out": int(len(x) * (k * 20 + 8)),
            "note": "costal_knn_exact from pageable host memory, outputs allocated per call, best of 3"}


def cpu_baseline(x: np.ndarray, k: int, target_s: float):
    """The oracle's blocked form on a bounded sample of query rows (all candidates)."""
    from oracle import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    n = len(x)
    # calibrate the per-row cost from two probes (the difference removes the fixed per-call cost)
    times = []
    probes = (min(n, 2000), min(n, 10000))
    for r in probes:
        t0 = time.perf_counter()
        oracle.knn_blocked(x, k, sort_by_norm=False, rows=(0, r), n_threads=threads)
        times.append(time.perf_counter() - t0)
    per_row = max((times[1] - times[0]) / max(probes[1] - probes[0], 1), 1e-7)
    rows = int(min(n, max(probes[1], target_s / per_row)))
    t0 = time.perf_counter()
    (idx, cos, norms), _ = oracle.knn_blocked(x, k, sort_by_norm=False, rows=(0, rows), n_threads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": rows / dt,
        "unit": "cells/s",
        "cores": threads,
        "kind": "port",
        "sample": f"query rows 0..{rows} of the same {n}x{x.shape[1]} input against all {n} rows, "
                  f"k={k}, {dt:.1f} s wall",
    }, (rows, idx[:rows], cos[:rows])


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this build "
              "(the path shards by rows; see the design notes)", file=sys.stderr)
        sys.exit(2)
    from costal_amd import _lib

    wk, wforms, wtrain = WORKLOADS[args.workload]
    k = args.k or wk
    forms = (args.forms or wforms).split(",")
    train_cells = args.train_cells or wtrain
    data, train_data = load_workload(args.workload, train_cells)
    batched = args.workload == "batched_64"
    tuner_info = None
    if batched:
        from costal_amd import tuner

        t0 = time.perf_counter()
        cache = os.path.join(HERE, "gpurun_out", "tune.json") if os.path.isdir(os.path.join(HERE, "gpurun_out")) else None
        tn = tuner.Tuner(cache)
        plan = tn.plan_batch([x.shape for x in data], k)
        tuner_info = {"precision": plan[0], "seconds": time.perf_counter() - t0, "buckets": len(tn.table)}
        run = BatchRun(data, k, plan)
        n = sum(len(x) for x in data)
        d = max(x.shape[1] for x in data)
        flops = sum(2.0 * len(x) * len(x) * x.shape[1] for x in data)
    else:
        n, d = data.shape
        run = DeviceRun(data, k)
        flops = 2.0 * n * n * d
    model_info = None
    if any(f.endswith("+learned") for f in forms):
        from costal_amd import learned

        t0 = time.perf_counter()
        if args.model_cache and os.path.exists(args.model_cache):
            model = learned.ThresholdModel.load(args.model_cache)
            model.meta.update({"k_label": None, "holdout_rmse": None, "quantile": None, "loaded_from": args.model_cache})
        else:
            model = learned.train(train_data(), k, n_target=n, seed=0, quantile=args.quantile)
            if args.model_cache:
                model.save(args.model_cache)
        run.set_model(model)
        model_info = {"train_cells": train_cells, "k_label": model.meta["k_label"],
                      "holdout_rmse": model.meta["holdout_rmse"], "offset": model.offset,
                      "quantile": model.meta["quantile"], "train_seconds": time.perf_counter() - t0,
                      "arch": f"{d}->64->64->1 relu, fp32 weights, bf16 inference"}
    results, outs = {}, {}
    for form in forms:
        ms, kern, st = time_form(run, form, args.steps, args.warmup)
        results[form] = {"ms_per_step": ms, "cells_per_s": n / (ms / 1000.0), "kernel_ms": kern,
                         "fallback_rows": int(st.fallback_rows), "fallback_rate": int(st.fallback_rows) / n,
                         "short_rows": int(st.short_rows),
                         "coarse_err_max": float(st.coarse_err_max), "coarse_eps": float(st.coarse_eps),
                         "col_splits": int(st.col_splits), "candidates": int(st.candidates),
                         "inserts": int(st.inserts), "rowscan_rows": int(st.rowscan_rows), "path": int(st.path),
                         "tile_pairs": int(st.blocks_total), "tile_pairs_skipped": int(st.blocks_skipped)}
        outs[form] = run.outputs()
    head = forms[-1]
    ref = outs[forms[0]]
    match = {}
    for form in forms[1:]:
        o = outs[form]
        match[f"{form}_vs_{forms[0]}"] = {
            "idx_identical": bool(np.array_equal(o[0], ref[0])),
            "cos_identical": bool(np.array_equal(o[1], ref[1])),
            "w_max_rel": float(np.max(np.abs(o[3] - ref[3]) / np.maximum(np.abs(ref[3]), 1e-300))),
        }

    # Roofline of the dominant kernel.  Algorithmic flops: 2*d per (query, candidate) pair the
    # kernel scores.  The dense algorithm scores all N*N ordered pairs (2*N*N*d, as the reference
    # computes them); the pruned scan scores only the (query block, candidate tile) pairs it does
    # not skip, so its flops are the scanned fraction of that (the dense-equivalent rate is
    # reported beside it).
    hk = results[head]["kernel_ms"]
    path = results[head]["path"]
    if batched:
        kern_name = "batched graph (prep+scan+rescore)"
    else:
        kern_name = {0: "coarse_kernel", 1: "rowscan_kernel", 2: "wide_gemm_kernel + wide_scan_kernel"}[path]
    kern_ms = hk["coarse"] if path != 1 else hk["fallback"]
    pairs, skipped = results[head]["tile_pairs"], results[head]["tile_pairs_skipped"]
    scanned = 1.0 - skipped / pairs if (path == 0 and not batched and pairs > 0) else 1.0
    exec_flops = flops * scanned
    achieved = exec_flops / (kern_ms / 1000.0) / 1e12 if kern_ms > 0 else 0.0
    peak = PEAK_TFLOPS[PRECISION_OF[head]]
    roofline = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": MEASURED_TRAFFIC.get((args.workload, head)), "kernel": kern_name,
                "flops_per_launch": exec_flops, "avg_launch_ms": kern_ms, "scanned_fraction": scanned,
                "dense_flops_per_launch": flops,
                "dense_equivalent_tflops": flops / (kern_ms / 1000.0) / 1e12 if kern_ms > 0 else 0.0}

    h2h = None
    if not batched:
        h2h = host_to_host(data, k, head, run.model)

    # Selection by Tanimoto (select="tanimoto") on the same input: its own kNN, timed like the
    # forms above, and checked equal to the unpruned scan of the same selection.
    tani = None
    if not batched and not args.no_tanimoto:
        ms, kern, st = time_form(run, "tanimoto-select", args.steps, args.warmup)
        got = run.outputs()
        run.step("tanimoto-select", prune=False)
        want = run.outputs()
        tani = {"ms_per_step": ms, "cells_per_s": n / (ms / 1000.0), "kernel_ms": kern,
                "fallback_rows": int(st.fallback_rows), "coarse_err_max": float(st.coarse_err_max),
                "coarse_eps": float(st.coarse_eps), "tile_pairs": int(st.blocks_total),
                "tile_pairs_skipped": int(st.blocks_skipped),
                "equals_unpruned": bool(np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]))}

    cpu = None
    if not args.no_cpu and not batched:
        cpu, (rows, cidx, ccos) = cpu_baseline(data, k, args.cpu_seconds)
        gidx, gcos = outs[head][0], outs[head][1]
        match["gpu_vs_cpu_sample"] = {"rows": rows,
                                      "idx_identical": bool(np.array_equal(gidx[:rows], cidx)),
                                      "cos_identical": bool(np.array_equal(gcos[:rows], ccos))}

    line = {
        "metric": METRIC if args.workload == "rare_396460x15" else f"cells/s exact k={k} cosine kNN+Tanimoto",
        "value": results[head]["cells_per_s"],
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": results[head]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if PRECISION_OF[head] == "fp32" else "fp16",
        "data": "synthetic",
        "config": {"workload": args.workload, "cells": n, "features": d, "k": k,
                   "form": FORM_DESC[head], "parallelism": "single GPU"},
        "forms": results,
        "match": match,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "host_to_host": h2h,
        "tanimoto_selection": tani,
        "threshold_model": model_info,
        "tuner": tuner_info,
        "device": _lib.device_info(),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
