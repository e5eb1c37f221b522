"""Benchmark: exact kNN graph cells/s at 1,000,000 x 50, k = 15 on one MI355X.

One step = one complete kNN build of a device-resident embedding (centring, MFMA
candidate pass, fp32 rerank, exact recomputation of uncertified rows) with outputs left
on the device. The headline `value` is the BF16+rerank path; the plain FP32 kernel is
timed on the same data and reported beside it, with the ratio.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu] [--no-fp32]

Prints ONE JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

BF16_DENSE_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 MFMA
FP32_MATRIX_PEAK_TFLOPS = 157.3  # MI355X fp32 MFMA


def parse_args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n", type=int, default=1_000_000)
    p.add_argument("--d", type=int, default=50)
    p.add_argument("--k", type=int, default=15)
    p.add_argument("--fp32-steps", type=int, default=2)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-fp32", action="store_true")
    p.add_argument("--cpu-rows", type=int, default=32768)
    p.add_argument("--parity-rows", type=int, default=256)
    p.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                   help="only copy a MIB-sized device buffer three times and exit: the "
                        "known-size copy that calibrates the HBM byte counters")
    return p.parse_args()


# HBM bytes per launch of the bf16 pass at the default workload, from the FETCH_SIZE and
# WRITE_SIZE counters (each scaled by the reading of a known-size copy); the summary file
# holds the commands and the raw readings.
TRAFFIC_SUMMARY = ROOT / "profiles" / "pass_kernel_bf16_1Mx50_k15_traffic.md"


def measured_traffic(n: int, d: int, k: int):
    if (n, d, k) != (1_000_000, 50, 15) or not TRAFFIC_SUMMARY.exists():
        return None
    for line in TRAFFIC_SUMMARY.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def copy_probe(mib: int) -> int:
    import torch

    src = torch.ones(mib << 20, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": src.numel(), "launches": 3}), flush=True)
    return 0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def main() -> int:
    args = parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                          "(the path shards by query rows; see the design notes)"}), flush=True)
        return 2
    if args.copy_probe:
        return copy_probe(args.copy_probe)

    import numpy as np
    import torch

    import mrustsc_amd
    from mrustsc_amd import synthetic

    n, d, k = args.n, args.d, args.k
    log(f"[bench] generating {n} x {d} clustered mixture")
    t0 = time.time()
    x = synthetic.clustered(n, d, seed=1)
    log(f"[bench] data ready in {time.time() - t0:.1f}s")

    ctx = mrustsc_amd.Context(0)
    dx = torch.from_numpy(x).to("cuda")
    ids = {p: torch.empty((n, k), dtype=torch.int32, device="cuda") for p in ("bf16", "fp32")}
    dist = {p: torch.empty((n, k), dtype=torch.float32, device="cuda") for p in ("bf16", "fp32")}
    torch.cuda.synchronize()

    def step(precision: str) -> dict:
        return mrustsc_amd.knn_device(dx.data_ptr(), n, d, k, ids[precision].data_ptr(),
                                      dist[precision].data_ptr(), precision=precision,
                                      context=ctx)

    def timed(precision: str, steps: int, warmup: int):
        for i in range(warmup):
            s = step(precision)
            log(f"[bench] {precision} warmup {i}: device {s['device_ms']:.1f} ms "
                f"(pass {s['pass_ms']:.1f}, fallback rows {s['fallback_rows']})")
        torch.cuda.synchronize()
        stats = []
        start = time.perf_counter()
        for i in range(steps):
            stats.append(step(precision))
            log(f"[bench] {precision} step {i}: device {stats[-1]['device_ms']:.1f} ms")
        torch.cuda.synchronize()
        wall = time.perf_counter() - start
        return wall / steps, stats

    sec_bf16, st_bf16 = timed("bf16", args.steps, args.warmup)
    cells_bf16 = n / sec_bf16
    pass_ms = sum(s["pass_ms"] for s in st_bf16) / len(st_bf16)
    flops = 2.0 * n * n * d
    achieved = flops / (pass_ms * 1e-3) / 1e12

    result = {
        "metric": "kNN-graph cells/s at 1M x 50, k=15 (BF16+rerank vs FP32); p99 request latency",
        "value": cells_bf16,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec_bf16 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": f"clustered mixture {n} x {d}, k={k}, euclidean, "
                               "bf16 hi/lo MFMA candidates (32) + fp32 rerank + fp64 ties",
                   "n_cells": n, "n_dims": d, "k": k, "clusters": 20,
                   "duplicate_fraction": 0.005},
        "stages_ms": {key: sum(s[key] for s in st_bf16) / len(st_bf16)
                      for key in ("prepare_ms", "pass_ms", "rerank_ms", "fallback_ms",
                                  "device_ms")},
        "fallback_rows": int(st_bf16[-1]["fallback_rows"]),
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": BF16_DENSE_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / BF16_DENSE_PEAK_TFLOPS,
                     "traffic": measured_traffic(n, d, k),
                     "kernel": "pass_kernel<1, 5> (bf16 operands, 5 chunks of 32)",
                     "algorithmic_flops": flops,
                     "algorithmic_flops_per_unit": "2 d per (query, reference) pair; n^2 pairs"},
    }

    if not args.no_fp32:
        sec_fp32, st_fp32 = timed("fp32", args.fp32_steps, 1)
        pass32 = sum(s["pass_ms"] for s in st_fp32) / len(st_fp32)
        ach32 = flops / (pass32 * 1e-3) / 1e12
        same = bool(torch.equal(ids["bf16"], ids["fp32"]) and torch.equal(dist["bf16"], dist["fp32"]))
        result["fp32"] = {"cells_per_s": n / sec_fp32, "ms_per_step": sec_fp32 * 1e3,
                          "pass_ms": pass32, "achieved_tflops": ach32,
                          "roofline_frac": ach32 / FP32_MATRIX_PEAK_TFLOPS,
                          "fallback_rows": int(st_fp32[-1]["fallback_rows"]),
                          "identical_to_bf16": same}
        result["bf16_over_fp32"] = (n / sec_bf16) / (n / sec_fp32)

    got_ids = ids["bf16"].cpu().numpy().view(np.uint32)
    got_dist = dist["bf16"].cpu().numpy()
    if args.parity_rows > 0 or not args.no_cpu:
        import oracle

        threads = min(16, os.cpu_count() or 1)
        rng = np.random.default_rng(0)
        if args.parity_rows > 0:
            rows = np.sort(rng.choice(n, min(n, args.parity_rows), replace=False)).astype(np.uint32)
            t0 = time.time()
            w_ids, w_dist = oracle.knn(x, k, queries=rows, threads=threads)
            result["parity"] = {
                "rows_checked": int(rows.size),
                "ids_equal": bool(np.array_equal(got_ids[rows], w_ids)),
                "distances_equal": bool(np.array_equal(got_dist[rows], w_dist)),
                "oracle_seconds": time.time() - t0,
            }
        if not args.no_cpu:
            rows = np.arange(min(n, args.cpu_rows), dtype=np.uint32)
            t0 = time.perf_counter()
            oracle.baseline_knn(x, k, queries=rows, threads=threads)
            sec = time.perf_counter() - t0
            result["cpu_baseline"] = {
                "value": rows.size / sec, "unit": "cells/s", "cores": threads, "kind": "port",
                "sample": f"{rows.size} query rows x all {n} references (expansion path of "
                          "neighbors.rs restated in C), scaled: cells/s = rows / seconds"}
    print(json.dumps(result), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
