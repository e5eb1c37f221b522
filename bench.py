"""Benchmark: exact kNN graph cells/s on one MI355X.

One step = one complete kNN build of a device-resident embedding (centring, MFMA
candidate pass, fp32 rerank, exact recomputation of uncertified rows) with outputs left
on the device. The default workload is 1,000,000 x 50, k = 15 (a 20-cluster anisotropic
mixture with 0.5% duplicate rows); its headline `value` is the BF16+rerank path, and the
plain FP32 kernel, the cosine metric and the host-to-host rate are measured beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1|c2|c3|c4|c5]
                    [--n N] [--no-cpu] [--no-fp32] [--no-extra]

Every configuration prints ONE JSON line:
    c1  2,638 x 50 (8 clusters, 20 duplicate rows), full CPU baseline (best of 3)
    c2  100,000 x 50 mixture, full CPU baseline
    c3  1,000,000 x 50 mixture (default): BF16 vs FP32, cosine, distance errors; the
        default run also carries blocks for c1, c2, c4 and c5 (p99 latency) beside it
    c4  64 datasets (n 500-50,000, d 30/50/100, k 15/30) in one graph launch
    c5  serving: Poisson arrivals, n 1,000-200,000, d 50, k 15, three offered loads
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
HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E
FP32_MATRIX_PEAK_TFLOPS = 157.3  # MI355X fp32 MFMA
METRIC = "kNN-graph cells/s at 1M x 50, k=15 (BF16+rerank vs FP32); p99 request latency"

# HBM bytes per launch of the bf16 pass at the default workload, from the FETCH_SIZE and
# WRITE_SIZE counters (each scaled by the reading of a known-size copy); the summary file
# holds the commands and the raw readings.
TRAFFIC_SUMMARY = ROOT / "profiles" / "pass_kernel_bf16_1Mx50_k15_traffic.md"
FP16_TRAFFIC_SUMMARY = ROOT / "profiles" / "pass_local_kernel_fp16_1Mx50_k15_traffic.md"


def parse_args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--config", default="c3", choices=("c1", "c2", "c3", "c4", "c5"))
    p.add_argument("--n", type=int, default=0, help="cells (c1-c3; default per config)")
    p.add_argument("--d", type=int, default=50)
    p.add_argument("--k", type=int, default=15)
    p.add_argument("--fp32-steps", type=int, default=2)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-fp32", action="store_true")
    p.add_argument("--no-cosine", action="store_true")
    p.add_argument("--no-fp16", action="store_true", help="skip the fp16 local-origin block (c3)")
    p.add_argument("--no-large-k", action="store_true", help="skip the k = 100 block (c2)")
    p.add_argument("--no-tsne", action="store_true", help="skip the t-SNE input distances block")
    p.add_argument("--cpu-rows", type=int, default=32768)
    p.add_argument("--parity-rows", type=int, default=256)
    p.add_argument("--requests", type=int, default=400, help="c5: requests per offered load")
    p.add_argument("--slots", type=int, default=4, help="c5: builds the server runs at once")
    p.add_argument("--block-requests", type=int, default=150,
                   help="requests per offered load of the serving block in the default run")
    p.add_argument("--no-extra", action="store_true",
                   help="default run: skip the C1, C2, C4 and C5 blocks beside the headline")
    p.add_argument("--gemm-ceiling", action="store_true",
                   help="time the platform's bf16 GEMM (torch / hipBLASLt) as a practical MFMA ceiling")
    p.add_argument("--hbm-probe", type=int, default=0, metavar="MB",
                   help="only time a device fill and a device copy of an MB-sized f32 buffer "
                        "(HIP events, 10 launches after 3) and exit: the practical HBM rates")
    p.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                   help="only copy a MIB-sized device buffer three times and exit: the "
                        "known-size copy that calibrates the HBM byte counters")
    return p.parse_args()


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_cores() -> int:
    """CPU cores this process may run on (the CPU baseline uses all of them)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


GOLDEN = ROOT / "tests" / "golden"


def fixture_check(name: str, x, ids, dist, index: int | None = None, key: str | None = None) -> dict:
    """The whole graph against a committed oracle fixture (digests of the oracle's ids and
    distances over every row, written by tests/golden/make_config_fixtures.py)."""
    import hashlib

    import numpy as np

    path = GOLDEN / f"{name}_oracle_fixture.json"
    if not path.exists():
        return {"fixture": path.name, "checked": False}
    fix = json.loads(path.read_text())
    if index is not None:
        fix = fix["datasets"][index]
    if key is not None:
        fix = fix["graphs"][key]
    h = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()  # noqa: E731
    return {"fixture": path.name, "checked": True,
            "embedding_equal": h(x) == fix["x_sha256"],
            "ids_equal": h(ids.astype(np.uint32)) == fix["ids_sha256"],
            "distances_equal": h(dist.astype(np.float32)) == fix["dist_sha256"],
            "rows": int(ids.shape[0])}


def profile_figure(summary: Path, key: str):
    """A `key: value` line of a committed profile summary, or None."""
    if not summary.exists():
        return None
    for line in summary.read_text().splitlines():
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def measured_traffic(n: int, d: int, k: int, summary: Path = TRAFFIC_SUMMARY):
    if (n, d, k) != (1_000_000, 50, 15):
        return None
    return profile_figure(summary, "traffic_bytes_per_launch")


# SQ_INSTS_MFMA of the bf16 pass at 1M x 50, k = 15 (committed counter profile): the MFMA work
# the two-stage pass issues depends on how many blocks it refines, so it is measured.
SQ_SUMMARY = ROOT / "profiles" / "pass_kernel_bf16_1Mx50_k15_sq_counters.md"


def executed_bf16_tflops(n: int, d: int, k: int, pass_ms: float):
    if (n, d, k) != (1_000_000, 50, 15):
        return None
    insts = profile_figure(SQ_SUMMARY, "mfma_instructions_per_launch")
    return None if insts is None else insts * 16 * 16 * 32 * 2 / (pass_ms * 1e-3) / 1e12


def gemm_ceiling() -> int:
    """The platform library's bf16 GEMM on this GPU (torch.matmul, hipBLASLt underneath): the
    practical MFMA rate under sustained load, beside the 2.5 PFLOP/s nominal peak. The last
    shape is the pass's contraction written out as a plain GEMM (K = 160 packed features)."""
    import torch

    for m, n, k in ((8192, 8192, 8192), (16384, 16384, 4096), (65536, 4096, 160)):
        a = torch.randn(m, k, dtype=torch.bfloat16, device="cuda")
        b = torch.randn(k, n, dtype=torch.bfloat16, device="cuda")
        for _ in range(3):
            c = a @ b
        torch.cuda.synchronize()
        iters = max(10, int(2.0e15 * 0.5 / (2 * m * n * k)))  # about half a second of work
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(iters):
            c = a @ b
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / iters
        print(json.dumps({"gemm": [m, n, k], "dtype": "bf16", "iters": iters, "ms": ms,
                          "tflops": 2.0 * m * n * k / ms / 1e9}), flush=True)
        del a, b, c
    return 0


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


def hbm_probe(mb: int) -> int:
    """A plain device fill and copy of mb MB: what streaming writes and reads reach on this GPU,
    beside the 8 TB/s nominal peak (the reference point for the write-bound t-SNE tile kernel,
    whose output at 20,000 cells is 1,600 MB)."""
    import torch

    n = mb * 1_000_000 // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda")
    b = torch.empty(n, dtype=torch.float32, device="cuda")
    for name, fn, nbytes in (("fill", lambda: a.fill_(1.0), 4 * n), ("copy", lambda: b.copy_(a), 8 * n)):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 10
        print(json.dumps({"op": name, "mb": mb, "ms": ms, "bytes_moved": nbytes,
                          "tb_per_s": nbytes / ms / 1e9}), flush=True)
    return 0


def base_line(args, value, ms, dtype="bf16"):
    return {"metric": METRIC, "value": value, "unit": "cells/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": dtype, "data": "synthetic"}


class Single:
    """Device-resident single-embedding builds through knn_device."""

    def __init__(self, x, k, ctx):
        import torch

        self.n, self.d = x.shape
        self.k = k
        self.ctx = ctx
        self.x = torch.from_numpy(x).to("cuda")
        self.ids = torch.empty((self.n, k), dtype=torch.int32, device="cuda")
        self.dist = torch.empty((self.n, k), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()

    def step(self, precision="bf16", metric="euclidean"):
        import mrustsc_amd

        return mrustsc_amd.knn_device(self.x.data_ptr(), self.n, self.d, self.k,
                                      self.ids.data_ptr(), self.dist.data_ptr(),
                                      precision=precision, metric=metric, context=self.ctx)

    def timed(self, steps, warmup, precision="bf16", metric="euclidean"):
        """Wall time per step (host clock around synchronised steps) and device stats."""
        import torch

        for i in range(warmup):
            s = self.step(precision, metric)
            log(f"[bench] {precision}/{metric} warmup {i}: device {s['device_ms']:.2f} ms "
                f"(pass {s['pass_ms']:.2f}, fallback rows {s['fallback_rows']})")
        torch.cuda.synchronize()
        stats, walls = [], []
        for i in range(steps):
            t0 = time.perf_counter()
            stats.append(self.step(precision, metric))
            walls.append(time.perf_counter() - t0)
            log(f"[bench] {precision}/{metric} step {i}: device {stats[-1]['device_ms']:.2f} ms")
        torch.cuda.synchronize()
        return walls, stats

    def result(self):
        import numpy as np

        return self.ids.cpu().numpy().view(np.uint32), self.dist.cpu().numpy()


def stage_means(stats):
    return {key: sum(s[key] for s in stats) / len(stats)
            for key in ("prepare_ms", "pass_ms", "rerank_ms", "fallback_ms", "device_ms")}


def parity_block(x, k, got_ids, got_dist, rows, metric="euclidean"):
    """Indices and distances against the oracle on `rows`; distance errors against fp64."""
    import numpy as np

    import oracle

    threads = host_cores()
    t0 = time.time()
    w_ids, w_dist = oracle.knn(x, k, metric=metric, queries=rows, threads=threads)
    f_ids, f_dist = oracle.knn_fp64(x, k, metric=metric, queries=rows, threads=threads)
    err = np.abs(got_dist[rows].astype(np.float64) - f_dist)
    return {"rows_checked": int(rows.size),
            "ids_equal": bool(np.array_equal(got_ids[rows], w_ids)),
            "distances_equal": bool(np.array_equal(got_dist[rows], w_dist)),
            "ids_equal_fp64_truth": float(np.mean(got_ids[rows] == f_ids)),
            "distance_error_vs_fp64": {"max": float(err.max()), "mean": float(err.mean())},
            "oracle_seconds": time.time() - t0}


def connectivities_block(single, ctx, with_oracle: bool) -> dict:
    """The fuzzy simplicial set of the graph just built, device-resident in and out;
    against the oracle (the reference's algorithm restated in C, one thread) when asked."""
    import numpy as np
    import torch

    import mrustsc_amd

    n, k = single.n, single.k
    cap = 2 * n * k
    indptr = torch.empty(n + 1, dtype=torch.int32, device="cuda")
    cols = torch.empty(cap, dtype=torch.int32, device="cuda")
    vals = torch.empty(cap, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    def run():
        return mrustsc_amd.connectivities_device(single.ids.data_ptr(), single.dist.data_ptr(), n, k,
                                                 indptr.data_ptr(), cols.data_ptr(), vals.data_ptr(),
                                                 cap, context=ctx)
    run()
    walls = []
    for _ in range(3):
        t0 = time.perf_counter()
        nnz = run()
        walls.append(time.perf_counter() - t0)
    best = min(walls)
    block = {"ms": best * 1e3, "cells_per_s": n / best, "nnz": nnz,
             "note": "device-resident kNN graph in, CSR out; best of 3 after 1 warm-up"}
    if with_oracle:
        import oracle

        ids, dist = single.result()
        t0 = time.perf_counter()
        w_indptr, w_cols, w_vals, _, _ = oracle.connectivities(ids, dist)
        sec = time.perf_counter() - t0
        g_indptr = indptr.cpu().numpy().view(np.uint32)
        g_cols = cols.cpu().numpy()[:nnz].view(np.uint32)
        g_vals = vals.cpu().numpy()[:nnz]
        same = bool(np.array_equal(g_indptr, w_indptr) and np.array_equal(g_cols, w_cols))
        rel = (np.abs(g_vals.astype(np.float64) - w_vals) / np.maximum(np.abs(w_vals), 1e-30)) if same else None
        block["parity"] = {"pattern_equal": same,
                           "max_relative_error": float(rel.max()) if same else None,
                           "bitwise_fraction": float(np.mean(g_vals == w_vals)) if same else None,
                           "bar": "pattern equal, relative <= 1e-4, >= 90% bit-identical"}
        block["cpu_baseline"] = {"value": n / sec, "unit": "cells/s", "cores": 1, "kind": "port",
                                 "sample": f"all {n} rows (neighbors.rs:193-372 restated in C)"}
    return block


def tsne_block(ctx, with_oracle: bool, n: int = 20000, d: int = 50) -> dict:
    """t-SNE's input distances (tsne.rs:86, :173-188) at the reference's exact-t-SNE cell
    cap: device-resident points in, the n x n float32 matrix out. Bound by the n^2 * 4
    bytes it writes; the rate below is over the whole call (centring, tiles, flag
    readback), a lower bound on the tile kernel's."""
    import numpy as np
    import torch

    import mrustsc_amd
    from mrustsc_amd import synthetic

    x = synthetic.clustered(n, d, seed=2)
    dx = torch.from_numpy(x).to("cuda")
    out = torch.empty((n, n), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    def run():
        mrustsc_amd.tsne_squared_distances_device(dx.data_ptr(), n, d, out.data_ptr(), context=ctx)
    run()
    walls = []
    for _ in range(5):
        t0 = time.perf_counter()
        run()
        walls.append(time.perf_counter() - t0)
    best = min(walls)
    written = n * n * 4.0
    block = {"n_cells": n, "n_dims": d, "call_ms": best * 1e3,
             "achieved_gb_per_s": written / best / 1e9,
             "roofline": {"bound": "hbm", "achieved": written / best / 1e9, "peak": HBM_PEAK_GBPS,
                          "unit": "GB/s", "frac": written / best / 1e9 / HBM_PEAK_GBPS,
                          "algorithmic_bytes": written,
                          "algorithmic_bytes_per_unit": "4 per matrix entry written; n^2 entries"},
             "tflops": 2.0 * n * n * d / best / 1e12,
             "note": "device-resident, best of 5 after 1 warm-up; wall clock of the whole call"}
    if with_oracle:
        import oracle

        threads = host_cores()
        rows = (0, 512)
        t0 = time.perf_counter()
        want = oracle.tsne_squared_distances(x, rows=rows, threads=threads)
        sec = time.perf_counter() - t0
        got = out[rows[0]:rows[1]].cpu().numpy().astype(np.float64)
        c = x.astype(np.float64) - x.astype(np.float64).mean(0)
        norms = (c * c).sum(1)
        lim = (2 * d + 8) * 2.0 ** -24 * (norms[rows[0]:rows[1], None] + norms[None, :]) * 1.0001
        err = np.abs(got - want)
        block["parity"] = {"rows_checked": rows[1] - rows[0], "within_bound": bool(np.all(err <= lim)),
                           "max_error_over_bound": float((err / lim).max()),
                           "bitwise_fraction": float(np.mean(got == want)),
                           "bar": "|device - oracle| <= (2 d + 8) 2^-24 (|a|^2 + |b|^2)"}
        block["cpu_baseline"] = {"value": (rows[1] - rows[0]) * n / sec, "unit": "entries/s",
                                 "cores": threads, "host_cpus": os.cpu_count(), "kind": "port",
                                 "sample": f"rows {rows[0]}-{rows[1]} of the {n} x {n} matrix "
                                           "(tsne.rs:173-188 restated in C)",
                                 "device_entries_per_s": n * n / best}
    del out, dx
    torch.cuda.synchronize()
    return block


def cpu_baseline(x, k, rows=None, repeats=1):
    import numpy as np

    import oracle

    threads = host_cores()
    n = x.shape[0]
    q = np.arange(n if rows is None else min(n, rows), dtype=np.uint32)
    best = None
    for _ in range(repeats):
        t0 = time.perf_counter()
        oracle.baseline_knn(x, k, queries=None if rows is None else q, threads=threads)
        sec = time.perf_counter() - t0
        best = sec if best is None else min(best, sec)
    sample = (f"all {n} rows" if rows is None or q.size == n else
              f"{q.size} query rows x all {n} references, scaled: cells/s = rows / seconds")
    return {"value": q.size / best, "unit": "cells/s", "cores": threads,
            "host_cpus": os.cpu_count(), "kind": "port",
            "sample": sample + (f", best of {repeats}" if repeats > 1 else "") +
            " (expansion path of neighbors.rs restated in C)"}


def run_single_config(args, name, x, workload, cpu_rows, cpu_repeats):
    import numpy as np
    import torch

    import mrustsc_amd

    n, d = x.shape
    k = args.k
    ctx = mrustsc_amd.Context(0)
    single = Single(x, k, ctx)
    walls, st = single.timed(args.steps, args.warmup, "bf16")
    sec = sum(walls) / len(walls)
    dev = stage_means(st)
    # Algorithmic work: 2 d flops per (query, reference) pair over the n^2 pairs. The bf16
    # pass scores only pairs_scored of them (every pair of the padded rows but those of the
    # tile-pair steps its ball bound proves empty); the fp32 and fp16 passes score them all.
    scored = int(st[-1]["pairs_scored"]) or n * n
    brute = 2.0 * n * n * d
    flops = brute
    achieved = flops / (dev["pass_ms"] * 1e-3) / 1e12
    on_scored = 2.0 * d * scored / (dev["pass_ms"] * 1e-3) / 1e12
    out = base_line(args, n / sec, sec * 1e3)
    out["config"] = {"workload": workload, "n_cells": n, "n_dims": d, "k": k}
    out["best_ms"] = min(walls) * 1e3
    out["device_cells_per_s_best"] = n / (min(s["device_ms"] for s in st) * 1e-3)
    out["stages_ms"] = dev
    out["fallback_rows"] = int(st[-1]["fallback_rows"])
    # Device memory of one build: the library's workspace (arena and launch metadata;
    # the embedding and the outputs are the caller's device buffers here).
    out["workspace_bytes"] = int(st[-1]["workspace_bytes"])
    out["persistent_pass"] = bool(st[-1]["persistent"])
    out["roofline"] = {"bound": "mfma", "achieved": achieved, "peak": BF16_DENSE_PEAK_TFLOPS,
                       "unit": "TFLOP/s", "frac": achieved / BF16_DENSE_PEAK_TFLOPS,
                       "traffic": measured_traffic(n, d, k),
                       "kernel": f"pass_kernel<1, {(3 * d + 3 + 31) // 32}, ...> "
                                 "(bf16 two-stage pass, chunks of 32)",
                       "algorithmic_flops": flops,
                       "algorithmic_flops_per_unit": "2 d per (query, reference) pair, n^2 pairs",
                       "pairs_scored": scored,
                       "scored_fraction": scored / (n * n),
                       # The same rate counted on the pairs the pass scored: what the matrix
                       # cores achieve on the work they are given.
                       "achieved_on_scored_pairs": on_scored,
                       "frac_on_scored_pairs": on_scored / BF16_DENSE_PEAK_TFLOPS,
                       # The MFMA work issued (coarse steps for every scored pair, the fine
                       # ones for refined blocks), from the committed SQ_INSTS_MFMA reading.
                       "executed_tflops": executed_bf16_tflops(n, d, k, dev["pass_ms"])}
    got_ids, got_dist = single.result()
    fixture = name if args.n == 0 and args.d == 50 and args.k == 15 else None
    if fixture:
        out["whole_graph"] = fixture_check(fixture, x, got_ids, got_dist)
    out["connectivities"] = connectivities_block(single, ctx, not args.no_cpu)

    if not args.no_fp32:
        walls32, st32 = single.timed(args.fp32_steps, 1, "fp32")
        sec32 = sum(walls32) / len(walls32)
        pass32 = stage_means(st32)["pass_ms"]
        ach32 = brute / (pass32 * 1e-3) / 1e12
        ids32, dist32 = single.result()
        out["fp32"] = {"cells_per_s": n / sec32, "ms_per_step": sec32 * 1e3, "pass_ms": pass32,
                       "achieved_tflops": ach32, "roofline_frac": ach32 / FP32_MATRIX_PEAK_TFLOPS,
                       "fallback_rows": int(st32[-1]["fallback_rows"]),
                       "identical_to_bf16": bool(np.array_equal(ids32, got_ids) and
                                                 np.array_equal(dist32, got_dist))}
        out["bf16_over_fp32"] = (n / sec) / (n / sec32)

    if not args.no_fp16 and name == "c3":
        # The local-origin fp16 candidate pass beside the bf16 headline: the same graph from
        # one fp16 product per feature (64 packed features per pair at d = 50).
        w16, st16 = single.timed(max(2, args.steps // 2), 1, "fp16")
        ids16, dist16 = single.result()
        dev16 = stage_means(st16)
        sec16 = sum(w16) / len(w16)
        ach16 = brute / (dev16["pass_ms"] * 1e-3) / 1e12
        nch16 = (d + 3 + 31) // 32
        nch16 = 4 if nch16 > 2 else nch16
        out["fp16"] = {"cells_per_s": n / sec16, "ms_per_step": sec16 * 1e3,
                       "device_cells_per_s_best": n / (min(x["device_ms"] for x in st16) * 1e-3),
                       "stages_ms": dev16, "fallback_rows": int(st16[-1]["fallback_rows"]),
                       "path": st16[-1]["path"], "candidates": int(st16[-1]["candidates"]),
                       "roofline": {"bound": "mfma", "achieved": ach16, "peak": BF16_DENSE_PEAK_TFLOPS,
                                    "unit": "TFLOP/s", "frac": ach16 / BF16_DENSE_PEAK_TFLOPS,
                                    "traffic": measured_traffic(n, d, k, FP16_TRAFFIC_SUMMARY),
                                    "kernel": f"pass_local_kernel<{nch16}> (fp16 local-origin operands)",
                                    "algorithmic_flops": brute,
                                    "executed_tflops": 2.0 * n * ((n + 511) // 512 * 512) * 32 * nch16
                                                       / (dev16["pass_ms"] * 1e-3) / 1e12},
                       "identical_to_bf16": bool(np.array_equal(ids16, got_ids) and np.array_equal(dist16, got_dist)),
                       "fp16_over_bf16": (n / sec16) / (n / sec)}
        if fixture:
            out["fp16"]["whole_graph"] = fixture_check(fixture, x, ids16, dist16)
        if "fp32" in out:
            out["fp16"]["fp16_over_fp32"] = (n / sec16) / out["fp32"]["cells_per_s"]
        if not args.no_cosine:
            single.step("fp16", "cosine")
            ci16, cd16 = single.result()
            out["fp16"]["cosine_whole_graph"] = fixture_check("c3_cosine", x, ci16, cd16) if fixture else None

    rng = np.random.default_rng(0)
    rows = np.sort(rng.choice(n, min(n, args.parity_rows), replace=False)).astype(np.uint32)
    if args.parity_rows > 0:
        out["parity"] = parity_block(x, k, got_ids, got_dist, rows)

    if not args.no_cosine:
        wc, stc = single.timed(max(1, args.steps // 2), 1, "bf16", "cosine")
        c_ids, c_dist = single.result()
        out["cosine"] = {"cells_per_s": n / (sum(wc) / len(wc)),
                         "ms_per_step": sum(wc) / len(wc) * 1e3,
                         "fallback_rows": int(stc[-1]["fallback_rows"])}
        if fixture == "c3":
            out["cosine"]["whole_graph"] = fixture_check("c3_cosine", x, c_ids, c_dist)
        if args.parity_rows > 0:
            out["cosine"]["parity"] = parity_block(x, k, c_ids, c_dist, rows, "cosine")

    # Host to host: host embedding in, host outputs back (PCIe transfers included).
    h2h = []
    for i in range(3):
        t0 = time.perf_counter()
        mrustsc_amd.knn_hip(x, k, context=ctx)
        h2h.append(time.perf_counter() - t0)
    out["host_to_host"] = {"cells_per_s": n / min(h2h[1:]), "best_ms": min(h2h[1:]) * 1e3,
                           "note": "knn_hip from host memory, best of 2 after 1 warm-up"}
    # The device half of pp.neighbors: graph + connectivities, host embedding in, both obsp
    # matrices out, the graph kept on the device between the steps.
    nb = []
    for i in range(3):
        t0 = time.perf_counter()
        mrustsc_amd.neighbors(x, n_neighbors=k + 1, context=ctx)
        nb.append(time.perf_counter() - t0)
    out["pp_neighbors_host_to_host"] = {"cells_per_s": n / min(nb[1:]), "best_ms": min(nb[1:]) * 1e3,
                                        "note": "mrustsc_amd.neighbors: kNN graph + connectivities, "
                                                "best of 2 after 1 warm-up"}
    if not args.no_tsne and name == "c3":
        out["tsne_squared_distances"] = tsne_block(ctx, not args.no_cpu)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(x, k, cpu_rows, cpu_repeats)
    del single
    if name == "c2" and k == 15 and not args.no_large_k:
        out["k100"] = large_k_block(args, x, ctx, 100)
    torch.cuda.synchronize()
    ctx.close()
    return out


def config_block(args, name, x, workload, ctx, cpu_repeats=1):
    """One single-embedding workload beside the headline: bf16 and fp32 builds of the
    device-resident embedding (best of `args.steps` after warm-ups), the whole graph against
    its oracle fixture, the pass's roofline and the CPU baseline on every row."""
    import numpy as np

    n, d = x.shape
    k = 15
    single = Single(x, k, ctx)
    walls, st = single.timed(args.steps, args.warmup, "bf16")
    best = min(walls)
    dev = stage_means(st)
    flops = 2.0 * n * n * d
    achieved = flops / (dev["pass_ms"] * 1e-3) / 1e12
    ids, dist = single.result()
    block = {"workload": workload, "cells_per_s": n / (sum(walls) / len(walls)),
             "cells_per_s_best": n / best, "ms_per_step": sum(walls) / len(walls) * 1e3,
             "stages_ms": dev, "fallback_rows": int(st[-1]["fallback_rows"]),
             "roofline": {"bound": "mfma", "achieved": achieved, "peak": BF16_DENSE_PEAK_TFLOPS,
                          "unit": "TFLOP/s", "frac": achieved / BF16_DENSE_PEAK_TFLOPS,
                          "kernel": "pass_kernel (bf16)", "algorithmic_flops": flops},
             "whole_graph": fixture_check(name, x, ids, dist)}
    if not args.no_fp32:
        w32, st32 = single.timed(max(2, args.fp32_steps), 1, "fp32")
        ids32, dist32 = single.result()
        pass32 = stage_means(st32)["pass_ms"]
        block["fp32"] = {"cells_per_s": n / (sum(w32) / len(w32)), "pass_ms": pass32,
                         "roofline_frac": flops / (pass32 * 1e-3) / 1e12 / FP32_MATRIX_PEAK_TFLOPS,
                         "identical_to_bf16": bool(np.array_equal(ids32, ids) and np.array_equal(dist32, dist))}
        block["bf16_over_fp32"] = block["cells_per_s"] / block["fp32"]["cells_per_s"]
    if not args.no_cpu:
        block["cpu_baseline"] = cpu_baseline(x, k, None, cpu_repeats)
    del single
    if name == "c2" and not args.no_large_k:
        block["k100"] = large_k_block(args, x, ctx, 100)
        block["k128"] = large_k_block(args, x, ctx, 128)
    return block


def large_k_block(args, x, ctx, k):
    """The same embedding at a large k (the 128- or 256-candidate pass): bf16 builds, the whole
    graph against its oracle fixture."""
    n, d = x.shape
    single = Single(x, k, ctx)
    walls, st = single.timed(max(3, args.steps // 2), 2, "bf16")
    dev = stage_means(st)
    ids, dist = single.result()
    flops = 2.0 * n * n * d
    out = {"k": k, "candidates": int(st[-1]["candidates"]), "cells_per_s": n / (sum(walls) / len(walls)),
           "cells_per_s_best": n / min(walls), "ms_per_step": sum(walls) / len(walls) * 1e3,
           "stages_ms": dev, "fallback_rows": int(st[-1]["fallback_rows"]),
           "roofline_frac": flops / (dev["pass_ms"] * 1e-3) / 1e12 / BF16_DENSE_PEAK_TFLOPS,
           "whole_graph": fixture_check("c2_large_k", x, ids, dist, key=str(k))}
    del single
    return out


def run_batch_config(args):
    """64 datasets of uneven n, d and k in one graph launch; each equals its own run."""
    import numpy as np
    import torch

    import mrustsc_amd
    from mrustsc_amd import synthetic

    specs = synthetic.batch_specs(64, seed=7)
    ctx = mrustsc_amd.Context(0)
    data, outs, entries = [], [], []
    t0 = time.time()
    for sp in specs:
        x = synthetic.batch_dataset(sp)
        dx = torch.from_numpy(x).to("cuda")
        di = torch.empty((sp["n"], sp["k"]), dtype=torch.int32, device="cuda")
        dd = torch.empty((sp["n"], sp["k"]), dtype=torch.float32, device="cuda")
        data.append((x, dx))
        outs.append((di, dd))
        entries.append((dx.data_ptr(), sp["n"], sp["d"], sp["k"], di.data_ptr(), dd.data_ptr()))
    torch.cuda.synchronize()
    log(f"[bench] 64 datasets ready in {time.time() - t0:.1f}s, "
        f"{sum(s['n'] for s in specs)} cells")
    cells = sum(s["n"] for s in specs)
    flops = sum(2.0 * s["n"] * s["n"] * s["d"] for s in specs)

    def time_batch(batch, label):
        for i in range(args.warmup):
            st = batch.run()
            log(f"[bench] {label} batch warmup {i}: {st}")
        walls, devs = [], []
        for i in range(args.steps):
            t0 = time.perf_counter()
            st = batch.run()
            walls.append(time.perf_counter() - t0)
            devs.append(st["device_ms"])
            log(f"[bench] {label} batch step {i}: device {st['device_ms']:.2f} ms")
        return sum(walls) / len(walls), devs

    # The creation shape first, then the batch tuner's choice (it times every shape of the
    # whole graph and keeps the fastest).
    with mrustsc_amd.Batch(entries, context=ctx) as batch:
        untuned_sec, untuned_devs = time_batch(batch, "untuned")
        base_ids = [di.cpu().numpy().view(np.uint32).copy() for di, _ in outs]
        base_dist = [dd.cpu().numpy().copy() for _, dd in outs]
        t0 = time.perf_counter()
        tuned = batch.tune()
        tune_sec = time.perf_counter() - t0
        log(f"[bench] batch tuner: {tuned} in {tune_sec:.1f} s")
        out = {}
        sec, devs = time_batch(batch, "tuned")
        # Each dataset against its own standalone run and the untuned batch.
        equal = equal_untuned = equal_oracle = 0
        for i, (sp, (x, dx), (di, dd), b_ids, b_dist) in enumerate(zip(specs, data, outs, base_ids, base_dist)):
            ids = di.cpu().numpy().view(np.uint32)
            dist = dd.cpu().numpy()
            w_ids, w_dist = mrustsc_amd.knn_hip(x, sp["k"], context=ctx)
            equal += int(np.array_equal(ids, w_ids) and np.array_equal(dist, w_dist))
            equal_untuned += int(np.array_equal(ids, b_ids) and np.array_equal(dist, b_dist))
            fc = fixture_check("c4", x, ids, dist, index=i)
            equal_oracle += int(fc.get("ids_equal", False) and fc.get("distances_equal", False))
        # The same 64 builds as separate device-resident calls, for comparison.
        seq = []
        for _ in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for sp, (x, dx), (di, dd) in zip(specs, data, outs):
                mrustsc_amd.knn_device(dx.data_ptr(), sp["n"], sp["d"], sp["k"], di.data_ptr(),
                                       dd.data_ptr(), context=ctx)
            torch.cuda.synchronize()
            seq.append(time.perf_counter() - t0)
        achieved = flops / (min(devs) * 1e-3) / 1e12
        out = base_line(args, cells / sec, sec * 1e3)
        out["config"] = {"workload": "64 datasets: n log-uniform 500-50,000, d in {30, 50, 100}, "
                                     "k in {15, 30}, 5-30 clusters, one graph launch",
                         "datasets": 64, "cells": cells}
        out["device_ms_best"] = min(devs)
        out["datasets_equal_to_standalone"] = equal
        out["datasets_equal_to_untuned"] = equal_untuned
        out["datasets_equal_to_oracle_fixture"] = equal_oracle
        out["tuning"] = {**tuned, "seconds": tune_sec, "untuned_ms": untuned_sec * 1e3,
                         "untuned_device_ms_best": min(untuned_devs),
                         "tuned_over_untuned": untuned_sec / sec,
                         "note": "batch tuner (longest range per work item x candidate buffer), "
                                 "run once before the timed replays, untimed"}
        out["sequential_calls_ms"] = min(seq) * 1e3
        out["batch_over_sequential"] = min(seq) / sec
        out["roofline"] = {"bound": "mfma", "achieved": achieved, "peak": BF16_DENSE_PEAK_TFLOPS,
                           "unit": "TFLOP/s", "frac": achieved / BF16_DENSE_PEAK_TFLOPS,
                           "traffic": None, "kernel": "whole batch (device time)",
                           "algorithmic_flops": flops}
    if not args.no_cpu:
        import oracle

        threads = host_cores()
        t0 = time.perf_counter()
        for sp, (x, _) in zip(specs, data):
            oracle.baseline_knn(x, sp["k"], threads=threads)
        sec = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cells / sec, "unit": "cells/s", "cores": threads,
                               "host_cpus": os.cpu_count(), "kind": "port",
                               "sample": "all 64 datasets, one after another (expansion path of "
                                         "neighbors.rs restated in C), summed"}
    del data, outs
    ctx.close()
    return out


def run_serving_config(args, requests: int | None = None):
    """Poisson arrivals at three offered loads; latency from arrival to host result."""
    import numpy as np

    import mrustsc_amd
    from mrustsc_amd import synthetic

    rng = np.random.default_rng(11)
    pool = []
    for i in range(24):
        n = int(round(np.exp(rng.uniform(np.log(1000), np.log(200000)))))
        pool.append(synthetic.clustered(n, 50, seed=100 + i, n_clusters=int(rng.integers(5, 31))))
    log(f"[bench] serving pool: {len(pool)} embeddings, n {min(p.shape[0] for p in pool)}"
        f"-{max(p.shape[0] for p in pool)}")
    import torch

    resident = [torch.from_numpy(x).to("cuda") for x in pool]  # embeddings kept on the device
    torch.cuda.synchronize()

    def submit(server, i):
        t = resident[i]
        return server.submit_device(t.data_ptr(), t.shape[0], t.shape[1], 15, keep_alive=t)
    out = base_line(args, 0.0, 0.0)
    with mrustsc_amd.Server(slots=args.slots) as server:
        # Saturation: a closed burst of requests, all submitted at once.
        order = rng.integers(0, len(pool), size=64)
        t0 = time.perf_counter()
        tickets = [submit(server, i) for i in order]
        for t in tickets:
            server.wait(t)
        sat = len(order) / (time.perf_counter() - t0)
        log(f"[bench] saturation {sat:.1f} requests/s")
        loads = []
        for frac in (0.5, 0.8, 1.1):
            rate = frac * sat
            count = requests or args.requests
            gaps = rng.exponential(1.0 / rate, size=count)
            picks = rng.integers(0, len(pool), size=count)
            # Arrivals and the server's timestamps share CLOCK_MONOTONIC (steady_clock).
            start = time.monotonic_ns()
            arrival = start + (np.cumsum(gaps) * 1e9).astype(np.int64)
            tickets = []
            for when, i in zip(arrival, picks):
                delay = (int(when) - time.monotonic_ns()) * 1e-9
                if delay > 0:
                    time.sleep(delay)
                tickets.append((int(when), submit(server, i)))
            lat = []
            for when, t in tickets:
                _, _, info = server.wait(t)
                lat.append((info["done_ns"] - when) * 1e-9)
            span = (time.monotonic_ns() - start) * 1e-9
            lat = np.array(lat) * 1e3
            loads.append({"offered_rps": rate, "fraction_of_saturation": frac,
                          "requests": count, "throughput_rps": count / span,
                          "p50_ms": float(np.percentile(lat, 50)),
                          "p99_ms": float(np.percentile(lat, 99))})
            log(f"[bench] load {frac}: {loads[-1]}")
        stats = server.stats()
    del resident
    mid = loads[1]
    out.update({"metric": METRIC + " (serving)", "unit": "requests/s",
                "value": mid["throughput_rps"], "ms_per_step": mid["p99_ms"],
                "p99_ms": mid["p99_ms"], "p50_ms": mid["p50_ms"],
                "saturation_rps": sat, "loads": loads, "server": stats,
                "config": {"workload": "Poisson arrivals, n log-uniform 1,000-200,000, d=50, "
                                       f"k=15, device-resident embeddings, results to host, {args.slots} slots",
                           "pool": len(pool)}})
    return out


def pca_block(args, n: int = 100_000, g: int = 2000, comps: int = 50, k: int = 15) -> dict:
    """PCA by randomised SVD of a device-resident CSR expression matrix (pca.rs:74-215 on the
    device), then the kNN graph of the embedding it leaves on the device: the producer and the
    consumer of X_pca with nothing crossing PCIe between them."""
    import numpy as np
    import torch

    import mrustsc_amd
    from mrustsc_amd import synthetic

    t0 = time.time()
    ip, ix, v = synthetic.expression_csr(n, g, seed=2)
    log(f"[bench] pca: {n} x {g} expression matrix, {ix.size} entries, ready in {time.time() - t0:.1f}s")
    ctx = mrustsc_amd.Context(0)
    d = [torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else a).cuda() for a in (ip, ix, v)]
    emb = torch.empty((n, comps), dtype=torch.float32, device="cuda")
    ids = torch.empty((n, k), dtype=torch.int32, device="cuda")
    dist = torch.empty((n, k), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    def run_pca():
        return mrustsc_amd.pca_device(d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), n, g, comps,
                                      emb.data_ptr(), context=ctx)
    res = run_pca()
    walls = []
    for _ in range(2):
        t0 = time.perf_counter()
        res = run_pca()
        walls.append(time.perf_counter() - t0)
    knn_walls = []
    for _ in range(3):
        t0 = time.perf_counter()
        st = mrustsc_amd.knn_device(emb.data_ptr(), n, comps, k, ids.data_ptr(), dist.data_ptr(), context=ctx)
        knn_walls.append(time.perf_counter() - t0)
    block = {"workload": f"{n} cells x {g} genes CSR ({ix.size} entries, log1p counts of 12 clusters), "
                         f"{comps} components, then k={k} graph of the embedding, all device-resident",
             "pca_ms": min(walls) * 1e3, "pca_cells_per_s": n / min(walls),
             "explained_variance_ratio_top5": [float(x) for x in res["explained_variance_ratio"][:5]],
             "knn_ms": min(knn_walls[1:]) * 1e3, "knn_fallback_rows": int(st["fallback_rows"]),
             "pca_then_knn_cells_per_s": n / (min(walls) + min(knn_walls[1:])),
             "note": "pca_device best of 2 after 1 warm-up (host clock, includes the host Jacobi "
                     "eigensolves); knn_device best of 2 after 1"}
    if not args.no_cpu:
        import oracle

        m = 20_000
        sip, six, sv = ip[:m + 1], ix[:int(ip[m])], v[:int(ip[m])]
        t0 = time.perf_counter()
        oracle.pca(sip, six, sv, g, comps)
        sec = time.perf_counter() - t0
        block["cpu_baseline"] = {"value": m / sec, "unit": "cells/s", "cores": 1, "host_cpus": os.cpu_count(),
                                 "kind": "port",
                                 "sample": f"the first {m} cells of the same matrix (pca.rs restated in C, f64 "
                                           "algebra, one thread)"}
    del d, emb, ids, dist
    torch.cuda.synchronize()
    ctx.close()
    return block


def extra_blocks(args, out: dict) -> None:
    """The other workloads of the metric beside the 1M headline: C1 and C2 single builds,
    the 64-dataset batch (C4) and the serving loop's latency at three offered loads (C5)."""
    import mrustsc_amd
    from mrustsc_amd import synthetic

    ctx = mrustsc_amd.Context(0)
    t0 = time.time()
    out["c1"] = config_block(args, "c1", synthetic.pbmc_like(),
                             "2638 x 50, 8 clusters (sd 5) + N(0, I), 20 duplicate rows, k=15", ctx,
                             cpu_repeats=3)
    log(f"[bench] c1 block in {time.time() - t0:.1f}s")
    t0 = time.time()
    out["c2"] = config_block(args, "c2", synthetic.clustered(100_000, 50, seed=1),
                             "clustered mixture 100000 x 50, k=15, euclidean", ctx)
    log(f"[bench] c2 block in {time.time() - t0:.1f}s")
    ctx.close()
    t0 = time.time()
    b = run_batch_config(args)
    out["batch_c4"] = {key: b[key] for key in b if key not in (
        "metric", "n_gpus", "steps", "warmup", "higher_is_better", "scaling", "vs_baseline", "data")}
    log(f"[bench] c4 block in {time.time() - t0:.1f}s")
    t0 = time.time()
    c5 = run_serving_config(args, requests=args.block_requests)
    out["serving_c5"] = {key: c5[key] for key in ("saturation_rps", "loads", "server", "config", "p50_ms", "p99_ms")}
    mid = c5["loads"][1]
    out["p99_request_latency_ms"] = {
        "value": mid["p99_ms"], "p50_ms": mid["p50_ms"], "offered_rps": mid["offered_rps"],
        "fraction_of_saturation": mid["fraction_of_saturation"], "throughput_rps": mid["throughput_rps"],
        "note": "serving loop (C5), Poisson arrivals at 0.8 of the measured saturation; "
                "latency from arrival to the result in host memory; all three loads under serving_c5"}
    log(f"[bench] c5 block in {time.time() - t0:.1f}s")
    t0 = time.time()
    out["pca_then_knn"] = pca_block(args)
    log(f"[bench] pca block in {time.time() - t0:.1f}s")


def main() -> int:
    args = parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "scrust_hip drives one GPU; --gpus must be 1 (across GPUs the "
                          "graph would shard by query rows, every GPU holding all references)"}),
              flush=True)
        return 2
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.hbm_probe:
        return hbm_probe(args.hbm_probe)
    if args.gemm_ceiling:
        return gemm_ceiling()

    import torch  # noqa: F401  (loads the HIP runtime the library then shares)

    from mrustsc_amd import synthetic

    if args.config == "c4":
        out = run_batch_config(args)
    elif args.config == "c5":
        out = run_serving_config(args)
    else:
        t0 = time.time()
        if args.config == "c1":
            x = synthetic.pbmc_like()
            workload = "2638 x 50, 8 clusters (sd 5) + N(0, I), 20 duplicate rows, k=15"
            cpu_rows, repeats = None, 3
        else:
            n = args.n or (100_000 if args.config == "c2" else 1_000_000)
            x = synthetic.clustered(n, args.d, seed=1)
            workload = (f"clustered mixture {n} x {args.d}, k={args.k}, euclidean, bf16 hi/lo "
                        "MFMA candidates + fp32 rerank + fp64 ties")
            cpu_rows = None if args.config == "c2" else args.cpu_rows
            repeats = 1
        log(f"[bench] {args.config}: {x.shape[0]} x {x.shape[1]} ready in {time.time() - t0:.1f}s")
        out = run_single_config(args, args.config, x, workload, cpu_rows, repeats)
        out["config"].update({"clusters": 8 if args.config == "c1" else 20,
                              "duplicate_fraction": 0.0076 if args.config == "c1" else 0.005})
        if args.config == "c3" and args.n == 0 and not args.no_extra:
            extra_blocks(args, out)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
