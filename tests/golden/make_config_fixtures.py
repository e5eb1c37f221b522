"""Writes the whole-graph oracle fixtures of the benchmark workloads (k = 15 unless stated):

  c2_oracle_fixture.json         100,000 x 50, Euclidean: mrustsc_amd.synthetic.clustered(100000, 50, seed=1)
  c3_oracle_fixture.json         1,000,000 x 50, Euclidean: synthetic.clustered(1000000, 50, seed=1)
  c3_cosine_oracle_fixture.json  the same embedding, cosine metric
  c4_oracle_fixture.json         the 64 datasets of synthetic.batch_specs(64, seed=7), each
                                 at its own k
  c2_large_k_oracle_fixture.json the C2 embedding at k = 64, 100, 111 and 128 (the 128-candidate
                                 pass and, at k = 128, the exact kernel)

Each graph is the oracle's (oracle.knn_fast, the vectorised oracle_knn with the fp64 tie rule)
and is stored as SHA-256 digests of the raw arrays (ids uint32, distances float32), digests of
every 65,536-row chunk (so a mismatch names its rows), and a few rows in full (distances as
float32 bit patterns). The input's digest is stored too: a test first checks that the
generator reproduced the embedding. About 25 minutes on 8 cores (each 1M graph ~10 min).

    python tests/golden/make_config_fixtures.py [c2 c3 c3_cosine c4 c2_large_k]
"""

from __future__ import annotations

import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from mrustsc_amd import synthetic  # noqa: E402

CHUNK = 65536
HERE = Path(__file__).parent


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def graph_record(x: np.ndarray, k: int, metric: str, n_rows: int) -> dict:
    ids, dist = oracle.knn_fast(x, k, metric=metric)
    n = x.shape[0]
    rows = sorted(set(np.linspace(0, n - 1, n_rows).astype(int).tolist()))
    return {
        "shape": list(x.shape),
        "k": k,
        "metric": metric,
        "x_sha256": digest(x.astype(np.float32)),
        "ids_sha256": digest(ids.astype(np.uint32)),
        "dist_sha256": digest(dist.astype(np.float32)),
        "chunk_rows": CHUNK,
        "chunks": [digest(ids[i:i + CHUNK]) + ":" + digest(dist[i:i + CHUNK]) for i in range(0, n, CHUNK)],
        "zero_distances": int((dist == 0).sum()),
        "rows": {str(r): {"ids": ids[r].astype(int).tolist(),
                          "dist_bits": [f"{b:08x}" for b in dist[r].astype(np.float32).view(np.uint32)]}
                 for r in rows},
    }


def single(name: str, n: int, metric: str) -> None:
    t0 = time.time()
    x = synthetic.clustered(n, 50, seed=1)
    rec = graph_record(x, 15, metric, 16)
    rec["_source"] = (f"oracle k=15 {metric} graph of mrustsc_amd.synthetic.clustered({n}, 50, seed=1); "
                      "written by tests/golden/make_config_fixtures.py")
    (HERE / f"{name}_oracle_fixture.json").write_text(json.dumps(rec, indent=1) + "\n")
    print(f"{name}: {time.time() - t0:.0f} s", flush=True)


def batch() -> None:
    t0 = time.time()
    out = {"_source": "oracle graphs of the 64 datasets of mrustsc_amd.synthetic.batch_specs(64, seed=7), "
                      "each at its own k; written by tests/golden/make_config_fixtures.py",
           "datasets": []}
    for spec in synthetic.batch_specs(64, seed=7):
        x = synthetic.batch_dataset(spec)
        rec = graph_record(x, spec["k"], "euclidean", 4)
        rec["spec"] = spec
        out["datasets"].append(rec)
    (HERE / "c4_oracle_fixture.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"c4: {time.time() - t0:.0f} s", flush=True)


LARGE_K = (64, 100, 111, 128)


def large_k() -> None:
    t0 = time.time()
    x = synthetic.clustered(100_000, 50, seed=1)
    out = {"_source": "oracle Euclidean graphs of mrustsc_amd.synthetic.clustered(100000, 50, seed=1) at "
                      "k = 64, 100, 111, 128; written by tests/golden/make_config_fixtures.py",
           "graphs": {str(k): graph_record(x, k, "euclidean", 8) for k in LARGE_K}}
    (HERE / "c2_large_k_oracle_fixture.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"c2_large_k: {time.time() - t0:.0f} s", flush=True)


def main(argv: list[str]) -> None:
    todo = argv or ["c2", "c4", "c3", "c3_cosine"]
    for name in todo:
        if name == "c2":
            single("c2", 100_000, "euclidean")
        elif name == "c3":
            single("c3", 1_000_000, "euclidean")
        elif name == "c3_cosine":
            single("c3_cosine", 1_000_000, "cosine")
        elif name == "c4":
            batch()
        elif name == "c2_large_k":
            large_k()
        else:
            raise SystemExit(f"unknown fixture {name}")


if __name__ == "__main__":
    main(sys.argv[1:])
