"""The benchmark workloads at their full sizes against the oracle, whole graphs.

Fixtures (tests/golden/*_oracle_fixture.json, written by tests/golden/make_config_fixtures.py
with the vectorised oracle) hold SHA-256 digests of the oracle's ids and distances, digests of
every 65,536-row chunk and a few rows in full. Each test first checks that the generator
reproduced the workload's embedding (its digest), then that the HIP path's whole graph equals
the oracle's bit for bit:

  C2  100,000 x 50, k = 15: the bf16 and the fp32 pass; k = 64, 100 and 111 on the
      128-candidate pass (bf16 and fp32), k = 128 on the exact kernel
  C3  1,000,000 x 50, k = 15: bf16 Euclidean and cosine, fp32 Euclidean
  C4  the 64 datasets of the batched workload in one graph launch: every dataset's graph
      equals the oracle's and its own single call
  C5  a Poisson stream of 40 requests (n log-uniform 1,000-200,000, d = 50, k = 15, host and
      device-resident embeddings mixed) into the server: each result equals its alone-run
      and sampled rows equal the oracle
The reference's own parity file compares whole graphs the same way
(tests/test_device_parity.py:76-92).
"""

from __future__ import annotations

import hashlib
import json
import time
from pathlib import Path

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import mrustsc_amd  # noqa: E402
import oracle  # noqa: E402
from mrustsc_amd import synthetic  # noqa: E402

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).parent / "golden"


def _fixture(name: str) -> dict:
    return json.loads((GOLDEN / f"{name}_oracle_fixture.json").read_text())


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def check_against(fix: dict, ids: np.ndarray, dist: np.ndarray, label: str) -> None:
    """Rows in full first (a readable failure), then the chunk digests (which rows), then
    the whole-graph digests."""
    for r, row in fix["rows"].items():
        np.testing.assert_array_equal(ids[int(r)], np.array(row["ids"], dtype=np.uint32),
                                      err_msg=f"{label} row {r}")
        bits = np.array([int(b, 16) for b in row["dist_bits"]], dtype=np.uint32)
        np.testing.assert_array_equal(dist[int(r)].view(np.uint32), bits, err_msg=f"{label} row {r}")
    step = fix["chunk_rows"]
    bad = [i for i, want in enumerate(fix["chunks"])
           if digest(ids[i * step:(i + 1) * step]) + ":" + digest(dist[i * step:(i + 1) * step]) != want]
    assert not bad, f"{label}: rows differ from the oracle in chunks {bad} of {step} rows"
    assert digest(ids.astype(np.uint32)) == fix["ids_sha256"], label
    assert digest(dist.astype(np.float32)) == fix["dist_sha256"], label


def embedding(fix: dict, n: int) -> np.ndarray:
    x = synthetic.clustered(n, 50, seed=1)
    assert list(x.shape) == fix["shape"]
    assert digest(x) == fix["x_sha256"], "the generator did not reproduce the workload's embedding"
    return x


def device_build(x: np.ndarray, k: int, context, precision: str, metric: str):
    n, d = x.shape
    dx = torch.from_numpy(x).to("cuda")
    di = torch.full((n, k), -1, dtype=torch.int32, device="cuda")
    dd = torch.full((n, k), -1.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    st = mrustsc_amd.knn_device(dx.data_ptr(), n, d, k, di.data_ptr(), dd.data_ptr(),
                                precision=precision, metric=metric, context=context)
    ids, dist = di.cpu().numpy().view(np.uint32), dd.cpu().numpy()
    del dx, di, dd
    return ids, dist, st


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
def test_c2_whole_graph_equals_the_oracle(context, precision):
    fix = _fixture("c2")
    x = embedding(fix, 100_000)
    print(f"[case] C2 100000 x 50 {precision}", flush=True)
    ids, dist, st = device_build(x, 15, context, precision, "euclidean")
    print(st, flush=True)
    check_against(fix, ids, dist, f"C2 {precision}")


@pytest.mark.parametrize("k,precision", [(64, "bf16"), (64, "fp32"), (100, "bf16"), (100, "fp32"),
                                         (111, "bf16"), (128, "bf16"), (128, "fp32")])
def test_c2_large_k_whole_graph_equals_the_oracle(context, k, precision):
    fix = _fixture("c2_large_k")["graphs"][str(k)]
    x = embedding(fix, 100_000)
    print(f"[case] C2 100000 x 50 k={k} {precision}", flush=True)
    ids, dist, st = device_build(x, k, context, precision, "euclidean")
    print(st, flush=True)
    # KP = 128 up to k = 111, then KP = 256 (1024-slot buffers).
    assert st["path"] == precision and st["candidates"] == (128 if k <= 111 else 256)
    assert st["fallback_rows"] < 100
    check_against(fix, ids, dist, f"C2 k={k} {precision}")


@pytest.fixture(scope="module")
def c3_embedding():
    return embedding(_fixture("c3"), 1_000_000)


@pytest.mark.parametrize("precision,metric,name", [("bf16", "euclidean", "c3"),
                                                   ("bf16", "cosine", "c3_cosine"),
                                                   ("fp32", "euclidean", "c3")])
def test_c3_whole_graph_equals_the_oracle(context, c3_embedding, precision, metric, name):
    fix = _fixture(name)
    assert digest(c3_embedding) == fix["x_sha256"]
    print(f"[case] C3 1000000 x 50 {precision} {metric}", flush=True)
    ids, dist, st = device_build(c3_embedding, 15, context, precision, metric)
    print(st, flush=True)
    assert st["fallback_rows"] == 0
    check_against(fix, ids, dist, f"C3 {precision} {metric}")


def test_c4_batch_of_64_datasets_equals_the_oracle_and_single_calls(context):
    fix = _fixture("c4")
    specs = synthetic.batch_specs(64, seed=7)
    assert [rec["spec"] for rec in fix["datasets"]] == specs
    data, outs, entries = [], [], []
    for spec, rec in zip(specs, fix["datasets"]):
        x = synthetic.batch_dataset(spec)
        assert digest(x) == rec["x_sha256"], f"dataset {spec} not reproduced"
        dx = torch.from_numpy(x).to("cuda")
        di = torch.full((spec["n"], spec["k"]), -1, dtype=torch.int32, device="cuda")
        dd = torch.full((spec["n"], spec["k"]), -1.0, dtype=torch.float32, device="cuda")
        data.append((x, dx))
        outs.append((di, dd))
        entries.append((dx.data_ptr(), spec["n"], spec["d"], spec["k"], di.data_ptr(), dd.data_ptr()))
    torch.cuda.synchronize()
    print(f"[case] C4 batch of 64 datasets, {sum(s['n'] for s in specs)} cells", flush=True)
    with mrustsc_amd.Batch(entries, context=context) as batch:
        for replay in range(2):
            st = batch.run()
            print(replay, st, flush=True)
            for i, (spec, rec, (x, _), (di, dd)) in enumerate(zip(specs, fix["datasets"], data, outs)):
                ids, dist = di.cpu().numpy().view(np.uint32), dd.cpu().numpy()
                check_against(rec, ids, dist, f"C4 dataset {i} {spec} replay {replay}")
                if replay == 0:
                    w_ids, w_dist = mrustsc_amd.knn_hip(x, spec["k"], context=context)
                    np.testing.assert_array_equal(ids, w_ids, err_msg=f"dataset {i}")
                    np.testing.assert_array_equal(dist, w_dist, err_msg=f"dataset {i}")


def test_c5_poisson_stream_results_equal_their_alone_runs(context):
    rng = np.random.default_rng(5)
    count = 40
    reqs = []
    for i in range(count):
        n = int(round(np.exp(rng.uniform(np.log(1000), np.log(200000)))))
        reqs.append(synthetic.clustered(n, 50, seed=500 + i, n_clusters=int(rng.integers(5, 31))))
    alone = [mrustsc_amd.knn_hip(x, 15, context=context) for x in reqs]
    resident = [torch.from_numpy(x).to("cuda") if i % 2 else None for i, x in enumerate(reqs)]
    torch.cuda.synchronize()
    # Offered load: 40 requests over about 0.35 s, above what 4 slots finish in that time, so
    # requests queue, share slots and overlap on the device.
    gaps = rng.exponential(0.35 / count, size=count)
    print(f"[case] C5 Poisson stream of {count} requests, n {min(x.shape[0] for x in reqs)}-"
          f"{max(x.shape[0] for x in reqs)}", flush=True)
    with mrustsc_amd.Server(slots=4) as server:
        tickets = []
        t0 = time.perf_counter()
        due = t0
        for i, (x, gap) in enumerate(zip(reqs, gaps)):
            due += gap
            delay = due - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            if resident[i] is None:
                tickets.append(server.submit(x, 15))
            else:
                t = resident[i]
                tickets.append(server.submit_device(t.data_ptr(), t.shape[0], 50, 15, keep_alive=t))
        for i, (t, (ids_a, dist_a)) in enumerate(zip(tickets, alone)):
            ids, dist, info = server.wait(t)
            np.testing.assert_array_equal(ids, ids_a, err_msg=f"request {i}")
            np.testing.assert_array_equal(dist, dist_a, err_msg=f"request {i}")
            assert info["submit_ns"] <= info["admit_ns"] <= info["done_ns"]
        stats = server.stats()
    assert stats["completed"] == count and stats["peak_in_flight"] >= 2
    for i in range(0, count, 5):
        x = reqs[i]
        rows = np.linspace(0, x.shape[0] - 1, 3).astype(np.uint32)
        o_ids, o_dist = oracle.knn(x, 15, queries=rows)
        np.testing.assert_array_equal(alone[i][0][rows], o_ids, err_msg=f"request {i}")
        np.testing.assert_array_equal(alone[i][1][rows], o_dist, err_msg=f"request {i}")
