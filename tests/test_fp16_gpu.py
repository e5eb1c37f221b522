"""The local-origin fp16 candidate pass (precision "fp16"): every graph equals the oracle's
bit for bit, on the shapes and data kinds that stress its certificate (duplicates, integer
ties, near ties, cosine, every compiled chunk count, tiny and huge coordinate scales, a
work item holding sentinel rows), and at 1M cells against the committed whole-graph
fixtures. Shapes it does not serve take the bf16 path (stats["path"] says which ran)."""

from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import mrustsc_amd  # noqa: E402
import oracle  # noqa: E402
from mrustsc_amd import synthetic  # noqa: E402

pytestmark = pytest.mark.gpu

FP16 = "fp16"


def _check(x, k, context, metric="euclidean", order="grouped", expect_path=FP16):
    ids, dist, st = mrustsc_amd.knn_hip(x, k, precision="fp16", metric=metric, order=order,
                                        context=context, stats=True)
    print(st, flush=True)
    want_ids, want_dist = oracle.knn_fast(x, k, metric=metric)
    np.testing.assert_array_equal(ids, want_ids)
    np.testing.assert_array_equal(dist.view(np.uint32), want_dist.view(np.uint32))
    assert st["path"] == expect_path
    return st


@pytest.mark.parametrize("n,d,k", [(3000, 50, 15), (5000, 50, 30), (4000, 100, 15), (2500, 125, 47),
                                   (3000, 29, 15), (3000, 61, 15), (3000, 7, 5), (2000, 1, 3),
                                   (16, 50, 15), (700, 50, 15)])
def test_fp16_graphs_equal_the_oracle(context, n, d, k):
    print(f"[case] fp16 n={n} d={d} k={k}", flush=True)
    _check(synthetic.clustered(n, d, seed=n + d + k, n_clusters=6), k, context)


@pytest.mark.parametrize("kind", ["duplicates", "lattice", "near_ties", "cosine", "tiny", "huge", "offset"])
def test_fp16_edge_data_equals_the_oracle(context, kind):
    print(f"[case] fp16 edge {kind}", flush=True)
    rng = np.random.default_rng(7)
    metric = "euclidean"
    if kind == "duplicates":
        x = synthetic.clustered(4000, 30, seed=3, duplicate_fraction=0.25)
    elif kind == "lattice":
        x = np.floor(rng.uniform(0, 5, (3000, 6))).astype(np.float32)
    elif kind == "near_ties":
        base = synthetic.clustered(1500, 20, seed=4)
        x = np.concatenate([base, base + np.float32(1e-6)]).astype(np.float32)
    elif kind == "cosine":
        x, metric = synthetic.clustered(5000, 50, seed=5), "cosine"
    elif kind == "tiny":
        x = (synthetic.clustered(3000, 50, seed=6).astype(np.float64) * 1e-25).astype(np.float32)
    elif kind == "huge":
        x = (synthetic.clustered(3000, 50, seed=6).astype(np.float64) * 1e10).astype(np.float32)
    else:
        x = synthetic.clustered(3000, 50, seed=8) + np.float32(1e4)
    _check(x, 15, context, metric=metric)


def test_fp16_at_70000_cells_orders_rows_by_itself(context):
    print("[case] fp16 70000 x 50 automatic order", flush=True)
    st = _check(synthetic.clustered(70000, 50, seed=9), 15, context, order="auto")
    assert st["order"] == 2


@pytest.mark.parametrize("n,d,order", [(3000, 50, "input"), (3000, 126, "grouped"), (3000, 50, "auto")])
def test_fp16_falls_back_to_bf16_where_it_does_not_apply(context, n, d, order):
    print(f"[case] fp16 fallback n={n} d={d} order={order}", flush=True)
    _check(synthetic.clustered(n, d, seed=10), 15, context, order=order, expect_path="bf16")


@pytest.mark.parametrize("metric,name", [("euclidean", "c3"), ("cosine", "c3_cosine")])
def test_fp16_1m_whole_graph_equals_the_oracle_fixture(context, metric, name):
    fix = json.loads((Path(__file__).parent / "golden" / f"{name}_oracle_fixture.json").read_text())
    x = synthetic.clustered(1_000_000, 50, seed=1)
    assert hashlib.sha256(x.tobytes()).hexdigest() == fix["x_sha256"]
    print(f"[case] fp16 1M {metric}", flush=True)
    dx = torch.from_numpy(x).to("cuda")
    di = torch.full((1_000_000, 15), -1, dtype=torch.int32, device="cuda")
    dd = torch.full((1_000_000, 15), -1.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    st = mrustsc_amd.knn_device(dx.data_ptr(), 1_000_000, 50, 15, di.data_ptr(), dd.data_ptr(),
                                precision="fp16", metric=metric, context=context)
    print(st, flush=True)
    assert st["path"] == FP16
    ids, dist = di.cpu().numpy().view(np.uint32), dd.cpu().numpy()
    assert hashlib.sha256(ids.tobytes()).hexdigest() == fix["ids_sha256"]
    assert hashlib.sha256(dist.tobytes()).hexdigest() == fix["dist_sha256"]
