"""The batched path: many embeddings of uneven n, d and k in one persistent MFMA launch
captured in a HIP graph. Every dataset's result must equal its own single-call result
(and the oracle's) bit for bit, on every replay of the graph."""

from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import mrustsc_amd  # noqa: E402
import oracle  # noqa: E402
from mrustsc_amd import synthetic  # noqa: E402

pytestmark = pytest.mark.gpu

SPECS = [  # (n, d, k, seed)
    (3000, 50, 15, 1),
    (700, 30, 30, 2),
    (5000, 100, 15, 3),
    (16, 5, 15, 4),      # n = k + 1
    (1200, 50, 30, 5),
    (257, 30, 15, 6),
    (900, 20, 60, 7),    # 47 < k <= 111: 128 candidates inside the batch
    (2048, 100, 30, 8),
    (600, 20, 120, 9),   # k > 111: exact path inside the batch
]


def _device(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def _make(specs, precision="bf16", metric="euclidean"):
    data, outs, entries = [], [], []
    for n, d, k, seed in specs:
        x = synthetic.clustered(n, d, seed=seed, n_clusters=6)
        dx = _device(x)
        di = torch.full((n, k), -1, dtype=torch.int32, device="cuda")
        dd = torch.full((n, k), -1.0, dtype=torch.float32, device="cuda")
        data.append((x, dx))
        outs.append((di, dd))
        entries.append((dx.data_ptr(), n, d, k, di.data_ptr(), dd.data_ptr()))
    torch.cuda.synchronize()
    return data, outs, mrustsc_amd.Batch(entries, precision=precision, metric=metric)


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
def test_batch_equals_single_calls_and_the_oracle(context, precision):
    print(f"[case] batch of {len(SPECS)} precision={precision}", flush=True)
    data, outs, batch = _make(SPECS, precision)
    with batch:
        for replay in range(3):
            stats = batch.run()
            print(replay, stats, flush=True)
            for (n, d, k, _), (x, _), (di, dd) in zip(SPECS, data, outs):
                ids = di.cpu().numpy().view(np.uint32)
                dist = dd.cpu().numpy()
                want_ids, want_dist = mrustsc_amd.knn_hip(x, k, precision=precision, context=context)
                np.testing.assert_array_equal(ids, want_ids, err_msg=f"n={n} d={d} k={k} replay {replay}")
                np.testing.assert_array_equal(dist, want_dist, err_msg=f"n={n} d={d} k={k} replay {replay}")
                if n <= 3000:
                    o_ids, o_dist = oracle.knn(x, k)
                    np.testing.assert_array_equal(ids, o_ids)
                    np.testing.assert_array_equal(dist, o_dist)
            assert stats["datasets"] == len(SPECS)
            assert stats["cells"] == sum(s[0] for s in SPECS)


def test_batch_replays_follow_new_input_contents(context):
    specs = SPECS[:3]
    data, outs, batch = _make(specs)
    with batch:
        batch.run()
        # Same pointers, new contents: the replay must compute the new graph.
        fresh = [synthetic.clustered(n, d, seed=seed + 100, n_clusters=6) for n, d, k, seed in specs]
        for (x, dx), y in zip(data, fresh):
            dx.copy_(torch.from_numpy(y))
        torch.cuda.synchronize()
        batch.run()
        for (n, d, k, _), y, (di, dd) in zip(specs, fresh, outs):
            want_ids, want_dist = oracle.knn(y, k)
            np.testing.assert_array_equal(di.cpu().numpy().view(np.uint32), want_ids)
            np.testing.assert_array_equal(dd.cpu().numpy(), want_dist)


def test_batch_cosine(context):
    specs = SPECS[:4]
    data, outs, batch = _make(specs, metric="cosine")
    with batch:
        batch.run()
        for (n, d, k, _), (x, _), (di, dd) in zip(specs, data, outs):
            want_ids, want_dist = oracle.knn(x, k, metric="cosine")
            np.testing.assert_array_equal(di.cpu().numpy().view(np.uint32), want_ids)
            np.testing.assert_array_equal(dd.cpu().numpy(), want_dist)


def test_batch_rejects_an_invalid_dataset(context):
    x = _device(synthetic.clustered(100, 8, seed=1))
    di = torch.empty((100, 15), dtype=torch.int32, device="cuda")
    dd = torch.empty((100, 15), dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError, match="dataset 1"):
        mrustsc_amd.Batch([(x.data_ptr(), 100, 8, 15, di.data_ptr(), dd.data_ptr()),
                           (x.data_ptr(), 10, 8, 15, di.data_ptr(), dd.data_ptr())])


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
@pytest.mark.parametrize("n", [3000, 70000])
def test_a_one_dataset_batch_equals_its_single_call_and_the_oracle(context, precision, n):
    # A batch of one dataset still runs the batched pass (work items and groups); a layout
    # that inferred a single call from the count launched no pass at all.
    print(f"[case] one-dataset batch n={n} precision={precision}", flush=True)
    data, outs, batch = _make([(n, 50, 15, 31)], precision)
    (x, _), (di, dd) = data[0], outs[0]
    with batch:
        stats = batch.run()
        assert stats["pass_items"] > 0
        ids = di.cpu().numpy().view(np.uint32)
        dist = dd.cpu().numpy()
    want_ids, want_dist = mrustsc_amd.knn_hip(x, 15, precision=precision, context=context)
    np.testing.assert_array_equal(ids, want_ids)
    np.testing.assert_array_equal(dist, want_dist)
    rows = np.linspace(0, n - 1, 200).astype(np.uint32)
    o_ids, o_dist = oracle.knn(x, 15, queries=rows)
    np.testing.assert_array_equal(ids[rows], o_ids)
    np.testing.assert_array_equal(dist[rows], o_dist)
