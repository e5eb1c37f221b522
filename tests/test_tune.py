"""The autotuner of the candidate pass's work shape and its process-wide table.

CPU: the table's text format round-trips, malformed files are rejected without touching
the table, and the tuner checks its arguments before it needs a device.
GPU: tuning times several shapes, every shape returns the same graph (the tuner compares
them itself, and the oracle checks the result here), later calls with automatic tiles take
the tuned shape, batches built after tuning equal the untuned ones bit for bit, and the
batch tuner's chosen graph replays to the same outputs as the creation shape's."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import mrustsc_amd
from mrustsc_amd import _native, synthetic

HEADER = "# scrust_hip tuning table v1: precision metric kp nch n_bin tile splits order ms"


@pytest.fixture()
def empty_table():
    mrustsc_amd.tune_clear()
    yield
    mrustsc_amd.tune_clear()


def test_table_round_trips_through_a_file(tmp_path, empty_table):
    src = tmp_path / "in.txt"
    src.write_text("# a comment\n\n0 0 32 5 40 256 1 2 239.5\n  1 1 64 13 33 512 2 1 12.25\n")
    mrustsc_amd.tune_load(src)
    assert mrustsc_amd.tune_size() == 2
    out = tmp_path / "out.txt"
    mrustsc_amd.tune_save(out)
    lines = out.read_text().splitlines()
    assert lines[0] == HEADER
    assert lines[1:] == ["0 0 32 5 40 256 1 2 239.5000", "1 1 64 13 33 512 2 1 12.2500"]
    # Loading a saved table again replaces the same buckets.
    mrustsc_amd.tune_load(out)
    assert mrustsc_amd.tune_size() == 2
    mrustsc_amd.tune_clear()
    assert mrustsc_amd.tune_size() == 0


@pytest.mark.parametrize("line", [
    "0 0 32 5 40 100 1 2 1.0",      # tile not 128/256/512
    "0 0 64 5 40 128 1 1 1.0",      # tile below 4 x kp
    "0 0 32 5 40 256 5 1 1.0",      # splits x tile above 1024
    "0 0 32 5 40 256 0 1 1.0",      # no range
    "0 0 32 5 40 256 1 3 1.0",      # unknown order
    "2 0 32 5 40 256 1 1 1.0",      # the exact path has no shape
    "0 2 32 5 40 256 1 1 1.0",      # unknown metric
    "0 0 48 5 40 256 1 1 1.0",      # kp is 32 or 64
    "0 0 32 5 40 256 1 1",          # missing field
    "0 0 32 5 40 256 1 1 1.0 7",    # extra field
    "0 0 32 5 40 256 1 1 nan",      # time not finite
    "0 0 32 5 40 256 1 1 -1",       # negative time
    "zero 0 32 5 40 256 1 1 1.0",   # not a number
])
def test_malformed_lines_are_rejected_and_leave_the_table_unchanged(tmp_path, empty_table, line):
    good = tmp_path / "good.txt"
    good.write_text("0 0 32 5 40 256 1 2 1.0\n")
    mrustsc_amd.tune_load(good)
    bad = tmp_path / "bad.txt"
    bad.write_text("1 0 32 13 33 128 2 1 1.0\n" + line + "\n")
    with pytest.raises(ValueError, match="line 2"):
        mrustsc_amd.tune_load(bad)
    assert mrustsc_amd.tune_size() == 1


def test_missing_and_unwritable_paths_are_invalid_parameters(tmp_path, empty_table):
    with pytest.raises(ValueError, match="readable"):
        mrustsc_amd.tune_load(tmp_path / "absent.txt")
    with pytest.raises(ValueError, match="writable"):
        mrustsc_amd.tune_save(tmp_path / "no_such_dir" / "t.txt")


def test_tuner_checks_arguments_before_the_device():
    lib = _native.load()
    x = np.zeros((8, 3), dtype=np.float32)
    res = _native.TuneResult()
    opts = _native.KnnOptions(0, 0, 0, 0)
    assert lib.scrust_hip_tune_device(None, x.ctypes.data, 8, 3, 0, ctypes.byref(opts), 0,
                                      ctypes.byref(res)) == _native.INVALID_PARAMETER
    assert lib.scrust_hip_tune_device(None, x.ctypes.data, 8, 3, 8, ctypes.byref(opts), 0,
                                      ctypes.byref(res)) == _native.INVALID_PARAMETER
    assert lib.scrust_hip_tune_device(None, x.ctypes.data, 8, 3, 2, ctypes.byref(opts), 0,
                                      None) == _native.INVALID_PARAMETER
    assert lib.scrust_hip_tune_device(None, x.ctypes.data, 8, 3, 2, ctypes.byref(opts), 6,
                                      ctypes.byref(res)) == _native.INVALID_PARAMETER
    bad = _native.KnnOptions(5, 0, 0, 0)
    assert lib.scrust_hip_tune(None, x.ctypes.data, 8, 3, 2, ctypes.byref(bad), 0,
                               ctypes.byref(res)) == _native.INVALID_PARAMETER
    assert lib.scrust_hip_tune(None, x.ctypes.data, 8, 3, 2, ctypes.byref(opts), 0,
                               ctypes.byref(res)) == _native.NO_GPU


# ---------------------------------------------------------------------------------------
# On the device.

def _oracle_equal(x, k, ids, dist, metric="euclidean"):
    import oracle

    want_ids, want_dist = oracle.knn(x, k, metric=metric)
    np.testing.assert_array_equal(ids, want_ids)
    np.testing.assert_array_equal(dist, want_dist)


@pytest.mark.gpu
@pytest.mark.parametrize(("n", "d", "k", "precision", "metric"), [
    (3000, 50, 15, "bf16", "euclidean"),
    (9000, 30, 30, "bf16", "euclidean"),   # kp 64: 256 and 512 slots only
    (6000, 100, 15, "fp32", "euclidean"),
    (5000, 50, 15, "bf16", "cosine"),
])
def test_tuning_keeps_the_graph_and_later_calls_take_the_tuned_shape(context, empty_table, n, d, k,
                                                                     precision, metric):
    print(f"[case] tune n={n} d={d} k={k} {precision} {metric}", flush=True)
    x = synthetic.clustered(n, d, seed=n + d, n_clusters=8)
    ids0, dist0, st0 = mrustsc_amd.knn_hip(x, k, precision=precision, metric=metric, context=context,
                                           stats=True)
    assert st0["tuned"] == 0
    r = mrustsc_amd.tune(x, k, precision=precision, metric=metric, context=context)
    print(r, flush=True)
    assert r["shapes"] >= 3 and r["reused"] == 0
    assert r["tile"] in (128, 256, 512) and r["tile"] >= (128 if k <= 15 else 256)
    assert r["order"] in (1, 2) and r["splits"] >= 1
    assert 0 < r["best_ms"] <= r["default_ms"] <= r["worst_ms"]
    assert mrustsc_amd.tune_size() == 1
    ids, dist, st = mrustsc_amd.knn_hip(x, k, precision=precision, metric=metric, context=context, stats=True)
    assert st["tuned"] == 1
    assert (st["tile"], st["splits"], st["order"]) == (r["tile"], r["splits"], r["order"])
    np.testing.assert_array_equal(ids, ids0)
    np.testing.assert_array_equal(dist, dist0)
    _oracle_equal(x, k, ids, dist, metric)
    # An explicit tile bypasses the table.
    _, _, st_fixed = mrustsc_amd.knn_hip(x, k, precision=precision, metric=metric, context=context,
                                         stats=True, tile=256)
    assert st_fixed["tuned"] == 0 and st_fixed["tile"] == 256
    # The same bucket again: skipped on request, re-tuned otherwise.
    again = mrustsc_amd.tune(x, k, precision=precision, metric=metric, context=context, skip_known=True)
    assert again["reused"] == 1 and again["shapes"] == 0
    assert (again["tile"], again["splits"], again["order"]) == (r["tile"], r["splits"], r["order"])


@pytest.mark.gpu
def test_a_neighbouring_bucket_is_untouched_and_the_exact_path_has_nothing_to_tune(context, empty_table):
    x = synthetic.clustered(4000, 50, seed=3, n_clusters=5)
    mrustsc_amd.tune(x, 15, context=context)
    y = synthetic.clustered(8000, 50, seed=4, n_clusters=5)  # two half-octaves up
    _, _, st = mrustsc_amd.knn_hip(y, 15, context=context, stats=True)
    assert st["tuned"] == 0
    z = synthetic.clustered(600, 20, seed=5, n_clusters=5)
    r = mrustsc_amd.tune(z, 120, context=context)  # k > 111: one shape (1024-slot buffers)
    assert r["shapes"] == 0 and mrustsc_amd.tune_size() == 1


@pytest.mark.gpu
def test_saved_table_drives_a_fresh_table_and_batches_stay_bit_exact(context, empty_table, tmp_path):
    torch = pytest.importorskip("torch")
    specs = [(3000, 50, 15, 1), (700, 30, 30, 2), (5000, 100, 15, 3), (1500, 50, 15, 4)]
    data, outs, entries = [], [], []
    for n, d, k, seed in specs:
        x = synthetic.clustered(n, d, seed=seed, n_clusters=6)
        dx = torch.from_numpy(x).to("cuda")
        di = torch.full((n, k), -1, dtype=torch.int32, device="cuda")
        dd = torch.full((n, k), -1.0, dtype=torch.float32, device="cuda")
        data.append((x, dx))
        outs.append((di, dd))
        entries.append((dx.data_ptr(), n, d, k, di.data_ptr(), dd.data_ptr()))
    torch.cuda.synchronize()
    with mrustsc_amd.Batch(entries, context=context) as batch:
        batch.run()
        before = [(di.cpu().numpy().copy(), dd.cpu().numpy().copy()) for di, dd in outs]
    for (n, d, k, _), (x, dx) in zip(specs, data):
        mrustsc_amd.tune_device(dx.data_ptr(), n, d, k, context=context)
    path = tmp_path / "table.txt"
    mrustsc_amd.tune_save(path)
    mrustsc_amd.tune_clear()
    mrustsc_amd.tune_load(path)
    assert mrustsc_amd.tune_size() == len(specs)
    for di, dd in outs:
        di.fill_(-1)
        dd.fill_(-1.0)
    torch.cuda.synchronize()
    with mrustsc_amd.Batch(entries, context=context) as batch:
        for replay in range(2):
            batch.run()
            for (n, d, k, _), (x, _), (di, dd), (b_ids, b_dist) in zip(specs, data, outs, before):
                np.testing.assert_array_equal(di.cpu().numpy(), b_ids, err_msg=f"n={n} replay {replay}")
                np.testing.assert_array_equal(dd.cpu().numpy(), b_dist, err_msg=f"n={n} replay {replay}")
    x, _ = data[0]
    _oracle_equal(x, specs[0][2], before[0][0].view(np.uint32), before[0][1])


def test_batch_tuner_checks_its_arguments():
    lib = _native.load()
    res = _native.BatchTuneResult()
    assert lib.scrust_hip_batch_tune(None, ctypes.byref(res)) == _native.INVALID_PARAMETER


@pytest.mark.gpu
@pytest.mark.parametrize(("precision", "metric"), [("bf16", "euclidean"), ("fp32", "euclidean"),
                                                   ("bf16", "cosine")])
def test_batch_tuner_keeps_every_graph_and_the_tuned_replays_stay_bit_exact(context, empty_table, precision,
                                                                            metric):
    torch = pytest.importorskip("torch")
    specs = [(3000, 50, 15, 1), (700, 30, 30, 2), (5000, 100, 15, 3), (16, 5, 15, 4),
             (900, 20, 60, 7), (2048, 100, 30, 8)]  # k = 60: exact path inside the batch
    data, outs, entries = [], [], []
    for n, d, k, seed in specs:
        x = synthetic.clustered(n, d, seed=seed, n_clusters=6)
        dx = torch.from_numpy(x).to("cuda")
        di = torch.full((n, k), -1, dtype=torch.int32, device="cuda")
        dd = torch.full((n, k), -1.0, dtype=torch.float32, device="cuda")
        data.append((x, dx))
        outs.append((di, dd))
        entries.append((dx.data_ptr(), n, d, k, di.data_ptr(), dd.data_ptr()))
    torch.cuda.synchronize()
    print(f"[case] batch tune of {len(specs)} datasets {precision} {metric}", flush=True)
    with mrustsc_amd.Batch(entries, precision=precision, metric=metric, context=context) as batch:
        batch.run()
        before = [(di.cpu().numpy().copy(), dd.cpu().numpy().copy()) for di, dd in outs]
        r = batch.tune()
        print(r, flush=True)
        assert r["shapes"] == 18 and (r["default_cut"], r["default_tile"]) == (3, 256)
        assert r["cut"] in (1, 2, 3, 4, 6, 8) and r["tile"] in (128, 256, 512)
        assert 0 < r["best_ms"] <= r["default_ms"] <= r["worst_ms"]
        for replay in range(2):
            for di, dd in outs:
                di.fill_(-1)
                dd.fill_(-1.0)
            torch.cuda.synchronize()
            st = batch.run()
            assert st["datasets"] == len(specs)
            for (n, d, k, _), (di, dd), (b_ids, b_dist) in zip(specs, outs, before):
                np.testing.assert_array_equal(di.cpu().numpy(), b_ids, err_msg=f"n={n} replay {replay}")
                np.testing.assert_array_equal(dd.cpu().numpy(), b_dist, err_msg=f"n={n} replay {replay}")
    for (n, d, k, _), (x, _), (b_ids, b_dist) in zip(specs, data, before):
        _oracle_equal(x, k, b_ids.view(np.uint32), b_dist, metric)
