"""The serving loop: admission by device memory, concurrent builds, and results that are
the same alone and in the mix (bit for bit: every build is deterministic)."""

from __future__ import annotations

import numpy as np
import pytest

import mrustsc_amd
from mrustsc_amd import synthetic

pytestmark = pytest.mark.gpu


def _requests(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(300), np.log(20000))))
        d = int(rng.choice([8, 30, 50]))
        k = int(rng.choice([5, 15, 30]))
        out.append((synthetic.clustered(n, d, seed=seed * 100 + i, n_clusters=5), k))
    return out


def test_results_in_the_mix_equal_results_alone(context):
    reqs = _requests(24, seed=1)
    alone = [mrustsc_amd.knn_hip(x, k, context=context) for x, k in reqs]
    print(f"[case] server 24 requests, 4 slots", flush=True)
    with mrustsc_amd.Server(slots=4) as server:
        tickets = [server.submit(x, k) for x, k in reqs]
        for (ids_a, dist_a), t in zip(alone, tickets):
            ids, dist, info = server.wait(t)
            np.testing.assert_array_equal(ids, ids_a)
            np.testing.assert_array_equal(dist, dist_a)
            assert info["submit_ns"] <= info["admit_ns"] <= info["done_ns"]
        stats = server.stats()
    assert stats["completed"] == 24 and stats["submitted"] == 24
    assert stats["peak_in_flight"] >= 2, "builds never overlapped"


def test_admission_respects_the_memory_budget(context):
    reqs = _requests(12, seed=2)
    budget = 96 << 20
    print(f"[case] server budget {budget}", flush=True)
    with mrustsc_amd.Server(slots=8, memory_bytes=budget) as server:
        tickets = []
        for x, k in reqs:
            try:
                tickets.append((server.submit(x, k), x, k))
            except RuntimeError as err:  # larger than the whole budget
                assert "budget" in str(err)
        assert tickets, "every request was larger than the budget"
        for t, x, k in tickets:
            ids, dist, _ = server.wait(t)
            want_ids, want_dist = mrustsc_amd.knn_hip(x, k, context=context)
            np.testing.assert_array_equal(ids, want_ids)
            np.testing.assert_array_equal(dist, want_dist)
        stats = server.stats()
    assert stats["peak_memory_bytes"] <= budget


def test_a_small_request_passes_a_head_that_does_not_fit(context):
    # Budget for one large build and a small one, not two large ones: the second large
    # request waits for memory at the head of the queue while the small one behind it is
    # admitted; every result still equals its alone-run.
    big = [synthetic.clustered(60000, 50, seed=70 + i) for i in range(2)]
    small = synthetic.clustered(3000, 50, seed=72)
    sizes = []
    for x in (big[0], small):
        with mrustsc_amd.Server(slots=2, memory_bytes=1 << 30) as probe:
            probe.wait(probe.submit(x, 15))
            sizes.append(probe.stats()["peak_memory_bytes"])
    budget = sizes[0] + sizes[1] + (sizes[0] - sizes[1]) // 2
    print(f"[case] server budget {budget} (large {sizes[0]}, small {sizes[1]})", flush=True)
    with mrustsc_amd.Server(slots=4, memory_bytes=budget) as server:
        tickets = [server.submit(big[0], 15), server.submit(big[1], 15), server.submit(small, 15)]
        results = [server.wait(t) for t in tickets]
        stats = server.stats()
    assert stats["peak_memory_bytes"] <= budget
    infos = [r[2] for r in results]
    assert infos[2]["admit_ns"] < infos[1]["admit_ns"], infos
    for (ids, dist, _), x in zip(results, (big[0], big[1], small)):
        want_ids, want_dist = mrustsc_amd.knn_hip(x, 15, context=context)
        np.testing.assert_array_equal(ids, want_ids)
        np.testing.assert_array_equal(dist, want_dist)


def test_a_request_larger_than_the_budget_is_refused():
    with mrustsc_amd.Server(slots=2, memory_bytes=16 << 20) as server:
        x = synthetic.clustered(60000, 50, seed=3)
        with pytest.raises(RuntimeError, match="budget"):
            server.submit(x, 15)
        assert server.stats()["rejected"] == 1


def test_a_bad_request_fails_alone(context):
    good = synthetic.clustered(2000, 16, seed=4)
    bad = good.copy()
    bad[3, 2] = np.inf
    with mrustsc_amd.Server(slots=2) as server:
        t_bad = server.submit(bad, 10)
        t_good = server.submit(good, 10)
        with pytest.raises(ValueError):
            server.wait(t_bad)
        ids, dist, _ = server.wait(t_good)
    want_ids, want_dist = mrustsc_amd.knn_hip(good, 10, context=context)
    np.testing.assert_array_equal(ids, want_ids)
    np.testing.assert_array_equal(dist, want_dist)


def test_device_resident_requests_equal_host_requests(context):
    torch = pytest.importorskip("torch")
    reqs = _requests(8, seed=5)
    alone = [mrustsc_amd.knn_hip(x, k, context=context) for x, k in reqs]
    tensors = [torch.from_numpy(x).to("cuda") for x, _ in reqs]
    torch.cuda.synchronize()
    print("[case] server device-resident requests", flush=True)
    with mrustsc_amd.Server(slots=3) as server:
        tickets = [server.submit_device(t.data_ptr(), t.shape[0], t.shape[1], k, keep_alive=t)
                   for t, (_, k) in zip(tensors, reqs)]
        for (ids_a, dist_a), t in zip(alone, tickets):
            ids, dist, _ = server.wait(t)
            np.testing.assert_array_equal(ids, ids_a)
            np.testing.assert_array_equal(dist, dist_a)


def test_a_large_request_in_the_mix_equals_its_input_order_build(context):
    # 80,000 cells: the server's build groups the rows (automatic order); the result equals
    # a build in input order.
    reqs = [(synthetic.clustered(80000, 50, seed=7), 15)] + _requests(6, seed=3)
    plain = [mrustsc_amd.knn_hip(x, k, context=context, order="input") for x, k in reqs]
    print("[case] server mix with an 80,000-cell request", flush=True)
    with mrustsc_amd.Server(slots=4) as server:
        tickets = [server.submit(x, k) for x, k in reqs]
        for (ids_a, dist_a), t in zip(plain, tickets):
            ids, dist, _ = server.wait(t)
            np.testing.assert_array_equal(ids, ids_a)
            np.testing.assert_array_equal(dist, dist_a)


def test_requests_take_the_tuning_table_and_still_equal_results_alone(context, tmp_path):
    # A table that moves every request's bucket off the default shape (smallest buffer, two
    # reference ranges, input order): server builds take it and give the untuned results.
    reqs = _requests(12, seed=3)
    alone = [mrustsc_amd.knn_hip(x, k, context=context) for x, k in reqs]
    lines = []
    for x, k in reqs:
        n, d = x.shape
        kp = 32 if k + 17 <= 32 else 64
        nch = (3 * d + 3 + 31) // 32
        nbin = int(round(2 * np.log2(n)))
        lines.append(f"0 0 {kp} {nch} {nbin} {4 * kp} 2 1 1.0")
    table = tmp_path / "table.txt"
    table.write_text("\n".join(lines) + "\n")
    mrustsc_amd.tune_clear()
    mrustsc_amd.tune_load(table)
    try:
        _, _, st = mrustsc_amd.knn_hip(reqs[0][0], reqs[0][1], context=context, stats=True)
        assert st["tuned"] == 1 and st["splits"] == 2 and st["order"] == 1
        print(f"[case] server 12 requests with a tuning table", flush=True)
        with mrustsc_amd.Server(slots=4) as server:
            tickets = [server.submit(x, k) for x, k in reqs]
            for (ids_a, dist_a), t in zip(alone, tickets):
                ids, dist, _ = server.wait(t)
                np.testing.assert_array_equal(ids, ids_a)
                np.testing.assert_array_equal(dist, dist_a)
    finally:
        mrustsc_amd.tune_clear()


def test_the_work_shape_is_fixed_at_submission(context, tmp_path):
    # Requests queued behind a long build were laid out (and their device regions sized)
    # with an empty tuning table; retuning their bucket to a larger buffer and more ranges
    # while they wait must not change the plan admission places into those regions.
    mrustsc_amd.tune_clear()
    first = synthetic.clustered(300000, 50, seed=41)
    small = [synthetic.clustered(6000, 50, seed=42 + i) for i in range(6)]
    alone = [mrustsc_amd.knn_hip(x, 15, context=context) for x in small]
    nbin = int(round(2 * np.log2(6000)))
    table = tmp_path / "table.txt"
    table.write_text(f"0 0 32 5 {nbin} 512 2 1 1.0\n")
    print("[case] server: tuning table changed between submission and admission", flush=True)
    try:
        with mrustsc_amd.Server(slots=1) as server:
            t_first = server.submit(first, 15)
            tickets = [server.submit(x, 15) for x in small]
            mrustsc_amd.tune_load(table)
            server.wait(t_first)
            for (ids_a, dist_a), t in zip(alone, tickets):
                ids, dist, _ = server.wait(t)
                np.testing.assert_array_equal(ids, ids_a)
                np.testing.assert_array_equal(dist, dist_a)
        # The retuned shape itself gives the same graph.
        ids, dist, st = mrustsc_amd.knn_hip(small[0], 15, context=context, stats=True)
        assert st["tuned"] == 1 and st["tile"] == 512
        np.testing.assert_array_equal(ids, alone[0][0])
        np.testing.assert_array_equal(dist, alone[0][1])
    finally:
        mrustsc_amd.tune_clear()
