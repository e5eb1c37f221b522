"""Parity of the HIP kNN path with the CPU oracle, through the C ABI.

Mirrors the reference's GPU tests (crates/scrust-gpu/src/kernels/knn.rs:327-424), its
device-parity file (tests/test_device_parity.py) and the kNN audits
(tests/test_neighbors_audit.py:137-153, :340-387). The bar is stricter than the
reference's: indices AND distances must equal the oracle's bit for bit, because both
sides form every distance with the same f32 operations.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

# torch must be loaded before libscrust_hip.so so the process has one HIP runtime and
# device pointers can be shared with torch tensors (torch ships its own libamdhip64).
torch = pytest.importorskip("torch")

import mrustsc_amd  # noqa: E402
import oracle
from mrustsc_amd import synthetic

pytestmark = pytest.mark.gpu

GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_known_answers.json").read_text())
PRECISIONS = ("bf16", "fp32", "exact")


def run(context, x, k, precision="bf16", metric="euclidean", stats=False):
    print(f"[case] n={x.shape[0]} d={x.shape[1]} k={k} precision={precision} metric={metric}",
          flush=True)
    return mrustsc_amd.knn_hip(x, k, precision=precision, metric=metric, context=context,
                               stats=stats)


def assert_matches_oracle(context, x, k, precision, metric="euclidean", queries=None):
    ids, dist = run(context, x, k, precision, metric)
    want_ids, want_dist = oracle.knn(x, k, metric=metric, queries=queries)
    if queries is not None:
        ids, dist = ids[queries], dist[queries]
    np.testing.assert_array_equal(ids, want_ids, err_msg=f"{precision}: neighbour ids differ")
    np.testing.assert_array_equal(dist, want_dist, err_msg=f"{precision}: distances differ")
    return ids, dist


# --- crates/scrust-gpu/src/kernels/knn.rs tests -------------------------------------------


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize(("n", "d", "k"), [(64, 8, 1), (100, 5, 7), (257, 3, 15), (300, 32, 16)])
def test_matches_the_oracle_on_random_embeddings(context, precision, n, d, k):
    x = synthetic.random_embedding(n, d, n)
    assert_matches_oracle(context, x, k, precision)


def test_matches_the_oracle_at_the_largest_supported_k(context):
    k = mrustsc_amd.max_supported_k(context)
    x = synthetic.random_embedding(k + 71, 4, 7)
    assert_matches_oracle(context, x, k, "bf16")


@pytest.mark.parametrize("precision", PRECISIONS)
def test_finds_the_obvious_neighbours_of_points_on_a_line(context, precision):
    ids, dist = run(context, synthetic.line(50), 2, precision)
    case = GOLDEN["line50_k2"]
    for row, want in case["rows"].items():
        assert ids[int(row)].tolist() == want
    for row, want in case["distances"].items():
        assert dist[int(row)].tolist() == want


@pytest.mark.parametrize("precision", PRECISIONS)
def test_duplicated_points_break_ties_by_the_smaller_index(context, precision):
    x = np.zeros((41, 2), dtype=np.float32)
    x[40, 0] = 100.0
    ids, _ = run(context, x, 3, precision)
    for row, want in GOLDEN["origin40_k3"]["rows"].items():
        assert ids[int(row)].tolist() == want


@pytest.mark.parametrize("precision", PRECISIONS)
def test_two_runs_produce_identical_output(context, precision):
    x = synthetic.random_embedding(500, 6, 11)
    a = run(context, x, 9, precision)
    b = run(context, x, 9, precision)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_rejects_parameters_it_cannot_serve(context):
    x = synthetic.random_embedding(8, 3, 3)
    for k in (0, 8, 9):
        with pytest.raises(ValueError):
            run(context, x, k)
    wide = synthetic.random_embedding(4096, 2, 4)
    with pytest.raises(ValueError):
        run(context, wide, mrustsc_amd.max_supported_k(context) + 1)


def test_rejects_non_finite_values(context):
    x = synthetic.random_embedding(50, 3, 3)
    x[7, 1] = np.nan
    with pytest.raises(ValueError):
        run(context, x, 4)


# --- crates/scrust-core/src/neighbors.rs known answers --------------------------------------


@pytest.mark.parametrize("precision", PRECISIONS)
def test_neighbours_on_a_line_and_a_grid(context, precision):
    ids, dist = run(context, synthetic.line(6), 2, precision)
    for row, want in GOLDEN["line6_k2"]["rows"].items():
        assert ids[int(row)].tolist() == want
    assert dist[0].tolist() == GOLDEN["line6_k2"]["distances"]["0"]
    ids, dist = run(context, synthetic.grid(3), 4, precision)
    for row, want in GOLDEN["grid3_k4"]["rows"].items():
        assert ids[int(row)].tolist() == want
    np.testing.assert_allclose(dist[0], GOLDEN["grid3_k4"]["corner_distances"], atol=1e-6)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_matches_scanpy_neighbour_sets(context, precision):
    case = GOLDEN["scanpy"]
    x = synthetic.reference_embedding(case["n_cells"], case["n_dims"], case["seed"])
    ids, _ = assert_matches_oracle(context, x, case["k"], precision)
    for cell, want in zip(case["sample_cells"], case["neighbors"]):
        shared = len(set(ids[cell].tolist()) & set(want))
        assert shared * 10 >= case["k"] * 9, f"cell {cell}: only {shared} shared"


# --- tests/test_device_parity.py ---------------------------------------------------------


def _normal(n, d, seed):
    return np.random.default_rng(seed).normal(size=(n, d)).astype(np.float32)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_duplicate_cells_are_at_distance_zero(context, precision):
    x = _normal(200, 6, 9)
    x[:20] = x[0]
    _, dist = assert_matches_oracle(context, x, 14, precision)
    assert np.count_nonzero(dist[:20]) == 0


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize(("n", "d", "seed"), [(400, 20, 3), (250, 50, 11)])
def test_the_graph_matches_the_oracle_on_normal_data(context, precision, n, d, seed):
    assert_matches_oracle(context, _normal(n, d, seed), 14, precision)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_a_knot_tighter_than_the_resolution_collapses_like_the_oracle(context, precision):
    x = _normal(150, 8, 21)
    x[:10] = x[0] + np.float32(1e-7) * np.arange(10, dtype=np.float32)[:, None]
    assert_matches_oracle(context, x, 5, precision)


# --- tests/test_neighbors_audit.py -------------------------------------------------------


@pytest.mark.parametrize("precision", PRECISIONS)
def test_ties_break_towards_the_smaller_index_on_a_lattice(context, precision):
    lat = synthetic.lattice(8)
    ids, dist = assert_matches_oracle(context, lat, 14, precision)
    _, exact = oracle.knn_fp64(lat, 14)
    np.testing.assert_allclose(dist, exact, atol=1e-6)
    for cell in range(len(lat)):
        for a in range(13):
            if np.isclose(dist[cell, a], dist[cell, a + 1], atol=1e-6):
                assert ids[cell, a] < ids[cell, a + 1]


@pytest.mark.parametrize("precision", ("bf16", "fp32"))
@pytest.mark.parametrize("offset", [0.0, 1e2, 1e4, 1e5])
def test_distances_survive_an_embedding_far_from_the_origin(context, precision, offset):
    x = (np.random.default_rng(2).normal(size=(300, 12)) + offset).astype(np.float32)
    ids, dist = assert_matches_oracle(context, x, 14, precision)
    want_ids, want_dist = oracle.knn_fp64(x, 14)
    for cell in range(x.shape[0]):
        assert set(ids[cell].tolist()) == set(want_ids[cell].tolist()), f"cell {cell}"
    np.testing.assert_allclose(dist, want_dist, rtol=1e-4, atol=1e-4 * max(1.0, offset * 1e-4))


# --- the benchmark's data at test sizes ---------------------------------------------------


@pytest.mark.parametrize("precision", PRECISIONS)
def test_clustered_mixture_matches_the_oracle(context, precision):
    x = synthetic.clustered(6000, 50, seed=3)
    assert_matches_oracle(context, x, 15, precision)


@pytest.mark.parametrize(("d", "k"), [(30, 30), (100, 15), (100, 30), (7, 3)])
def test_other_dimensions_and_k(context, d, k):
    x = synthetic.clustered(3000, d, seed=d + k)
    for precision in PRECISIONS:
        assert_matches_oracle(context, x, k, precision)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_cosine_matches_the_oracle(context, precision):
    x = synthetic.clustered(4000, 50, seed=5)
    x[11] = 0.0  # a zero row stays a zero vector: every cell is a near-tie from it
    ids, dist = assert_matches_oracle(context, x, 15, precision, metric="cosine")
    want_ids, want = oracle.knn_fp64(x, 15, metric="cosine")
    rows = np.arange(x.shape[0]) != 11
    np.testing.assert_allclose(dist[rows], want[rows], atol=2e-6)


def test_all_paths_agree_on_a_larger_mixture(context):
    x = synthetic.clustered(40000, 50, seed=9)
    results = {p: run(context, x, 15, p, stats=True) for p in PRECISIONS}
    for p in ("fp32", "exact"):
        np.testing.assert_array_equal(results["bf16"][0], results[p][0])
        np.testing.assert_array_equal(results["bf16"][1], results[p][1])
    queries = np.random.default_rng(0).choice(x.shape[0], 400, replace=False)
    want_ids, want_dist = oracle.knn(x, 15, queries=queries)
    np.testing.assert_array_equal(results["bf16"][0][queries], want_ids)
    np.testing.assert_array_equal(results["bf16"][1][queries], want_dist)
    for p in ("bf16", "fp32"):
        s = results[p][2]
        print(p, s, flush=True)
        assert s["fallback_rows"] <= x.shape[0] // 100


def test_device_entry_point_matches_the_host_entry_point(context):
    x = synthetic.clustered(5000, 50, seed=4)
    ids, dist = run(context, x, 15)
    dx = torch.from_numpy(x).to("cuda")
    di = torch.empty((5000, 15), dtype=torch.int32, device="cuda")
    dd = torch.empty((5000, 15), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    mrustsc_amd.knn_device(dx.data_ptr(), 5000, 50, 15, di.data_ptr(), dd.data_ptr(),
                           context=context)
    np.testing.assert_array_equal(di.cpu().numpy().view(np.uint32), ids)
    np.testing.assert_array_equal(dd.cpu().numpy(), dist)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_boundary_near_ties_are_resolved_in_fp64(context, precision):
    # ids 1 and 2 sit at f32 distance exactly 1 from cell 0; in f64 id 2 is nearer.
    a, b = np.float32(1e-4), np.float32(1e-5)
    x = np.array([[0, 0], [1, a], [1, b], [-51, -a], [49, -b]], dtype=np.float32)
    ids, dist = assert_matches_oracle(context, x, 1, precision)
    assert ids[0].tolist() == [2] and dist[0].tolist() == [1.0]
    ids, _ = assert_matches_oracle(context, x, 2, precision)
    assert ids[0].tolist() == [1, 2]


@pytest.mark.parametrize("precision", ("bf16", "fp32"))
@pytest.mark.parametrize("tile", [128, 256, 512])
@pytest.mark.parametrize("k", [5, 15, 30, 47, 48, 100, 111, 128])
def test_every_candidate_buffer_size_matches_the_oracle(context, precision, tile, k):
    # The buffer size changes when rows are compacted, never the result: each size, with
    # KP = 32 (k <= 15), KP = 64 (k <= 47, where 128 slots are raised to 256), KP = 128
    # (k <= 111, where every size is raised to 512) and KP = 256 (every size raised to 1024).
    x = synthetic.clustered(2500, 24, seed=k + tile)
    print(f"[case] tile={tile} k={k} precision={precision}", flush=True)
    ids, dist = mrustsc_amd.knn_hip(x, k, precision=precision, tile=tile, context=context)
    want_ids, want_dist = oracle.knn(x, k)
    np.testing.assert_array_equal(ids, want_ids)
    np.testing.assert_array_equal(dist, want_dist)


# --- the pass's row order ---------------------------------------------------------------
# Rows grouped by nearest pivot (automatic from 65,536 cells, forced here on small inputs)
# change only the order the pass visits pairs in: every case must still equal the oracle.


def grouped_cases():
    rng = np.random.default_rng(17)
    dup = rng.normal(size=(600, 12)).astype(np.float32)
    dup[100:140] = dup[5]
    dup[300:303] = dup[299]
    return {
        "random": (rng.normal(size=(300, 32)).astype(np.float32), 16, "euclidean"),
        "duplicates": (dup, 15, "euclidean"),
        "lattice": (synthetic.lattice(8), 6, "euclidean"),
        "line": (synthetic.line(500), 9, "euclidean"),
        "clustered": (synthetic.clustered(6000, 50, seed=31), 15, "euclidean"),
        "d100_k30": (synthetic.clustered(3000, 100, seed=32), 30, "euclidean"),
        "cosine": (synthetic.clustered(4000, 50, seed=33), 15, "cosine"),
        "near_ties": (np.array([[0, 0], [1, 1e-4], [1, 1e-5], [-51, -1e-4], [49, -1e-5]],
                               dtype=np.float32), 2, "euclidean"),
    }


@pytest.mark.parametrize("precision", ("bf16", "fp32"))
@pytest.mark.parametrize("case", sorted(grouped_cases()))
def test_grouped_row_order_matches_the_oracle(context, precision, case):
    x, k, metric = grouped_cases()[case]
    print(f"[case] grouped order {case} precision={precision}", flush=True)
    ids, dist = mrustsc_amd.knn_hip(x, k, precision=precision, metric=metric, context=context,
                                    order="grouped")
    want_ids, want_dist = oracle.knn(x, k, metric=metric)
    np.testing.assert_array_equal(ids, want_ids)
    np.testing.assert_array_equal(dist, want_dist)


@pytest.mark.parametrize("precision", ("bf16", "fp32"))
def test_automatic_order_at_scale_equals_input_order_and_the_oracle(context, precision):
    # 70,000 cells: the automatic setting groups the rows and rotates each block's scan.
    x = synthetic.clustered(70000, 50, seed=34)
    print(f"[case] automatic order n=70000 precision={precision}", flush=True)
    got = mrustsc_amd.knn_hip(x, 15, precision=precision, context=context)
    plain = mrustsc_amd.knn_hip(x, 15, precision=precision, context=context, order="input")
    np.testing.assert_array_equal(got[0], plain[0])
    np.testing.assert_array_equal(got[1], plain[1])
    queries = np.random.default_rng(1).choice(x.shape[0], 300, replace=False)
    want_ids, want_dist = oracle.knn(x, 15, queries=queries)
    np.testing.assert_array_equal(got[0][queries], want_ids)
    np.testing.assert_array_equal(got[1][queries], want_dist)


def test_unknown_order_is_rejected():
    with pytest.raises(ValueError, match="order"):
        mrustsc_amd.knn_hip(np.zeros((10, 2), dtype=np.float32), 3, order="sorted")


def test_four_million_cells_match_the_oracle_on_sampled_rows(context):
    # Past 2^22 rows: every index computation on the path in 64 bits where it must be
    # (buffers of npad x cap keys, candidate and coordinate offsets).
    n = 4_000_000
    x = synthetic.clustered(n, 50, seed=35)
    print(f"[case] n={n} d=50 k=15 bf16 (sampled rows)", flush=True)
    ids, dist, stats = mrustsc_amd.knn_hip(x, 15, context=context, stats=True)
    print(stats, flush=True)
    queries = np.random.default_rng(2).choice(n, 40, replace=False)
    queries = np.concatenate([queries, [0, n - 1]])
    want_ids, want_dist = oracle.knn(x, 15, queries=queries)
    np.testing.assert_array_equal(ids[queries], want_ids)
    np.testing.assert_array_equal(dist[queries], want_dist)
    assert stats["fallback_rows"] <= n // 1000


@pytest.mark.parametrize(("precision", "n", "blocks", "tail_blocks", "nbin"), [
    ("bf16", 200_000, 782, 270, 35),     # 782 query blocks on 512 workgroup slots
    ("fp32", 1_000_000, 3907, 323, 40),  # (fp32 at 200,000 already cuts every block in 3)
])
@pytest.mark.parametrize("order", ["input", "grouped"])
def test_last_round_blocks_scan_in_ranges_and_match_the_oracle(context, tmp_path, precision, n, blocks,
                                                                tail_blocks, nbin, order):
    # The launch's last round fills more than half of the 512 workgroup slots, so its
    # blocks scan in three reference ranges each. Every row equals a plan without that split
    # (two ranges for every block, forced through the tuning table), and sampled rows, a
    # third of them in the split blocks, equal the oracle.
    x = synthetic.clustered(n, 50, seed=41)
    print(f"[case] n={n} d=50 k=15 {precision} order={order} (last-round split)", flush=True)
    mrustsc_amd.tune_clear()
    ids, dist, stats = mrustsc_amd.knn_hip(x, 15, precision=precision, order=order, context=context, stats=True)
    print(stats, flush=True)
    assert (stats["splits"], stats["tail_blocks"], stats["tail_splits"]) == (1, tail_blocks, 3)
    table = tmp_path / "two_ranges.txt"
    nch = 5 if precision == "bf16" else 13
    table.write_text(f"{0 if precision == 'bf16' else 1} 0 32 {nch} {nbin} 256 2 {1 if order == 'input' else 2} 1.0\n")
    try:
        mrustsc_amd.tune_load(table)
        ids2, dist2, stats2 = mrustsc_amd.knn_hip(x, 15, precision=precision, context=context, stats=True)
    finally:
        mrustsc_amd.tune_clear()
    assert (stats2["splits"], stats2["tail_blocks"], stats2["tuned"]) == (2, 0, 1)
    if precision == "bf16":
        # More blocks than workgroup slots: the split-tile pass runs persistent, with one
        # buffer region per slot, under both plans (range items keep per-range lists).
        assert (stats["persistent"], stats2["persistent"]) == (1, 1)
    np.testing.assert_array_equal(ids, ids2)
    np.testing.assert_array_equal(dist, dist2)
    tail_row0 = (blocks - tail_blocks) * 256
    rng = np.random.default_rng(5)
    queries = np.concatenate([rng.choice(n, 40, replace=False), rng.choice(np.arange(tail_row0, n), 20, replace=False),
                              [0, tail_row0 - 1, tail_row0, n - 1]])
    want_ids, want_dist = oracle.knn(x, 15, queries=queries)
    np.testing.assert_array_equal(ids[queries], want_ids)
    np.testing.assert_array_equal(dist[queries], want_dist)


def test_random_shapes_options_and_data_match_the_oracle(context):
    # 200 seeded cases across n 20-30,000, d 1-129 (both sides of every chunk boundary and of
    # the pass's d <= 128 limit), k 1-64 (both sides of KP = 32 / 64 / 128),
    # both metrics, all three paths, every row order and buffer size, on clustered,
    # integer-lattice (ties) and scaled-normal data; up to 300 rows checked per case.
    rng = np.random.default_rng(2024)
    for case in range(200):
        n = int(np.exp(rng.uniform(np.log(20), np.log(30000))))
        d = int(rng.choice([1, 2, 3, 5, 8, 13, 20, 30, 31, 32, 41, 42, 50, 52, 53, 64, 100, 128, 129]))
        k = int(min(n - 1, rng.choice([1, 2, 5, 14, 15, 16, 30, 47, 48, 64])))
        metric = str(rng.choice(["euclidean", "cosine"]))
        precision = str(rng.choice(["bf16", "bf16", "fp32", "exact"]))
        order = str(rng.choice(["auto", "input", "grouped"]))
        tile = int(rng.choice([0, 0, 128, 256, 512]))
        kind = int(rng.integers(0, 3))
        if kind == 0:
            x = synthetic.clustered(n, d, seed=case, n_clusters=int(rng.integers(2, 12)))
        elif kind == 1:
            x = rng.integers(-3, 4, size=(n, d)).astype(np.float32)
        else:
            x = rng.normal(size=(n, d)).astype(np.float32) * np.float32(rng.choice([1e-3, 1.0, 1e3]))
        tag = f"case {case}: n={n} d={d} k={k} {metric} {precision} order={order} tile={tile} data={kind}"
        print(f"[case] {tag}", flush=True)
        ids, dist = mrustsc_amd.knn_hip(x, k, metric=metric, precision=precision, order=order, tile=tile,
                                        context=context)
        q = np.arange(n) if n <= 3000 else rng.choice(n, 300, replace=False)
        want_ids, want_dist = oracle.knn(x, k, metric=metric, queries=q)
        np.testing.assert_array_equal(ids[q], want_ids, err_msg=tag)
        np.testing.assert_array_equal(dist[q], want_dist, err_msg=tag)


def test_short_lists_of_the_wide_candidate_passes_match_the_oracle(context):
    # n - 1 < KP (128 or 256): every other cell is a candidate and the list's tail is empty.
    # A larger build first leaves the context's candidate slots full of its keys.
    big = synthetic.clustered(3000, 20, seed=9)
    mrustsc_amd.knn_hip(big, 100, context=context)
    for n, k in [(50, 49), (65, 48), (100, 60), (111, 48), (112, 111), (128, 100), (129, 111),
                 (113, 112), (200, 150), (256, 239), (257, 239), (300, 128)]:
        for metric in ("euclidean", "cosine"):
            for precision in ("bf16", "fp32"):
                x = np.random.default_rng(n + k).normal(size=(n, 5)).astype(np.float32)
                tag = f"n={n} k={k} {metric} {precision}"
                print(f"[case] {tag}", flush=True)
                ids, dist = mrustsc_amd.knn_hip(x, k, metric=metric, precision=precision, context=context)
                want_ids, want_dist = oracle.knn(x, k, metric=metric)
                np.testing.assert_array_equal(ids, want_ids, err_msg=tag)
                np.testing.assert_array_equal(dist, want_dist, err_msg=tag)


def test_rows_of_129_to_256_features_take_the_fp32_pass_and_match_the_oracle(context):
    # The packed 16-bit operands hold up to 128 features; single builds of 129-256 take the
    # fp32 pass (49, 57 or 65 chunks) whatever precision was asked, past 256 the exact kernel.
    # Both sides of each chunk boundary, one-round and multi-round builds (one range on the
    # spatial order, with the ball bound), both metrics, k across the candidate tiers.
    cases = [(3000, 129, 15, "euclidean", "bf16"), (5000, 150, 30, "cosine", "fp16"),
             (20000, 192, 15, "euclidean", "fp32"), (9000, 193, 100, "euclidean", "bf16"),
             (30000, 200, 5, "cosine", "bf16"), (7000, 256, 15, "euclidean", "bf16"),
             (4000, 256, 200, "euclidean", "fp32"), (2500, 257, 15, "euclidean", "bf16"),
             (6000, 224, 15, "euclidean", "bf16"), (6000, 225, 30, "cosine", "bf16"),
             (150000, 200, 15, "euclidean", "bf16"), (140000, 256, 30, "cosine", "bf16")]
    rng = np.random.default_rng(129)
    for n, d, k, metric, precision in cases:
        tag = f"n={n} d={d} k={k} {metric} {precision}"
        print(f"[case] {tag}", flush=True)
        x = synthetic.clustered(n, d, seed=d, n_clusters=12)
        ids, dist, st = mrustsc_amd.knn_hip(x, k, metric=metric, precision=precision, context=context,
                                            stats=True)
        assert st["path"] == ("fp32" if d <= 256 else "exact"), tag
        if d <= 256:
            assert st["fallback_rows"] < n // 100 + 8, tag
        if n > 131_072:
            # Multi-round builds on the spatial order skip tile pairs by their ball bound: every
            # row equal to the input-order build (no balls).
            assert st["order"] == 2, tag
            ids_in, dist_in = mrustsc_amd.knn_hip(x, k, metric=metric, precision=precision, order="input",
                                                  context=context)
            np.testing.assert_array_equal(ids, ids_in, err_msg=tag)
            np.testing.assert_array_equal(dist, dist_in, err_msg=tag)
        q = np.arange(n) if n <= 3000 else rng.choice(n, 200, replace=False)
        want_ids, want_dist = oracle.knn(x, k, metric=metric, queries=q)
        np.testing.assert_array_equal(ids[q], want_ids, err_msg=tag)
        np.testing.assert_array_equal(dist[q], want_dist, err_msg=tag)


def test_random_large_k_cases_match_the_oracle(context):
    # 48 seeded cases with k 48-240 (the 128-candidate pass up to k = 111, the 256-candidate
    # pass up to k = 239, the exact kernel above), n 130-20,000 (at least 2 k), both metrics
    # and pass precisions, every row order, on clustered, lattice (ties, wide tie bands) and
    # scaled-normal data.
    rng = np.random.default_rng(4096)
    for case in range(48):
        n = int(np.exp(rng.uniform(np.log(130), np.log(20000))))
        d = int(rng.choice([2, 3, 8, 20, 31, 50, 64, 100, 128]))
        k = int(rng.choice([48, 63, 64, 65, 80, 100, 110, 111, 112, 128, 150, 200, 239, 240]))
        n = max(n, 2 * k)
        metric = str(rng.choice(["euclidean", "cosine"]))
        precision = str(rng.choice(["bf16", "fp32"]))
        order = str(rng.choice(["auto", "input", "grouped"]))
        kind = int(rng.integers(0, 3))
        if kind == 0:
            x = synthetic.clustered(n, d, seed=500 + case, n_clusters=int(rng.integers(2, 12)))
        elif kind == 1:
            x = rng.integers(-2, 3, size=(n, d)).astype(np.float32)
        else:
            x = rng.normal(size=(n, d)).astype(np.float32) * np.float32(rng.choice([1e-3, 1.0, 1e3]))
        tag = f"case {case}: n={n} d={d} k={k} {metric} {precision} order={order} data={kind}"
        print(f"[case] {tag}", flush=True)
        ids, dist, st = mrustsc_amd.knn_hip(x, k, metric=metric, precision=precision, order=order,
                                            context=context, stats=True)
        assert st["candidates"] == (128 if k <= 111 else 256 if k <= 239 else 0), tag
        q = np.arange(n) if n <= 2000 else rng.choice(n, 300, replace=False)
        want_ids, want_dist = oracle.knn(x, k, metric=metric, queries=q)
        np.testing.assert_array_equal(ids[q], want_ids, err_msg=tag)
        np.testing.assert_array_equal(dist[q], want_dist, err_msg=tag)


def test_persistent_pass_equals_per_row_buffers_and_the_oracle(context):
    # 300,000 cells (1,172 query blocks on 512 slots): the persistent pass against per-row
    # buffers (k = 15: buffers of 128 slots, single-tile steps; k = 30, cosine, 64
    # candidates: the fp32 pass), every row equal; sampled rows equal to the oracle.
    n = 300_000
    x = synthetic.clustered(n, 50, seed=61)
    for k, metric in ((15, "euclidean"), (30, "cosine")):
        print(f"[case] n={n} d=50 k={k} {metric} persistent pass vs per-row buffers", flush=True)
        ids, dist, st = mrustsc_amd.knn_hip(x, k, metric=metric, context=context, stats=True)
        if k == 15:
            ids2, dist2, st2 = mrustsc_amd.knn_hip(x, k, metric=metric, context=context, stats=True, tile=128)
        else:
            ids2, dist2, st2 = mrustsc_amd.knn_hip(x, k, metric=metric, precision="fp32", context=context, stats=True)
        print(st, st2, flush=True)
        assert (st["persistent"], st2["persistent"]) == (1, 0)
        assert st["workspace_bytes"] < st2["workspace_bytes"]
        np.testing.assert_array_equal(ids, ids2)
        np.testing.assert_array_equal(dist, dist2)
        queries = np.concatenate([np.random.default_rng(k).choice(n, 24, replace=False), [0, n - 1]])
        want_ids, want_dist = oracle.knn(x, k, metric=metric, queries=queries)
        np.testing.assert_array_equal(ids[queries], want_ids)
        np.testing.assert_array_equal(dist[queries], want_dist)


def test_ball_bound_keeps_every_graph_equal_to_a_scan_of_every_pair(context):
    # Builds of more than one round on the spatial order skip tile pairs their ball bound
    # proves out of reach: the persistent pass (d <= 52) and the register-staged pass of wider
    # bf16 rows. Seeded cases past 131,072 cells, clustered (the bound skips pairs) and without
    # structure (it skips few or none), both metrics, d 10-100, k 1-47: every row equal to the
    # fp32 pass (which keeps no balls and scans every pair), sampled rows equal to the oracle.
    rng = np.random.default_rng(2024)
    cases = [(140_000, 50, 15, "euclidean", "clustered"), (260_000, 17, 5, "cosine", "clustered"),
             (200_000, 32, 30, "euclidean", "normal"), (150_000, 10, 1, "euclidean", "clustered"),
             (520_000, 50, 47, "cosine", "clustered"), (180_000, 52, 15, "euclidean", "duplicates"),
             (300_000, 64, 15, "euclidean", "clustered"), (160_000, 100, 30, "cosine", "clustered"),
             (140_000, 100, 5, "euclidean", "normal")]
    for n, d, k, metric, kind in cases:
        if kind == "clustered":
            x = synthetic.clustered(n, d, seed=int(rng.integers(1 << 30)), n_clusters=int(rng.integers(5, 31)))
        elif kind == "normal":
            x = rng.normal(size=(n, d)).astype(np.float32)
        else:
            x = synthetic.clustered(n, d, seed=int(rng.integers(1 << 30)), n_clusters=8)
            x[rng.choice(n, n // 10, replace=False)] = x[rng.choice(n, n // 10, replace=False)]
        tag = f"n={n} d={d} k={k} {metric} {kind}"
        print(f"[case] {tag}", flush=True)
        ids, dist, st = mrustsc_amd.knn_hip(x, k, metric=metric, context=context, stats=True)
        ids32, dist32, st32 = mrustsc_amd.knn_hip(x, k, metric=metric, precision="fp32", context=context,
                                                  stats=True)
        print(st, flush=True)
        assert st["persistent"] == (1 if d <= 52 else 0), tag
        every = st32["pairs_scored"]
        assert 0 < st["pairs_scored"] <= every, tag
        if kind == "clustered":
            assert st["pairs_scored"] < every, tag
        np.testing.assert_array_equal(ids, ids32, err_msg=tag)
        np.testing.assert_array_equal(dist, dist32, err_msg=tag)
        queries = np.concatenate([rng.choice(n, 24, replace=False), [0, n - 1]])
        want_ids, want_dist = oracle.knn(x, k, metric=metric, queries=queries)
        np.testing.assert_array_equal(ids[queries], want_ids, err_msg=tag)
        np.testing.assert_array_equal(dist[queries], want_dist, err_msg=tag)


def test_persistent_pass_skipped_for_huge_coordinates_lists_every_row(context):
    # A dataset past the pass's norm limit skips the scan; the persistent pass still gives
    # every row an empty candidate list, so every row takes the exact kernel.
    n = 140_000
    x = (synthetic.clustered(n, 10, seed=52, n_clusters=6).astype(np.float64) * 1e16).astype(np.float32)
    print(f"[case] n={n} d=10 k=15 bf16 huge coordinates, persistent pass", flush=True)
    ids, dist, st = mrustsc_amd.knn_hip(x, 15, context=context, stats=True)
    print(st, flush=True)
    assert st["persistent"] == 1 and st["fallback_rows"] == n
    queries = np.concatenate([np.random.default_rng(3).choice(n, 16, replace=False), [0, n - 1]])
    want_ids, want_dist = oracle.knn(x, 15, queries=queries)
    np.testing.assert_array_equal(ids[queries], want_ids)
    np.testing.assert_array_equal(dist[queries], want_dist)


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
@pytest.mark.parametrize("scale", [1e12, 1e15, 1e20])
def test_huge_coordinates_match_the_oracle(context, precision, scale):
    # Past a squared norm of 2^100 (coordinates beyond about 1e15) the MFMA pass is skipped
    # and every row takes the exact kernel; at 1e20 the f32 norms and distances overflow to
    # +inf, which the oracle's f32 arithmetic does the same way.
    print(f"[case] huge coordinates scale={scale:g} precision={precision}", flush=True)
    x = (synthetic.clustered(3000, 20, seed=51, n_clusters=5).astype(np.float64) * scale).astype(np.float32)
    assert np.isfinite(x).all()
    ids, dist, st = mrustsc_amd.knn_hip(x, 15, precision=precision, context=context, stats=True)
    want_ids, want_dist = oracle.knn(x, 15)
    np.testing.assert_array_equal(ids, want_ids)
    np.testing.assert_array_equal(dist, want_dist)
    if scale >= 1e15:
        assert st["fallback_rows"] == 3000
    else:
        assert st["fallback_rows"] == 0
