"""Pins the CPU oracle (oracle/knn_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the oracle fails it:
  * exact column sums against Python's correctly rounded ``math.fsum``;
  * every squared distance against an exact-rational restatement of the f32 fma chain;
  * the selection against ``sorted`` on (d^2, id), self excluded;
  * the reference's own known answers (neighbors.rs:408-428, knn.rs:361-388,
    tests/test_device_parity.py, tests/test_neighbors_audit.py) and scanpy sets
    (neighbors.rs:493-561);
  * set equality with an fp64 brute force on generic data;
  * the fp64 boundary rule on a constructed near tie, in both directions.
"""

from __future__ import annotations

import json
import math
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

import oracle
from mrustsc_amd import synthetic

GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_known_answers.json").read_text())


# --- exact arithmetic helpers (independent of the C code) ---------------------------------


def round_f32(value: Fraction) -> float:
    """Correctly rounded (nearest, ties to even) f32 value of an exact rational."""
    if value == 0:
        return 0.0
    sign = -1 if value < 0 else 1
    v = abs(value)
    e = math.floor(math.log2(v.numerator) - math.log2(v.denominator))
    while Fraction(2) ** e > v:
        e -= 1
    while Fraction(2) ** (e + 1) <= v:
        e += 1
    e = max(e, -126)  # subnormal range keeps the 2^-149 quantum
    quantum = Fraction(2) ** (e - 23)
    q, r = divmod(v, quantum)
    if r > quantum / 2 or (r == quantum / 2 and q % 2 == 1):
        q += 1
    return float(np.float32(sign * float(q * quantum)))


def fmaf(a: float, b: float, c: float) -> float:
    return round_f32(Fraction(a) * Fraction(b) + Fraction(c))


def direct_square(a, b, na, nb) -> float:
    s = 0.0
    for x, y in zip(a, b):
        delta = float(np.float32(x) - np.float32(y))
        s = fmaf(delta, delta, s)
    d = len(a)
    threshold = np.float32(np.float32(np.float32(d) + np.float32(2.0)) * np.float32(2.0 ** -23)) \
        * np.float32(np.float32(na) + np.float32(nb))
    return 0.0 if s < float(threshold) else s


# --- exact sums / centring ------------------------------------------------------------------


def test_exact_sum_is_the_correctly_rounded_sum():
    rng = np.random.default_rng(0)
    for trial in range(300):
        n = int(rng.integers(1, 400))
        v = (rng.normal(size=n) * 10.0 ** rng.integers(-40, 38)).astype(np.float32)
        if trial % 3 == 0:
            v = np.concatenate([v, -v[: n // 2], np.float32([3e38, -3e38, 1e-45, -2e-45])])
        v = v.astype(np.float32)
        assert oracle.exact_sum(v) == math.fsum(v.astype(np.float64)), trial


def test_column_means_match_fsum_and_the_sequential_sum_when_that_is_exact():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1000, 7)).astype(np.float32)
    means = oracle.column_means(x)
    for j in range(7):
        assert means[j] == math.fsum(x[:, j].astype(np.float64)) / 1000.0
    # Small integers: the reference's sequential f64 sum is exact, so the two agree.
    xi = rng.integers(-1000, 1000, size=(500, 5)).astype(np.float32)
    seq = np.zeros(5)
    for row in xi.astype(np.float64):
        seq += row
    np.testing.assert_array_equal(oracle.column_means(xi), seq / 500.0)


def test_prepare_centres_once_and_forms_fma_norms():
    x = synthetic.random_embedding(12, 5, 3) * np.float32(7.0) + np.float32(100.0)
    coords, norms = oracle.prepare(x)
    means = np.array([math.fsum(x[:, j].astype(np.float64)) / 12.0 for j in range(5)])
    np.testing.assert_array_equal(coords, (x.astype(np.float64) - means).astype(np.float32))
    for r in range(12):
        s = 0.0
        for c in coords[r]:
            s = fmaf(float(c), float(c), s)
        assert norms[r] == np.float32(s)


def test_cosine_prepare_scales_rows_to_unit_f64_norm():
    x = synthetic.random_embedding(20, 6, 5) - np.float32(0.5)
    x[3] = 0.0
    coords, _ = oracle.prepare(x, "cosine")
    length = np.sqrt((x.astype(np.float64) ** 2).sum(axis=1))
    want = np.where(length[:, None] > 0, x / np.maximum(length, 1e-300)[:, None], 0.0)
    np.testing.assert_allclose(coords, want.astype(np.float32), rtol=0, atol=1e-7)
    assert not coords[3].any()


# --- distances and selection against exact restatements -----------------------------------


@pytest.mark.parametrize(("n", "d", "k", "seed"), [(24, 3, 5, 1), (30, 7, 4, 2), (17, 1, 16, 3)])
def test_every_distance_and_the_order_match_an_exact_restatement(n, d, k, seed):
    x = (synthetic.random_embedding(n, d, seed) * np.float32(4.0)).astype(np.float32)
    x[5] = x[2]  # one exact duplicate pair
    coords, norms = oracle.prepare(x)
    ids, dist = oracle.knn(x, k, tie_rule=False)
    for q in range(n):
        keys = sorted((direct_square(coords[q], coords[r], norms[q], norms[r]), r)
                      for r in range(n) if r != q)
        assert ids[q].tolist() == [r for _, r in keys[:k]], q
        assert dist[q].tolist() == [float(np.sqrt(np.float32(s))) for s, _ in keys[:k]], q


def test_fp64_truth_agrees_on_generic_data():
    x = np.random.default_rng(11).normal(size=(240, 10)).astype(np.float32)
    ids, dist = oracle.knn(x, 14)
    t_ids, t_dist = oracle.knn_fp64(x, 14)
    for q in range(240):
        assert set(ids[q].tolist()) == set(t_ids[q].tolist())
    np.testing.assert_allclose(dist, t_dist, rtol=1e-5, atol=1e-6)
    # and the fp64 truth is itself a brute force
    x64 = x.astype(np.float64)
    full = np.sqrt(((x64[:, None, :] - x64[None, :, :]) ** 2).sum(-1))
    for q in (0, 17, 239):
        order = sorted((j for j in range(240) if j != q), key=lambda j: (full[q, j], j))[:14]
        assert t_ids[q].tolist() == order


# --- the fp64 boundary rule -----------------------------------------------------------------


def near_tie():
    """Query 0 at the origin; ids 1 and 2 are at f32 distance exactly 1 from it, while in
    f64 id 2 (offset 1e-5) is nearer than id 1 (offset 1e-4). Columns sum to zero exactly,
    so centring leaves every coordinate unchanged."""
    a, b = np.float32(1e-4), np.float32(1e-5)
    return np.array([[0, 0], [1, a], [1, b], [-51, -a], [49, -b]], dtype=np.float32)


def test_boundary_near_ties_are_resolved_in_fp64():
    x = near_tie()
    ids, dist = oracle.knn(x, 1)
    assert ids[0].tolist() == [2]
    assert dist[0].tolist() == [1.0]
    plain, _ = oracle.knn(x, 1, tie_rule=False)
    assert plain[0].tolist() == [1]  # (d^2, id) order alone keeps the smaller id
    both, _ = oracle.knn(x, 2)
    assert both[0].tolist() == [1, 2]  # a band inside the first k is not reordered


# --- the reference's own known answers -------------------------------------------------------


def test_neighbours_on_a_line_and_a_grid():
    ids, dist = oracle.knn(synthetic.line(6), 2)
    for row, want in GOLDEN["line6_k2"]["rows"].items():
        assert ids[int(row)].tolist() == want
    assert dist[0].tolist() == GOLDEN["line6_k2"]["distances"]["0"]
    ids, dist = oracle.knn(synthetic.grid(3), 4)
    for row, want in GOLDEN["grid3_k4"]["rows"].items():
        assert ids[int(row)].tolist() == want
    np.testing.assert_allclose(dist[0], GOLDEN["grid3_k4"]["corner_distances"], atol=1e-6)
    ids, dist = oracle.knn(synthetic.line(50), 2)
    for row, want in GOLDEN["line50_k2"]["rows"].items():
        assert ids[int(row)].tolist() == want
    x = np.zeros((41, 2), dtype=np.float32)
    x[40, 0] = 100.0
    ids, _ = oracle.knn(x, 3)
    for row, want in GOLDEN["origin40_k3"]["rows"].items():
        assert ids[int(row)].tolist() == want


def test_matches_scanpy_neighbour_sets():
    case = GOLDEN["scanpy"]
    x = synthetic.reference_embedding(case["n_cells"], case["n_dims"], case["seed"])
    ids, _ = oracle.knn(x, case["k"])
    for cell, want in zip(case["sample_cells"], case["neighbors"]):
        shared = len(set(ids[cell].tolist()) & set(want))
        assert shared * 10 >= case["k"] * 9, f"cell {cell}: only {shared} shared"


def test_duplicates_are_at_distance_zero():
    x = np.random.default_rng(5).normal(size=(180, 6)).astype(np.float32)
    x[20:30] = x[20]
    _, dist = oracle.knn(x, 14)
    assert np.count_nonzero(dist == 0.0) == 10 * 9  # tests/test_neighbors_audit.py:194-203
    x = np.random.default_rng(9).normal(size=(200, 6)).astype(np.float32)
    x[:20] = x[0]
    _, dist = oracle.knn(x, 14)
    assert np.count_nonzero(dist[:20]) == 0


def test_lattice_ties_break_towards_the_smaller_index():
    lat = synthetic.lattice(8)
    ids, dist = oracle.knn(lat, 14)
    _, exact = oracle.knn_fp64(lat, 14)
    np.testing.assert_allclose(dist, exact, atol=1e-6)
    for cell in range(len(lat)):
        for a in range(13):
            if np.isclose(dist[cell, a], dist[cell, a + 1], atol=1e-6):
                assert ids[cell, a] < ids[cell, a + 1]


@pytest.mark.parametrize("offset", [0.0, 1e2, 1e4, 1e5])
def test_distances_survive_an_embedding_far_from_the_origin(offset):
    x = (np.random.default_rng(2).normal(size=(300, 12)) + offset).astype(np.float32)
    ids, dist = oracle.knn(x, 14)
    want_ids, want_dist = oracle.knn_fp64(x, 14)
    for cell in range(300):
        assert set(ids[cell].tolist()) == set(want_ids[cell].tolist()), f"cell {cell}"
    np.testing.assert_allclose(dist, want_dist, rtol=1e-4, atol=1e-4 * max(1.0, offset * 1e-4))


def test_rejects_degenerate_arguments():
    with pytest.raises(ValueError):
        oracle.knn(synthetic.line(6), 0)
    with pytest.raises(ValueError):
        oracle.knn(synthetic.line(6), 6)
    with pytest.raises(ValueError):
        oracle.knn(np.zeros((5, 0), np.float32), 3)


def test_query_subsets_equal_the_full_rows():
    x = synthetic.clustered(900, 20, seed=3)
    ids, dist = oracle.knn(x, 15)
    q = np.array([5, 17, 899, 0], dtype=np.uint32)
    s_ids, s_dist = oracle.knn(x, 15, queries=q)
    np.testing.assert_array_equal(s_ids, ids[q])
    np.testing.assert_array_equal(s_dist, dist[q])


def test_cosine_oracle_against_fp64():
    x = synthetic.clustered(500, 16, seed=8)
    ids, dist = oracle.knn(x, 10, metric="cosine")
    t_ids, t_dist = oracle.knn_fp64(x, 10, metric="cosine")
    agree = np.mean([len(set(a) & set(b)) / 10 for a, b in zip(ids.tolist(), t_ids.tolist())])
    assert agree > 0.999
    np.testing.assert_allclose(dist, t_dist, atol=2e-6)


# --- the timed CPU baseline (the reference's expansion path) --------------------------------


def test_baseline_tiling_does_not_change_the_result():
    x = synthetic.reference_embedding(200, 8, 7)  # neighbors.rs:430-437
    a = oracle.baseline_knn(x, 10, tile_rows=200)
    b = oracle.baseline_knn(x, 10, tile_rows=7)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_baseline_agrees_with_the_oracle_away_from_near_ties():
    ids, _ = oracle.baseline_knn(synthetic.line(6), 2)
    assert ids[2].tolist() == [1, 3]
    x = np.random.default_rng(3).normal(size=(400, 20)).astype(np.float32)
    b_ids, b_dist = oracle.baseline_knn(x, 14)
    o_ids, o_dist = oracle.knn(x, 14)
    same = np.mean([set(a) == set(b) for a, b in zip(b_ids.tolist(), o_ids.tolist())])
    assert same > 0.99
    np.testing.assert_allclose(b_dist, o_dist, rtol=1e-4, atol=1e-5)


def test_a_band_wider_than_the_cap_keeps_the_plain_order():
    # Under cosine every cell is a near-tie seen from a zero row: the band is far wider
    # than 64 members, so that row keeps the (d^2, id) order.
    x = synthetic.clustered(300, 8, seed=4)
    x[7] = 0.0
    ids, _ = oracle.knn(x, 5, metric="cosine")
    plain, _ = oracle.knn(x, 5, metric="cosine", tie_rule=False)
    np.testing.assert_array_equal(ids[7], plain[7])
    assert ids[7].tolist() == sorted(ids[7].tolist())  # all at d^2 = |y|^2 ~ 1: id order


def test_oracle_orders_overflowed_distances_by_id():
    # Coordinates near 1e21: every f32 squared distance overflows to +inf except between
    # exact copies (0), so each row lists its copies first, then the smallest other ids, all
    # at +inf: a tie of infinities is not re-ordered by the fp64 boundary rule.
    x = (synthetic.clustered(400, 6, seed=5, n_clusters=3, duplicate_fraction=0.02)
         .astype(np.float64) * 1e20).astype(np.float32)
    k = 5
    ids, dist = oracle.knn(x, k)
    for q in range(x.shape[0]):
        twins = [r for r in range(x.shape[0]) if r != q and np.array_equal(x[r], x[q])]
        others = [r for r in range(x.shape[0]) if r != q and r not in twins]
        want = (twins + others)[:k]
        assert ids[q].tolist() == want, q
        assert all(dist[q][j] == (0.0 if want[j] in twins else np.inf) for j in range(k))


@pytest.mark.parametrize("case", ["clustered", "cosine", "lattice", "duplicates", "tiny", "k_large"])
def test_vectorised_oracle_equals_the_scalar_oracle(case):
    # oracle.knn_fast runs 16 references per vector and 16 rows per block; each pair's f32
    # difference / fma chain / snap and the kept-key selection are the scalar oracle's, so
    # every id and distance bit must agree, ties and bands included.
    metric, k = "euclidean", 15
    if case == "clustered":
        x = synthetic.clustered(3000, 50, seed=3)
    elif case == "cosine":
        x, metric, k = synthetic.clustered(2000, 7, seed=4), "cosine", 30
    elif case == "lattice":  # heavy exact ties and tie bands
        x = np.floor(np.random.default_rng(1).uniform(0, 4, (517, 3))).astype(np.float32)
        k = 5
    elif case == "duplicates":
        x = synthetic.clustered(1500, 20, seed=5, duplicate_fraction=0.2)
    elif case == "tiny":  # n = k + 1, fewer rows than a block, a ragged last block
        x, k = synthetic.clustered(21, 4, seed=6), 20
    else:
        x, k = synthetic.clustered(700, 9, seed=7), 100
    want = oracle.knn(x, k, metric=metric)
    got = oracle.knn_fast(x, k, metric=metric)
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_array_equal(got[1].view(np.uint32), want[1].view(np.uint32))
    rows = np.array([0, 3, len(x) - 1], dtype=np.uint32)
    sub = oracle.knn_fast(x, k, metric=metric, queries=rows)
    np.testing.assert_array_equal(sub[0], want[0][rows])
    np.testing.assert_array_equal(sub[1], want[1][rows])


@pytest.mark.parametrize("name,n,metric", [("c2", 100_000, "euclidean"), ("c3", 1_000_000, "euclidean"),
                                           ("c3_cosine", 1_000_000, "cosine")])
def test_workload_fixtures_hold_the_scalar_oracles_rows(name, n, metric):
    # The whole-graph fixtures were written with the vectorised oracle; their rows in full
    # must be the scalar oracle's, on the embedding the generator reproduces.
    import hashlib

    fix = json.loads((Path(__file__).parent / "golden" / f"{name}_oracle_fixture.json").read_text())
    x = synthetic.clustered(n, 50, seed=1)
    assert hashlib.sha256(x.tobytes()).hexdigest() == fix["x_sha256"]
    rows = np.array(sorted(int(r) for r in fix["rows"]), dtype=np.uint32)
    ids, dist = oracle.knn(x, fix["k"], metric=metric, queries=rows)
    for i, r in enumerate(rows):
        row = fix["rows"][str(r)]
        assert ids[i].tolist() == row["ids"]
        assert [f"{b:08x}" for b in dist[i].view(np.uint32)] == row["dist_bits"]


def test_large_k_fixture_holds_the_scalar_oracles_rows():
    # The C2 embedding at k = 64, 100, 111 and 128: rows in full against the scalar oracle.
    import hashlib

    fix = json.loads((Path(__file__).parent / "golden" / "c2_large_k_oracle_fixture.json").read_text())
    x = synthetic.clustered(100_000, 50, seed=1)
    assert sorted(int(k) for k in fix["graphs"]) == [64, 100, 111, 128]
    for k, rec in fix["graphs"].items():
        assert rec["k"] == int(k) and hashlib.sha256(x.tobytes()).hexdigest() == rec["x_sha256"]
        rows = np.array(sorted(int(r) for r in rec["rows"]), dtype=np.uint32)
        ids, dist = oracle.knn(x, int(k), queries=rows)
        for i, r in enumerate(rows):
            assert ids[i].tolist() == rec["rows"][str(r)]["ids"], (k, r)
            assert [f"{b:08x}" for b in dist[i].view(np.uint32)] == rec["rows"][str(r)]["dist_bits"], (k, r)


def test_batch_fixture_holds_the_scalar_oracles_rows():
    import hashlib

    fix = json.loads((Path(__file__).parent / "golden" / "c4_oracle_fixture.json").read_text())
    specs = synthetic.batch_specs(64, seed=7)
    assert [rec["spec"] for rec in fix["datasets"]] == specs
    for spec, rec in zip(specs[::8], fix["datasets"][::8]):
        x = synthetic.batch_dataset(spec)
        assert hashlib.sha256(x.tobytes()).hexdigest() == rec["x_sha256"]
        rows = np.array(sorted(int(r) for r in rec["rows"]), dtype=np.uint32)
        ids, dist = oracle.knn(x, spec["k"], queries=rows)
        for i, r in enumerate(rows):
            assert ids[i].tolist() == rec["rows"][str(r)]["ids"]
            assert [f"{b:08x}" for b in dist[i].view(np.uint32)] == rec["rows"][str(r)]["dist_bits"]
