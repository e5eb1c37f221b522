"""Connectivities (UMAP's fuzzy simplicial set of the kNN graph): the oracle pinned against
the reference's golden vectors and the mathematics, then the device pipeline against the
oracle.

The oracle restates crates/scrust-core/src/neighbors.rs:193-372 with the C library's expf,
as the reference's f32::exp is. The device rounds an f64 exponential to f32, so weights are
held to the reference's own audit bar for two faithful implementations
(tests/test_neighbors_audit.py:16-44 there): identical sparsity pattern, relative error at
most 1e-4, at least 90% of weights bit-identical.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from mrustsc_amd import synthetic

GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_known_answers.json").read_text())
WEIGHT_TOLERANCE = 1e-4
MIN_BITWISE_FRACTION = 0.90


def dense(indptr, cols, vals, n):
    m = np.zeros((n, n), dtype=np.float64)
    for r in range(n):
        for p in range(indptr[r], indptr[r + 1]):
            m[r, cols[p]] = vals[p]
    return m


def generic(n=240, d=10, seed=11):
    return np.random.default_rng(seed).normal(size=(n, d)).astype(np.float32)


# --- the oracle ------------------------------------------------------------------------


def test_oracle_matches_the_reference_scanpy_connectivities():
    case = GOLDEN["scanpy_connectivities"]
    x = synthetic.reference_embedding(300, 20, 42)
    ids, dist = oracle.knn(x, 14)
    indptr, cols, vals, _, _ = oracle.connectivities(ids, dist)
    tol = case["relative_tolerance"]
    for cell, row in zip(case["sample_cells"], case["rows"]):
        ours = dict(zip(cols[indptr[cell]:indptr[cell + 1]].tolist(), vals[indptr[cell]:indptr[cell + 1]].tolist()))
        shared = 0
        for col, expected in row:
            if col in ours:
                shared += 1
                assert abs(ours[col] - expected) <= tol * abs(expected), (cell, col, ours[col], expected)
        assert shared >= 0.9 * len(row)
    assert len(vals) == case["nnz"]
    assert abs(vals.sum(dtype=np.float32) - case["weight_sum"]) <= tol * case["weight_sum"]
    assert abs((vals * vals).sum(dtype=np.float32) - case["weight_square_sum"]) <= tol * case["weight_square_sum"]


def test_oracle_hand_computed_line():
    # Points 0, 1, 3 with k = 1: rho is each row's only distance, every weight is 1, and the
    # union keeps 0-1 and 1-2 in both directions.
    x = np.array([[0.0], [1.0], [3.0]], dtype=np.float32)
    ids, dist = oracle.knn(x, 1)
    assert ids.ravel().tolist() == [1, 0, 1]
    indptr, cols, vals, sigmas, rhos = oracle.connectivities(ids, dist)
    assert indptr.tolist() == [0, 1, 3, 4]
    assert cols.tolist() == [1, 0, 2, 1]
    assert vals.tolist() == [1.0, 1.0, 1.0, 1.0]
    assert rhos.tolist() == [1.0, 1.0, 2.0]


@pytest.mark.parametrize("k", [2, 5, 14, 30])
def test_oracle_fuzzy_set_invariants(k):
    x = generic()
    ids, dist = oracle.knn(x, k)
    indptr, cols, vals, sigmas, rhos = oracle.connectivities(ids, dist)
    n = x.shape[0]
    m = dense(indptr, cols, vals, n)
    np.testing.assert_array_equal(m, m.T)             # fuzzy union is symmetric
    assert np.count_nonzero(np.diag(m)) == 0          # no self edges
    assert vals.min() > 0.0 and vals.max() <= 1.0     # memberships in (0, 1]
    for r in range(n):
        assert np.all(np.diff(cols[indptr[r]:indptr[r + 1]].astype(np.int64)) > 0)  # sorted, unique
    # rho is the first strictly positive distance; sigma solves the cardinality equation
    # (evaluated independently in f64) unless the floor holds it.
    target = np.log2(k + 1)
    d64 = dist.astype(np.float64)
    first_pos = np.array([row[row > 0][0] if np.any(row > 0) else 0.0 for row in dist], dtype=np.float32)
    np.testing.assert_array_equal(rhos, first_pos)
    card = np.where(d64 - rhos[:, None] > 0, np.exp(-(d64 - rhos[:, None]) / sigmas[:, None].astype(np.float64)), 1.0).sum(1)
    floor = 1e-3 * dist.astype(np.float64).sum(1) / (k + 1)
    free = sigmas > floor * (1 + 1e-6)
    assert free.mean() > 0.9
    assert np.abs(card[free] - target).max() < 2e-5
    # Directed weights from sigma and rho reproduce every entry: a + a^T - a*a^T.
    a = np.zeros((n, n))
    for r in range(n):
        off = d64[r] - rhos[r]
        w = np.where(off <= 0, 1.0, np.exp(-off / sigmas[r]))
        a[r, ids[r]] = w
    union = a + a.T - a * a.T
    np.testing.assert_allclose(m, union, rtol=1e-5, atol=1e-7)


def test_oracle_rho_skips_duplicates_and_is_zero_when_all_neighbours_are_duplicates():
    x = generic(200, 6, 9)
    x[:20] = x[0]                 # 20 copies: every neighbour of a copy is at distance 0 (k = 14)
    x[100:105] = x[100]           # 5 copies: rho skips the 4 zero distances
    ids, dist = oracle.knn(x, 14)
    indptr, cols, vals, sigmas, rhos = oracle.connectivities(ids, dist)
    assert np.count_nonzero(rhos == 0.0) == 20
    for r in range(20):
        np.testing.assert_array_equal(vals[indptr[r]:indptr[r + 1]][np.isin(cols[indptr[r]:indptr[r + 1]], ids[r])], 1.0)
    for r in range(100, 105):
        assert rhos[r] > 0.0 and rhos[r] == dist[r][dist[r] > 0][0]


def test_oracle_sigma_floor_uses_the_row_mean_over_k_plus_one():
    # Satellites of a tight cluster see 14 equidistant neighbours: the bisection collapses
    # and only the row floor 1e-3 * sum(row) / (k + 1) keeps sigma finite.
    rng = np.random.default_rng(6)
    x = rng.normal(size=(200, 6)).astype(np.float32)
    x[:20] = x[0]
    x[20:25] = x[0] + (rng.normal(size=(5, 6)) * 1e-2).astype(np.float32)
    ids, dist = oracle.knn(x, 14)
    _, _, _, sigmas, rhos = oracle.connectivities(ids, dist)
    floor = np.float32(1e-3) * (np.array([oracle.ndarray_sum(r) for r in dist], dtype=np.float32) / np.float32(15))
    held = (rhos > 0) & (sigmas == floor)
    assert np.count_nonzero(held) >= 2


def test_ndarray_sum_is_a_sum():
    rng = np.random.default_rng(3)
    for length in (1, 7, 8, 9, 15, 16, 17, 100, 1001):
        v = rng.uniform(0.5, 2.0, size=length).astype(np.float32)
        exact = float(np.sum(v.astype(np.float64)))
        assert abs(oracle.ndarray_sum(v) - exact) <= length * 2 ** -23 * exact


def test_oracle_rejects_invalid_graphs():
    ids = np.zeros((4, 2), dtype=np.uint32)
    ids[0, 0] = 9
    with pytest.raises(ValueError):
        oracle.connectivities(ids, np.ones((4, 2), dtype=np.float32))


# --- the device pipeline ------------------------------------------------------------------

torch = pytest.importorskip("torch")
import mrustsc_amd  # noqa: E402


def assert_same_weights(got, want):
    g_indptr, g_cols, g_vals = got[:3]
    w_indptr, w_cols, w_vals = want[:3]
    np.testing.assert_array_equal(g_indptr, w_indptr)
    np.testing.assert_array_equal(g_cols, w_cols)
    rel = np.abs(g_vals.astype(np.float64) - w_vals) / np.maximum(np.abs(w_vals), 1e-30)
    assert rel.max() <= WEIGHT_TOLERANCE, rel.max()
    bitwise = np.count_nonzero(g_vals == w_vals) / len(w_vals)
    assert bitwise >= MIN_BITWISE_FRACTION, bitwise
    return bitwise


def test_device_rejects_invalid_graphs_before_touching_a_device():
    with pytest.raises(ValueError):
        mrustsc_amd.connectivities(np.zeros((0, 3), dtype=np.uint32), np.zeros((0, 3), dtype=np.float32))
    with pytest.raises(ValueError):
        mrustsc_amd.connectivities(np.zeros((3, 2), dtype=np.int64), np.zeros((3, 2), dtype=np.float32))
    with pytest.raises(ValueError):
        mrustsc_amd.connectivities(np.zeros((3, 2), dtype=np.uint32), np.zeros((3, 3), dtype=np.float32))


@pytest.mark.gpu
def test_device_matches_the_oracle_on_the_reference_graph(context):
    x = synthetic.reference_embedding(300, 20, 42)
    ids, dist = oracle.knn(x, 14)
    print("[case] connectivities reference graph", flush=True)
    got = mrustsc_amd.connectivities(ids, dist, context=context)
    assert got[3] == 300
    assert len(got[2]) == GOLDEN["scanpy_connectivities"]["nnz"]
    assert_same_weights(got, oracle.connectivities(ids, dist))


@pytest.mark.gpu
@pytest.mark.parametrize("k", [1, 2, 5, 14, 30])
def test_device_matches_the_oracle_on_generic_graphs(context, k):
    x = generic()
    ids, dist = oracle.knn(x, k)
    print(f"[case] connectivities generic k={k}", flush=True)
    assert_same_weights(mrustsc_amd.connectivities(ids, dist, context=context), oracle.connectivities(ids, dist))


@pytest.mark.gpu
def test_device_matches_the_oracle_with_duplicates_and_floors(context):
    x = generic(200, 6, 9)
    x[:20] = x[0]
    x[100:105] = x[100]
    rng = np.random.default_rng(6)
    y = rng.normal(size=(200, 6)).astype(np.float32)
    y[:20] = y[0]
    y[20:25] = y[0] + (rng.normal(size=(5, 6)) * 1e-2).astype(np.float32)
    for data in (x, y):
        ids, dist = oracle.knn(data, 14)
        print("[case] connectivities duplicates", flush=True)
        assert_same_weights(mrustsc_amd.connectivities(ids, dist, context=context), oracle.connectivities(ids, dist))


@pytest.mark.gpu
def test_device_matches_the_oracle_on_a_large_mixture(context):
    x = synthetic.clustered(30000, 50, seed=12)
    ids, dist = mrustsc_amd.knn_hip(x, 15, context=context)
    print("[case] connectivities 30000 x 15", flush=True)
    bitwise = assert_same_weights(mrustsc_amd.connectivities(ids, dist, context=context),
                                  oracle.connectivities(ids, dist))
    print("bitwise fraction", bitwise, flush=True)


@pytest.mark.gpu
def test_device_entry_point_matches_the_host_entry_point(context):
    x = synthetic.clustered(5000, 30, seed=2)
    ids, dist = mrustsc_amd.knn_hip(x, 15, context=context)
    indptr, cols, vals, _ = mrustsc_amd.connectivities(ids, dist, context=context)
    n, k = ids.shape
    d_ids = torch.from_numpy(ids.view(np.int32)).to("cuda")
    d_dist = torch.from_numpy(dist).to("cuda")
    d_indptr = torch.empty(n + 1, dtype=torch.int32, device="cuda")
    d_cols = torch.empty(2 * n * k, dtype=torch.int32, device="cuda")
    d_vals = torch.empty(2 * n * k, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    nnz = mrustsc_amd.connectivities_device(d_ids.data_ptr(), d_dist.data_ptr(), n, k, d_indptr.data_ptr(),
                                            d_cols.data_ptr(), d_vals.data_ptr(), 2 * n * k, context=context)
    assert nnz == len(vals)
    np.testing.assert_array_equal(d_indptr.cpu().numpy().view(np.uint32), indptr)
    np.testing.assert_array_equal(d_cols.cpu().numpy()[:nnz].view(np.uint32), cols)
    np.testing.assert_array_equal(d_vals.cpu().numpy()[:nnz], vals)


@pytest.mark.gpu
def test_device_rejects_bad_ids_and_distances(context):
    ids = np.tile(np.arange(1, 4, dtype=np.uint32), (5, 1)) % 5
    dist = np.ones((5, 3), dtype=np.float32)
    bad = ids.copy()
    bad[2, 1] = 7
    with pytest.raises(ValueError, match="neighbour ids"):
        mrustsc_amd.connectivities(bad, dist, context=context)
    neg = dist.copy()
    neg[1, 0] = -1.0
    with pytest.raises(ValueError):
        mrustsc_amd.connectivities(ids, neg, context=context)
    mrustsc_amd.connectivities(ids, dist, context=context)  # the context still works


@pytest.mark.gpu
def test_device_rejects_a_repeated_neighbour_id(context):
    # The reference would emit the column once per repeat (neighbors.rs:340-370); a kNN graph
    # never repeats an id in a row, and the device refuses one rather than fold it.
    ids = np.array([[1, 2], [0, 2], [0, 1], [1, 1]], dtype=np.uint32)
    dist = np.ones((4, 2), dtype=np.float32)
    with pytest.raises(ValueError, match="distinct neighbour ids"):
        mrustsc_amd.connectivities(ids, dist, context=context)
    mrustsc_amd.connectivities(ids[:3], dist[:3], context=context)


def test_neighbors_rejects_fewer_than_two_neighbours():
    with pytest.raises(ValueError, match="at least 2"):
        mrustsc_amd.neighbors(generic(), n_neighbors=1)


@pytest.mark.gpu
def test_neighbors_equals_knn_then_connectivities(context):
    # The pp.neighbors device call: the graph (n_neighbors - 1 columns) and its
    # connectivities, as the two separate calls give them.
    x = synthetic.clustered(8000, 50, seed=21)
    print("[case] neighbors 8000 x 50", flush=True)
    out = mrustsc_amd.neighbors(x, n_neighbors=15, context=context)
    ids, dist = mrustsc_amd.knn_hip(x, 14, context=context)
    indptr, cols, vals, _ = mrustsc_amd.connectivities(ids, dist, context=context)
    d_indptr, d_cols, d_vals, shape = out["obsp"]["distances"]
    assert shape == (8000, 8000)
    np.testing.assert_array_equal(d_indptr, np.arange(0, 8000 * 14 + 1, 14))
    np.testing.assert_array_equal(d_cols, ids.ravel())
    np.testing.assert_array_equal(d_vals, dist.ravel())
    c_indptr, c_cols, c_vals, _ = out["obsp"]["connectivities"]
    np.testing.assert_array_equal(c_indptr, indptr)
    np.testing.assert_array_equal(c_cols, cols)
    np.testing.assert_array_equal(c_vals, vals)
    assert out["uns"]["neighbors"]["params"] == {"n_neighbors": 15, "method": "umap", "use_rep": "X_pca"}


@pytest.mark.gpu
def test_device_resident_neighbors_equal_the_host_call(context):
    # The embedding stays on the device (a PCA output kept there) and so do both graphs.
    import torch

    x = synthetic.clustered(6000, 50, seed=23)
    n, k = 6000, 14
    print("[case] neighbors_device 6000 x 50", flush=True)
    host = mrustsc_amd.neighbors(x, n_neighbors=15, context=context)
    dx = torch.from_numpy(x).to("cuda")
    ids = torch.empty((n, k), dtype=torch.int32, device="cuda")
    dist = torch.empty((n, k), dtype=torch.float32, device="cuda")
    cap = 2 * n * k
    indptr = torch.empty(n + 1, dtype=torch.int32, device="cuda")
    cols = torch.empty(cap, dtype=torch.int32, device="cuda")
    vals = torch.empty(cap, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    nnz = mrustsc_amd.neighbors_device(dx.data_ptr(), n, 50, 15, ids.data_ptr(), dist.data_ptr(),
                                       indptr.data_ptr(), cols.data_ptr(), vals.data_ptr(), cap,
                                       context=context)
    d_indptr, d_cols, d_vals, _ = host["obsp"]["distances"]
    np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32).ravel(), d_cols)
    np.testing.assert_array_equal(dist.cpu().numpy().ravel(), d_vals)
    c_indptr, c_cols, c_vals, _ = host["obsp"]["connectivities"]
    assert nnz == len(c_cols)
    np.testing.assert_array_equal(indptr.cpu().numpy().view(np.uint32), c_indptr)
    np.testing.assert_array_equal(cols.cpu().numpy()[:nnz].view(np.uint32), c_cols)
    np.testing.assert_array_equal(vals.cpu().numpy()[:nnz], c_vals)


def test_device_resident_neighbors_checks_arguments_first():
    with pytest.raises(ValueError, match="at least 2"):
        mrustsc_amd.neighbors_device(0, 10, 5, 1, 0, 0, 0, 0, 0, 0)
