"""PCA by randomised SVD (crates/scrust-core/src/pca.rs:74-215): the oracle (its steps in
C with f64 algebra) pinned by what the mathematics fixes, and the device PCA (f32 on the
GPU, the embedding left on the device) against the oracle, then feeding the kNN graph.

What pins the oracle: an exact fp64 SVD of the explicitly centred dense matrix (singular
values, loadings, scores up to the sign convention) on spectra with a clear gap at the
truncation; the sign convention of pca.rs:471-490; both explained-variance statistics of
pca.rs:192-208 and the ratio's denominator of pca.rs:226-255 (tests/test_pca_audit.py:337-
445 in the reference state the same facts); the Jacobi eigensolver against numpy's eigh."""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

import oracle


def known_spectrum(n_cells=400, n_genes=300, n_comps=20, ratio=1e2, seed=0):
    """Singular values geometric from 100 to 100/ratio over the first n_comps, then a tail
    decaying by 0.5 per index (a clear gap past the truncation)."""
    rng = np.random.default_rng(seed)
    rank = min(n_cells, n_genes)
    s = np.empty(rank)
    s[:n_comps] = 100.0 * np.exp(-np.log(ratio) * np.arange(n_comps) / (n_comps - 1))
    s[n_comps:] = s[n_comps - 1] * 0.5 ** np.arange(1, rank - n_comps + 1)
    u, _ = np.linalg.qr(rng.standard_normal((n_cells, n_cells)))
    v, _ = np.linalg.qr(rng.standard_normal((n_genes, n_genes)))
    return ((u[:, :rank] * s) @ v[:, :rank].T).astype(np.float32)


def sparse_counts(n_cells=500, n_genes=120, density=0.3, seed=1):
    rng = np.random.default_rng(seed)
    dense = (rng.random((n_cells, n_genes)) < density) * rng.integers(1, 30, (n_cells, n_genes))
    # A low-rank signal on top, so the leading components are well defined.
    dense = dense + np.outer(rng.integers(0, 5, n_cells), rng.integers(0, 8, n_genes))
    return dense.astype(np.float32)


def csr(x):
    m = sp.csr_matrix(x)
    return m.indptr.astype(np.uint32), m.indices.astype(np.uint32), m.data.astype(np.float32)


def exact(x, zero_center=True):
    x64 = x.astype(np.float64)
    c = x64 - x64.mean(0) if zero_center else x64
    u, s, vt = np.linalg.svd(c, full_matrices=False)
    return u, s, vt


def test_jacobi_matches_numpy_eigh():
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 40):
        a = rng.standard_normal((n, n))
        a = a + a.T
        vals, vecs = oracle.jacobi_eigen(a)
        want = np.linalg.eigvalsh(a)
        np.testing.assert_allclose(np.sort(vals), want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
        np.testing.assert_allclose(vecs.T @ vecs, np.eye(n), atol=1e-12)
        np.testing.assert_allclose(a @ vecs, vecs * vals, atol=1e-10 * np.abs(want).max())


def test_sketch_is_deterministic_standard_normal():
    a = oracle.pca_sketch(300, 74, 0)
    assert np.array_equal(a, oracle.pca_sketch(300, 74, 0))
    assert not np.array_equal(a, oracle.pca_sketch(300, 74, 1))
    assert abs(a.mean()) < 0.02 and abs(a.std() - 1.0) < 0.02
    # Box-Muller pairs: consecutive values share a radius.
    flat = a.ravel().astype(np.float64)
    r2 = flat[0::2] ** 2 + flat[1::2] ** 2
    assert abs(r2.mean() - 2.0) < 0.05  # E[chi^2_2] = 2


@pytest.mark.parametrize("n_comps", [5, 20])
def test_oracle_matches_an_exact_svd_of_the_centred_matrix(n_comps):
    x = known_spectrum(n_comps=20)
    res = oracle.pca(*csr(x), x.shape[1], n_comps)
    u, s, vt = exact(x)
    n = x.shape[0]
    np.testing.assert_allclose(res["explained_variance"], (s[:n_comps] ** 2 / (n - 1)).astype(np.float32), rtol=2e-6)
    for c in range(n_comps):
        cos = abs(float(res["components"][c].astype(np.float64) @ vt[c]))
        assert cos > 1 - 1e-6, (c, cos)
        sign = np.sign(res["components"][c].astype(np.float64) @ vt[c])
        np.testing.assert_allclose(res["embedding"][:, c], (sign * u[:, c] * s[c]).astype(np.float32),
                                   atol=2e-5 * s[0])
    np.testing.assert_allclose(res["components"] @ res["components"].T, np.eye(n_comps), atol=1e-5)


def test_sign_convention_and_variance_ratio_denominators():
    x = sparse_counts()
    n = x.shape[0]
    for zero_center, ddof in ((True, 1), (False, 0)):
        res = oracle.pca(*csr(x), x.shape[1], 8, zero_center=zero_center)
        comp = res["components"]
        for c in range(8):
            assert comp[c][np.argmax(np.abs(comp[c]))] > 0
        total = x.astype(np.float64).var(axis=0, ddof=ddof).sum()
        np.testing.assert_allclose(res["explained_variance_ratio"],
                                   (res["explained_variance"] / total).astype(np.float32), rtol=1e-6)
        if zero_center:
            # The leading component is well separated and exact; the noise components are
            # near-degenerate, where a randomised SVD is within a fraction of a percent.
            _, s, _ = exact(x)
            want = s[:8] ** 2 / (n - 1)
            np.testing.assert_allclose(res["explained_variance"][0], want[0], rtol=1e-5)
            np.testing.assert_allclose(res["explained_variance"], want, rtol=1e-2)
            assert np.all(np.diff(res["explained_variance"]) <= 0)
        else:
            # Uncentred: TruncatedSVD's statistic, the ddof = 0 variance of each score column.
            np.testing.assert_allclose(res["explained_variance"],
                                       res["embedding"].astype(np.float64).var(axis=0).astype(np.float32), rtol=1e-5)
            _, s, vt = exact(x, zero_center=False)
            assert abs(float(res["components"][0].astype(np.float64) @ vt[0])) > 1 - 1e-6


def test_oracle_rejects_what_the_reference_rejects():
    x = sparse_counts(20, 10)
    args = csr(x)
    for n_comps in (0, 11):
        with pytest.raises(ValueError):
            oracle.pca(*args, 10, n_comps)
    one = sparse_counts(1, 10)
    with pytest.raises(ValueError):
        oracle.pca(*csr(one), 10, 1)


def test_device_pca_checks_arguments_before_the_device():
    import mrustsc_amd

    x = sparse_counts(20, 10)
    ip, ix, v = csr(x)
    for n_comps, match in ((0, "at least 1"), (11, "at most min")):
        with pytest.raises(ValueError, match=match):
            mrustsc_amd.pca(ip, ix, v, 10, n_comps)
    one = csr(sparse_counts(1, 10))
    with pytest.raises(ValueError, match="n_cells"):
        mrustsc_amd.pca(*one, 10, 1)
    big = csr(sparse_counts(300, 200))
    with pytest.raises(ValueError, match="104"):
        mrustsc_amd.pca(*big, 200, 105)
    bad = ip.copy()
    bad[3], bad[4] = bad[4], bad[3]
    with pytest.raises(ValueError, match="shape mismatch"):
        mrustsc_amd.pca(bad, ix, v, 10, 3)


# --------------------------------------------------------------------------- device

def _device_vs_oracle(x, n_comps, zero_center, context):
    import mrustsc_amd

    ip, ix, v = csr(x)
    got = mrustsc_amd.pca(ip, ix, v, x.shape[1], n_comps, zero_center, 0, context=context)
    want = oracle.pca(ip, ix, v, x.shape[1], n_comps, zero_center=zero_center)
    return got, want


@pytest.mark.gpu
@pytest.mark.parametrize("n_comps", [5, 20])
def test_device_pca_matches_the_oracle_on_a_clear_spectrum(context, n_comps):
    print(f"[case] device PCA known spectrum n_comps={n_comps}", flush=True)
    x = known_spectrum(n_comps=20)
    got, want = _device_vs_oracle(x, n_comps, True, context)
    # f32 algebra on the device against f64 in the oracle: singular values to ~1e-5, loadings
    # and scores to the same precision relative to the leading singular value.
    np.testing.assert_allclose(got["explained_variance"], want["explained_variance"], rtol=2e-4)
    np.testing.assert_allclose(got["explained_variance_ratio"], want["explained_variance_ratio"], rtol=2e-4)
    for c in range(n_comps):
        cos = float(got["components"][c].astype(np.float64) @ want["components"][c].astype(np.float64))
        assert cos > 1 - 1e-4, (c, cos)  # same direction and the same sign
    s0 = float(np.sqrt(want["explained_variance"][0] * (x.shape[0] - 1)))
    np.testing.assert_allclose(got["embedding"], want["embedding"], atol=5e-4 * s0)
    np.testing.assert_allclose(got["components"] @ got["components"].T, np.eye(n_comps), atol=1e-4)


@pytest.mark.gpu
@pytest.mark.parametrize("zero_center", [True, False])
def test_device_pca_matches_the_oracle_on_sparse_counts(context, zero_center):
    print(f"[case] device PCA sparse counts zero_center={zero_center}", flush=True)
    x = sparse_counts(2000, 300, density=0.1, seed=4)
    got, want = _device_vs_oracle(x, 10, zero_center, context)
    np.testing.assert_allclose(got["explained_variance"][:3], want["explained_variance"][:3], rtol=1e-3)
    np.testing.assert_allclose(got["explained_variance"], want["explained_variance"], rtol=2e-2)
    assert float(got["components"][0] @ want["components"][0]) > 1 - 1e-5
    for c in range(10):
        comp = got["components"][c]
        assert comp[np.argmax(np.abs(comp))] > 0
    total = x.astype(np.float64).var(axis=0, ddof=1 if zero_center else 0).sum()
    np.testing.assert_allclose(got["explained_variance_ratio"], got["explained_variance"] / total, rtol=1e-5)


@pytest.mark.gpu
def test_device_pca_is_deterministic_and_the_two_entry_points_agree(context):
    import torch

    import mrustsc_amd
    from mrustsc_amd import synthetic

    print("[case] device PCA determinism and entry points", flush=True)
    ip, ix, v = synthetic.expression_csr(5000, 800, seed=3)
    a = mrustsc_amd.pca(ip, ix, v, 800, 30, context=context)
    b = mrustsc_amd.pca(ip, ix, v, 800, 30, context=context)
    for key in a:
        assert np.array_equal(a[key], b[key]), key
    d_ip, d_ix, d_v = (torch.from_numpy(t.view(np.int32) if t.dtype == np.uint32 else t).cuda() for t in (ip, ix, v))
    emb = torch.empty((5000, 30), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    c = mrustsc_amd.pca_device(d_ip.data_ptr(), d_ix.data_ptr(), d_v.data_ptr(), 5000, 800, 30, emb.data_ptr(),
                               context=context)
    assert np.array_equal(emb.cpu().numpy(), a["embedding"])
    for key in c:
        assert np.array_equal(c[key], a[key]), key


@pytest.mark.gpu
def test_pca_to_knn_on_the_device_equals_the_oracle(context):
    """The f3 chain: a device-resident CSR matrix -> PCA -> the kNN graph of its embedding,
    with nothing crossing PCIe in between. The embedding is checked against the PCA oracle
    (randomised-SVD accuracy) and the graph against the kNN oracle on the same embedding (bit
    for bit)."""
    import torch

    import mrustsc_amd
    from mrustsc_amd import synthetic

    n, g, comps, k = 20000, 1000, 50, 15
    print(f"[case] PCA -> kNN on the device, {n} x {g}", flush=True)
    ip, ix, v = synthetic.expression_csr(n, g, seed=5)
    d_ip, d_ix, d_v = (torch.from_numpy(t.view(np.int32) if t.dtype == np.uint32 else t).cuda() for t in (ip, ix, v))
    emb = torch.empty((n, comps), dtype=torch.float32, device="cuda")
    ids = torch.empty((n, k), dtype=torch.int32, device="cuda")
    dist = torch.empty((n, k), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    res = mrustsc_amd.pca_device(d_ip.data_ptr(), d_ix.data_ptr(), d_v.data_ptr(), n, g, comps, emb.data_ptr(),
                                 context=context)
    mrustsc_amd.knn_device(emb.data_ptr(), n, comps, k, ids.data_ptr(), dist.data_ptr(), context=context)
    e = emb.cpu().numpy()
    want = oracle.pca(ip, ix, v, g, comps)
    np.testing.assert_allclose(res["explained_variance"][:5], want["explained_variance"][:5], rtol=1e-3)
    s0 = float(np.sqrt(want["explained_variance"][0] * (n - 1)))
    np.testing.assert_allclose(e[:, :5], want["embedding"][:, :5], atol=2e-3 * s0)
    w_ids, w_dist = oracle.knn_fast(e, k)
    np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32), w_ids)
    np.testing.assert_array_equal(dist.cpu().numpy(), w_dist)
