"""t-SNE's input distances, squared_distances(&centred(&points))
(crates/scrust-core/src/tsne.rs:86, :173-188): the oracle pinned against closed forms and an
f64 evaluation, then the device matrix against the oracle.

The reference's summation orders are candle's and gemm's, so the bar is the expansion's
rounding bound, not bitwise. Both the oracle and the device take the column sums exactly
and the squared norms in dimension order, so their coordinates and norms agree bit for bit
and only the Gram product differs: |device - oracle| <= (2 d + 8) 2^-24 (|a|^2 + |b|^2).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from mrustsc_amd import synthetic

U = 2.0**-24


def centred_f32(x):
    """Coordinates as tsne.rs:173-177 forms them: exact f32 column sums rounded once, times
    f32(1/n), subtracted in f32; and the f32 squared norms summed in dimension order."""
    n, d = x.shape
    sums = np.array([math.fsum(x[:, j].astype(np.float64)) for j in range(d)])
    mean = sums.astype(np.float32) * np.float32(1.0 / n)
    c = (x - mean).astype(np.float32)
    norms = np.zeros(n, dtype=np.float32)
    for t in range(d):
        norms = (norms + (c[:, t] * c[:, t]).astype(np.float32)).astype(np.float32)
    return c, norms


def truth(c):
    c64 = c.astype(np.float64)
    g = c64 @ c64.T
    sq = (c64 * c64).sum(1)
    return np.maximum(sq[:, None] + sq[None, :] - 2.0 * g, 0.0)


def bound(norms, d, slack):
    n64 = norms.astype(np.float64)
    return (2 * d + slack) * U * (n64[:, None] + n64[None, :]) + 1e-30


def gaussian(n, d, seed, offset=0.0):
    return (np.random.default_rng(seed).normal(size=(n, d)) + offset).astype(np.float32)


# --- the oracle ------------------------------------------------------------------------


def test_oracle_is_exact_on_the_lattice_and_the_line():
    # Integer points with a dyadic mean: every product and sum is exact, so the matrix is
    # the integer squared distances. A lost factor of 2, a sign or a transposed index fails.
    for pts in (synthetic.lattice(8), synthetic.line(100), synthetic.grid(6)):
        got = oracle.tsne_squared_distances(pts)
        want = ((pts[:, None, :].astype(np.float64) - pts[None, :, :]) ** 2).sum(-1)
        assert np.array_equal(got, want.astype(np.float32))


def test_oracle_is_within_the_expansion_bound_of_an_f64_evaluation():
    for x in (synthetic.clustered(500, 50, seed=3), gaussian(300, 7, 1), gaussian(200, 10, 2, offset=500.0)):
        d = x.shape[1]
        c, norms = centred_f32(x)
        got = oracle.tsne_squared_distances(x).astype(np.float64)
        assert np.all(np.abs(got - truth(c)) <= bound(norms, d, 3))


def test_oracle_matrix_is_symmetric_with_a_zero_diagonal():
    x = synthetic.clustered(300, 20, seed=5)
    got = oracle.tsne_squared_distances(x)
    assert np.array_equal(got, got.T)
    assert np.all(np.diag(got) == 0.0)
    assert np.all(got >= 0.0)


def test_oracle_centring_removes_the_offset_cancellation():
    # tsne.rs:163-172: on unit-variance points offset by 500 the centred expansion keeps
    # its relative accuracy; the same points without the offset give the same distances
    # up to the rounding of the offset inputs.
    x = gaussian(200, 10, 1, offset=500.0)
    got = oracle.tsne_squared_distances(x).astype(np.float64)
    exact = ((x[:, None, :].astype(np.float64) - x[None, :, :]) ** 2).sum(-1)
    off = ~np.eye(200, dtype=bool)
    assert (np.abs(got - exact)[off] / exact[off]).max() < 1e-5


def test_oracle_row_ranges_match_the_full_matrix():
    x = synthetic.clustered(257, 13, seed=8)
    full = oracle.tsne_squared_distances(x)
    assert np.array_equal(oracle.tsne_squared_distances(x, rows=(100, 157)), full[100:157])


def test_arguments_are_checked_before_touching_a_device():
    import mrustsc_amd

    with pytest.raises(ValueError, match="feature"):
        mrustsc_amd.tsne_squared_distances(np.zeros((5, 0), dtype=np.float32))
    with pytest.raises(ValueError, match="n_cells"):
        mrustsc_amd.tsne_squared_distances(np.zeros((20001, 1), dtype=np.float32))
    with pytest.raises(ValueError):
        mrustsc_amd.tsne_squared_distances(np.zeros((5, 3), dtype=np.float64))


# --- the device -------------------------------------------------------------------------


def assert_within_bound(got, x):
    d = x.shape[1]
    want = oracle.tsne_squared_distances(x).astype(np.float64)
    _, norms = centred_f32(x)
    err = np.abs(got.astype(np.float64) - want)
    lim = bound(norms, d, 8)
    worst = np.unravel_index(np.argmax(err / lim), err.shape)
    assert np.all(err <= lim), (worst, got[worst], want[worst], lim[worst])
    assert np.all(got >= 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("n,d", [(1, 1), (2, 3), (129, 13), (300, 50), (777, 33), (1000, 64), (513, 100)])
def test_device_matches_the_oracle(context, n, d):
    import mrustsc_amd

    x = synthetic.clustered(n, d, seed=n + d) if n > 20 else gaussian(n, d, 4)
    print(f"[case] tsne distances n={n} d={d}", flush=True)
    got = mrustsc_amd.tsne_squared_distances(x, context=context)
    assert got.shape == (n, n)
    assert_within_bound(got, x)


@pytest.mark.gpu
def test_device_is_exact_on_integer_points(context):
    import mrustsc_amd

    for pts in (synthetic.lattice(8), synthetic.line(200), synthetic.grid(13)):
        print(f"[case] tsne distances integer points {pts.shape}", flush=True)
        got = mrustsc_amd.tsne_squared_distances(pts, context=context)
        want = ((pts[:, None, :].astype(np.float64) - pts[None, :, :]) ** 2).sum(-1)
        assert np.array_equal(got, want.astype(np.float32))


@pytest.mark.gpu
def test_device_keeps_accuracy_on_offset_points(context):
    import mrustsc_amd

    x = gaussian(200, 10, 1, offset=500.0)
    print("[case] tsne distances offset 500", flush=True)
    got = mrustsc_amd.tsne_squared_distances(x, context=context)
    assert_within_bound(got, x)
    exact = ((x[:, None, :].astype(np.float64) - x[None, :, :]) ** 2).sum(-1)
    off = ~np.eye(200, dtype=bool)
    assert (np.abs(got.astype(np.float64) - exact)[off] / exact[off]).max() < 1e-5


@pytest.mark.gpu
def test_device_entry_matches_the_host_entry(context):
    import torch

    import mrustsc_amd

    x = synthetic.clustered(700, 50, seed=12)
    print("[case] tsne distances device entry", flush=True)
    host = mrustsc_amd.tsne_squared_distances(x, context=context)
    dx = torch.from_numpy(x).to("cuda")
    out = torch.empty((700, 700), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    mrustsc_amd.tsne_squared_distances_device(dx.data_ptr(), 700, 50, out.data_ptr(), context=context)
    assert np.array_equal(out.cpu().numpy(), host)


@pytest.mark.gpu
def test_device_at_the_cell_cap_matches_oracle_rows(context):
    import mrustsc_amd

    n = 20000
    x = synthetic.clustered(n, 50, seed=2)
    print("[case] tsne distances n=20000", flush=True)
    got = mrustsc_amd.tsne_squared_distances(x, context=context)
    _, norms = centred_f32(x)
    n64 = norms.astype(np.float64)
    lim_of = lambda r, c: (2 * 50 + 8) * U * (n64[r][:, None] + n64[c][None, :])  # noqa: E731
    for lo in (0, 9984, n - 64):
        want = oracle.tsne_squared_distances(x, rows=(lo, lo + 64)).astype(np.float64)
        sl = np.arange(lo, lo + 64)
        assert np.all(np.abs(got[lo:lo + 64].astype(np.float64) - want) <= lim_of(sl, np.arange(n)))
    # Size-independent properties: non-negative everywhere, a diagonal within the bound of
    # zero, and blocks symmetric to the bound.
    assert np.all(got >= 0.0)
    assert np.all(np.diag(got) <= (2 * 50 + 8) * U * 2 * n64)
    for a, b in ((0, 18000), (5000, 12000), (19000, 1000)):
        ra, rb = np.arange(a, a + 1000), np.arange(b, b + 1000)
        asym = np.abs(got[a:a + 1000, b:b + 1000].astype(np.float64) - got[b:b + 1000, a:a + 1000].T)
        assert np.all(asym <= lim_of(ra, rb))


@pytest.mark.gpu
def test_device_rejects_non_finite_points(context):
    import mrustsc_amd

    x = gaussian(50, 4, 3)
    x[7, 2] = np.nan
    with pytest.raises(ValueError, match="finite"):
        mrustsc_amd.tsne_squared_distances(x, context=context)
    # The context stays usable.
    y = gaussian(50, 4, 3)
    assert_within_bound(mrustsc_amd.tsne_squared_distances(y, context=context), y)
