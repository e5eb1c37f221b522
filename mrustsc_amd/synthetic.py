"""Deterministic synthetic embeddings for tests and benchmarks.

The generators the reference's tests use are restated bit for bit (they are defined by
integer arithmetic, so any language reproduces them):
  * ``reference_embedding``: splitmix64 (crates/scrust-core/src/neighbors.rs:380-392)
  * ``random_embedding``: the PCG-style LCG (crates/scrust-gpu/src/kernels/knn.rs:317-325)
  * ``line``, ``grid`` (neighbors.rs:394-406) and the 8x8 lattice
    (tests/test_neighbors_audit.py:347-349)
The Gaussian mixtures are the benchmark workloads: seeded numpy draws, f64 then one
rounding to f32.
"""

from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_MIX1 = np.uint64(0xBF58476D1CE4E5B9)
_MIX2 = np.uint64(0x94D049BB133111EB)


def reference_embedding(n_cells: int, n_dims: int, seed: int) -> np.ndarray:
    """splitmix64 stream, top 24 bits of each output scaled to [0, 1)."""
    count = n_cells * n_dims
    with np.errstate(over="ignore"):
        steps = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + steps * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _MIX1
        z = (z ^ (z >> np.uint64(27))) * _MIX2
        z = z ^ (z >> np.uint64(31))
    values = (z >> np.uint64(40)).astype(np.float32) / np.float32(1 << 24)
    return values.reshape(n_cells, n_dims)


def random_embedding(n_cells: int, n_dims: int, seed: int) -> np.ndarray:
    """state <- state * 6364136223846793005 + 1442695040888963407 (mod 2^64), starting
    from seed * 2 + 1; each value is (state >> 40) / 2^24."""
    mask = (1 << 64) - 1
    state = (seed * 2 + 1) & mask
    out = np.empty(n_cells * n_dims, dtype=np.float32)
    for i in range(out.size):
        state = (state * 6364136223846793005 + 1442695040888963407) & mask
        out[i] = np.float32(state >> 40) / np.float32(1 << 24)
    return out.reshape(n_cells, n_dims)


def line(n_cells: int) -> np.ndarray:
    return np.arange(n_cells, dtype=np.float32)[:, None]


def grid(side: int) -> np.ndarray:
    cells = np.arange(side * side)
    return np.stack([cells // side, cells % side], axis=1).astype(np.float32)


def lattice(side: int = 8) -> np.ndarray:
    return np.array([[i, j] for i in range(side) for j in range(side)], dtype=np.float32)


def isotropic_mixture(n_cells: int, n_dims: int, n_clusters: int, seed: int,
                      centre_sd: float = 5.0, spread: float = 1.0,
                      n_duplicates: int = 0) -> np.ndarray:
    """Centres ~ N(0, centre_sd^2 I); points = centre + N(0, spread^2 I); then
    `n_duplicates` rows overwritten by exact copies of other rows."""
    rng = np.random.default_rng(seed)
    centres = rng.normal(0.0, centre_sd, size=(n_clusters, n_dims))
    z = rng.integers(0, n_clusters, size=n_cells)
    x = centres[z] + rng.normal(0.0, spread, size=(n_cells, n_dims))
    _duplicate_rows(x, n_duplicates, rng)
    return x.astype(np.float32)


def anisotropic_mixture(n_cells: int, n_dims: int, n_clusters: int, seed: int,
                        duplicate_fraction: float = 0.005, centre_sd: float = 5.0,
                        eig_low: float = 0.1, eig_high: float = 10.0) -> np.ndarray:
    """Centres ~ N(0, centre_sd^2 I); cluster covariance Q diag(lambda) Q^T with lambda
    log-uniform in [eig_low, eig_high] and Q from the QR of a Gaussian matrix; then
    round(duplicate_fraction * n) rows overwritten by exact copies of other rows."""
    rng = np.random.default_rng(seed)
    centres = rng.normal(0.0, centre_sd, size=(n_clusters, n_dims))
    z = rng.integers(0, n_clusters, size=n_cells)
    x = np.empty((n_cells, n_dims), dtype=np.float64)
    for c in range(n_clusters):
        q, _ = np.linalg.qr(rng.normal(size=(n_dims, n_dims)))
        lam = np.exp(rng.uniform(np.log(eig_low), np.log(eig_high), size=n_dims))
        rows = np.nonzero(z == c)[0]
        x[rows] = centres[c] + (rng.normal(size=(rows.size, n_dims)) * np.sqrt(lam)) @ q.T
    _duplicate_rows(x, int(round(duplicate_fraction * n_cells)), rng)
    return x.astype(np.float32)


def _duplicate_rows(x: np.ndarray, count: int, rng: np.random.Generator) -> None:
    if count <= 0:
        return
    n = x.shape[0]
    dst = rng.choice(n, size=count, replace=False)
    src = rng.integers(0, n, size=count)
    x[dst] = x[src]


# Workloads ---------------------------------------------------------------

def pbmc_like(seed: int = 0) -> np.ndarray:
    """2,638 x 50: 8 isotropic Gaussians (centres sd 5, spread sd 1) + 20 exact duplicates."""
    return isotropic_mixture(2638, 50, 8, seed, n_duplicates=20)


def clustered(n_cells: int, n_dims: int = 50, seed: int = 1, n_clusters: int = 20,
              duplicate_fraction: float = 0.005) -> np.ndarray:
    """The 20-cluster anisotropic mixture with 0.5% duplicate rows (100k and 1M workloads)."""
    return anisotropic_mixture(n_cells, n_dims, n_clusters, seed, duplicate_fraction)


def batch_specs(count: int = 64, seed: int = 7) -> list[dict]:
    """Dataset shapes of the batched workload: n log-uniform in [500, 50000],
    d from {30, 50, 100}, k from {15, 30}, 5-30 clusters, one seed each."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        n = int(round(np.exp(rng.uniform(np.log(500), np.log(50000)))))
        specs.append({
            "n": n,
            "d": int(rng.choice([30, 50, 100])),
            "k": int(rng.choice([15, 30])),
            "clusters": int(rng.integers(5, 31)),
            "seed": int(seed * 1000 + i),
        })
    return specs


def batch_dataset(spec: dict) -> np.ndarray:
    return anisotropic_mixture(spec["n"], spec["d"], spec["clusters"], spec["seed"],
                               duplicate_fraction=0.005)


def expression_csr(n_cells: int, n_genes: int, n_clusters: int = 12, density: float = 0.1,
                   seed: int = 0):
    """A cells x genes expression matrix in CSR form (indptr, indices, values as uint32,
    uint32, float32): each cluster has its own mean profile over the genes; a cell keeps
    each gene with probability `density` (raised for the cluster's marker genes) at
    log1p(Poisson(profile)) + 1. The input of the PCA that produces the kNN embedding."""
    rng = np.random.default_rng(seed)
    profiles = np.exp(rng.normal(0.0, 1.0, size=(n_clusters, n_genes)))
    markers = rng.random((n_clusters, n_genes)) < 0.05
    z = rng.integers(0, n_clusters, size=n_cells)
    indptr = np.zeros(n_cells + 1, dtype=np.uint64)
    cols, vals = [], []
    block = 8192
    for r0 in range(0, n_cells, block):
        zz = z[r0:r0 + block]
        keep = rng.random((zz.size, n_genes)) < np.where(markers[zz], 0.8, density)
        rows, cc = np.nonzero(keep)
        lam = profiles[zz[rows], cc] * (1.0 + 4.0 * markers[zz[rows], cc])
        cols.append(cc.astype(np.uint32))
        vals.append(np.log1p(rng.poisson(lam) + 1).astype(np.float32))
        indptr[r0 + 1:r0 + zz.size + 1] = np.cumsum(keep.sum(axis=1))
    offsets = np.cumsum([0] + [c.size for c in cols]).astype(np.uint64)
    for i, r0 in enumerate(range(0, n_cells, block)):
        indptr[r0 + 1:min(n_cells, r0 + block) + 1] += offsets[i]
    return indptr.astype(np.uint32), np.concatenate(cols), np.concatenate(vals)
