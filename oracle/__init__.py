"""CPU oracle for the exact kNN graph -- test infrastructure, never the product.

Only the tests, the smoke check and the benchmark's CPU-baseline leg import this
package. It wraps ``build/liboracle.so`` (``knn_oracle.c``), a C restatement of
the reference's kNN semantics; each C function cites the reference lines it follows.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

EUCLIDEAN = 0
COSINE = 1
METRICS = {"euclidean": EUCLIDEAN, "cosine": COSINE}

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "build" / "liboracle.so"
_lib = None


def build() -> Path:
    """Compile the oracle library (gcc); returns its path."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        build()
    lib = ctypes.CDLL(str(_LIB_PATH))
    f32p = ctypes.POINTER(ctypes.c_float)
    f64p = ctypes.POINTER(ctypes.c_double)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    sz = ctypes.c_size_t
    lib.oracle_exact_sum_f32.argtypes = [f32p, sz, sz]
    lib.oracle_exact_sum_f32.restype = ctypes.c_double
    lib.oracle_column_means.argtypes = [f32p, sz, sz, f64p]
    lib.oracle_column_means.restype = None
    lib.oracle_prepare.argtypes = [f32p, sz, sz, ctypes.c_int, f32p, f32p]
    lib.oracle_prepare.restype = None
    lib.oracle_knn.argtypes = [f32p, sz, sz, sz, ctypes.c_int, ctypes.c_int, u32p, sz,
                               ctypes.c_int, u32p, f32p]
    lib.oracle_knn.restype = ctypes.c_int
    lib.oracle_knn_fast.argtypes = [f32p, sz, sz, sz, ctypes.c_int, u32p, sz, ctypes.c_int, u32p,
                                    f32p]
    lib.oracle_knn_fast.restype = ctypes.c_int
    lib.oracle_pca.argtypes = [u32p, u32p, f32p, sz, sz, sz, ctypes.c_int, ctypes.c_uint64, f32p, f32p, f32p, f32p]
    lib.oracle_pca.restype = ctypes.c_int
    lib.oracle_pca_sketch.argtypes = [sz, sz, ctypes.c_uint64, f32p]
    lib.oracle_pca_sketch.restype = None
    lib.oracle_jacobi_eigen.argtypes = [ctypes.POINTER(ctypes.c_double), sz, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double)]
    lib.oracle_jacobi_eigen.restype = ctypes.c_int
    lib.oracle_knn_fp64.argtypes = [f32p, sz, sz, sz, ctypes.c_int, u32p, sz, ctypes.c_int,
                                    u32p, f64p]
    lib.oracle_knn_fp64.restype = ctypes.c_int
    lib.baseline_expansion_knn.argtypes = [f32p, sz, sz, sz, sz, u32p, sz, ctypes.c_int, u32p,
                                           f32p]
    lib.baseline_expansion_knn.restype = ctypes.c_int
    lib.oracle_default_threads.argtypes = []
    lib.oracle_default_threads.restype = ctypes.c_int
    lib.oracle_ndarray_sum_f32.argtypes = [f32p, sz]
    lib.oracle_ndarray_sum_f32.restype = ctypes.c_float
    lib.oracle_connectivities.argtypes = [u32p, f32p, sz, sz, u32p, u32p, f32p, sz, f32p, f32p]
    lib.oracle_connectivities.restype = ctypes.c_long
    lib.oracle_tsne_squared_distances.argtypes = [f32p, sz, sz, sz, sz, f32p, ctypes.c_int]
    lib.oracle_tsne_squared_distances.restype = ctypes.c_int
    _lib = lib
    return lib


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _f32(x) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    if x.ndim != 2:
        raise ValueError("embedding must be 2-D")
    return x


def _queries(queries, n):
    if queries is None:
        return None, n
    q = np.ascontiguousarray(queries, dtype=np.uint32)
    if q.size and int(q.max()) >= n:
        raise ValueError("query id out of range")
    return q, q.size


def default_threads() -> int:
    return int(_load().oracle_default_threads())


def exact_sum(values) -> float:
    v = np.ascontiguousarray(values, dtype=np.float32).ravel()
    return float(_load().oracle_exact_sum_f32(_ptr(v, ctypes.c_float), v.size, 1))


def column_means(x) -> np.ndarray:
    x = _f32(x)
    out = np.empty(x.shape[1], dtype=np.float64)
    _load().oracle_column_means(_ptr(x, ctypes.c_float), x.shape[0], x.shape[1],
                                _ptr(out, ctypes.c_double))
    return out


def prepare(x, metric: str = "euclidean") -> tuple[np.ndarray, np.ndarray]:
    """(coordinates, f32 squared norms) every oracle distance is computed from."""
    x = _f32(x)
    n, d = x.shape
    coords = np.empty((n, d), dtype=np.float32)
    norms = np.empty(n, dtype=np.float32)
    _load().oracle_prepare(_ptr(x, ctypes.c_float), n, d, METRICS[metric],
                           _ptr(coords, ctypes.c_float), _ptr(norms, ctypes.c_float))
    return coords, norms


def knn(x, k: int, *, metric: str = "euclidean", tie_rule: bool = True, queries=None,
        threads: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """Oracle kNN: (uint32 ids, float32 distances) for every row (or `queries`)."""
    x = _f32(x)
    n, d = x.shape
    q, nq = _queries(queries, n)
    idx = np.empty((nq, k), dtype=np.uint32)
    dist = np.empty((nq, k), dtype=np.float32)
    rc = _load().oracle_knn(_ptr(x, ctypes.c_float), n, d, k, METRICS[metric], int(tie_rule),
                            None if q is None else _ptr(q, ctypes.c_uint32), nq, threads,
                            _ptr(idx, ctypes.c_uint32), _ptr(dist, ctypes.c_float))
    if rc != 0:
        raise ValueError(f"invalid arguments: n={n} d={d} k={k}")
    return idx, dist


def knn_fast(x, k: int, *, metric: str = "euclidean", queries=None,
             threads: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """knn (tie rule on) vectorised across references: the same output bit for bit, fast
    enough for whole graphs of 10^6 cells."""
    x = _f32(x)
    n, d = x.shape
    q, nq = _queries(queries, n)
    idx = np.empty((nq, k), dtype=np.uint32)
    dist = np.empty((nq, k), dtype=np.float32)
    rc = _load().oracle_knn_fast(_ptr(x, ctypes.c_float), n, d, k, METRICS[metric],
                                 None if q is None else _ptr(q, ctypes.c_uint32), nq, threads,
                                 _ptr(idx, ctypes.c_uint32), _ptr(dist, ctypes.c_float))
    if rc != 0:
        raise ValueError(f"invalid arguments: n={n} d={d} k={k}")
    return idx, dist


def knn_fp64(x, k: int, *, metric: str = "euclidean", queries=None,
             threads: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """fp64 ground truth on the raw input: (uint32 ids, float64 distances)."""
    x = _f32(x)
    n, d = x.shape
    q, nq = _queries(queries, n)
    idx = np.empty((nq, k), dtype=np.uint32)
    dist = np.empty((nq, k), dtype=np.float64)
    rc = _load().oracle_knn_fp64(_ptr(x, ctypes.c_float), n, d, k, METRICS[metric],
                                 None if q is None else _ptr(q, ctypes.c_uint32), nq, threads,
                                 _ptr(idx, ctypes.c_uint32), _ptr(dist, ctypes.c_double))
    if rc != 0:
        raise ValueError(f"invalid arguments: n={n} d={d} k={k}")
    return idx, dist


def baseline_knn(x, k: int, *, tile_rows: int = 0, queries=None,
                 threads: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """The reference's CPU expansion path restated (the timed CPU baseline)."""
    x = _f32(x)
    n, d = x.shape
    q, nq = _queries(queries, n)
    idx = np.empty((nq, k), dtype=np.uint32)
    dist = np.empty((nq, k), dtype=np.float32)
    rc = _load().baseline_expansion_knn(_ptr(x, ctypes.c_float), n, d, k, tile_rows,
                                        None if q is None else _ptr(q, ctypes.c_uint32), nq,
                                        threads, _ptr(idx, ctypes.c_uint32),
                                        _ptr(dist, ctypes.c_float))
    if rc != 0:
        raise ValueError(f"invalid arguments: n={n} d={d} k={k}")
    return idx, dist


def ndarray_sum(values) -> float:
    """ndarray 0.16.1's sum of a contiguous f32 array (eight-way unrolled fold)."""
    v = np.ascontiguousarray(values, dtype=np.float32).ravel()
    return float(_load().oracle_ndarray_sum_f32(_ptr(v, ctypes.c_float), v.size))


def connectivities(indices, distances):
    """UMAP's fuzzy simplicial set of a kNN graph (neighbors.rs:193-372):
    (indptr uint32[n+1], indices uint32[nnz], values float32[nnz], sigmas, rhos)."""
    idx = np.ascontiguousarray(indices, dtype=np.uint32)
    dist = np.ascontiguousarray(distances, dtype=np.float32)
    if idx.ndim != 2 or idx.shape != dist.shape:
        raise ValueError(f"indices {idx.shape} and distances {dist.shape} must be equal 2-D shapes")
    n, k = idx.shape
    cap = 2 * n * k
    indptr = np.empty(n + 1, dtype=np.uint32)
    cols = np.empty(max(cap, 1), dtype=np.uint32)
    vals = np.empty(max(cap, 1), dtype=np.float32)
    sigmas = np.empty(n, dtype=np.float32)
    rhos = np.empty(n, dtype=np.float32)
    nnz = _load().oracle_connectivities(_ptr(idx, ctypes.c_uint32), _ptr(dist, ctypes.c_float), n, k,
                                        _ptr(indptr, ctypes.c_uint32), _ptr(cols, ctypes.c_uint32),
                                        _ptr(vals, ctypes.c_float), cap, _ptr(sigmas, ctypes.c_float),
                                        _ptr(rhos, ctypes.c_float))
    if nnz < 0:
        raise ValueError(f"invalid kNN graph: n={n} k={k}")
    return indptr, cols[:nnz].copy(), vals[:nnz].copy(), sigmas, rhos


def tsne_squared_distances(points, *, rows=None, threads: int = 0) -> np.ndarray:
    """t-SNE's input distances (tsne.rs:86, :173-188): the (n, n) float32 matrix, or the
    rows ``range(*rows)`` of it."""
    x = _f32(points)
    n, d = x.shape
    begin, end = (0, n) if rows is None else rows
    out = np.empty((max(end - begin, 0), n), dtype=np.float32)
    rc = _load().oracle_tsne_squared_distances(_ptr(x, ctypes.c_float), n, d, begin, end,
                                               _ptr(out, ctypes.c_float), threads)
    if rc != 0:
        raise ValueError(f"invalid arguments: n={n} d={d} rows={rows}")
    return out


def threads_env_default() -> int:
    value = os.environ.get("OMP_NUM_THREADS")
    return int(value) if value and value.isdigit() else 0


def pca(indptr, indices, values, n_cols: int, n_components: int, *, zero_center: bool = True,
        seed: int = 0) -> dict:
    """Randomised-SVD PCA of a CSR matrix (pca.rs:74-215 restated with f64 algebra): the
    dict the reference's _scrust.pca returns (embedding, components, explained_variance,
    explained_variance_ratio)."""
    indptr = np.ascontiguousarray(indptr, dtype=np.uint32)
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    values = np.ascontiguousarray(values, dtype=np.float32)
    n = indptr.size - 1
    emb = np.empty((n, n_components), dtype=np.float32)
    comp = np.empty((n_components, n_cols), dtype=np.float32)
    ev = np.empty(n_components, dtype=np.float32)
    evr = np.empty(n_components, dtype=np.float32)
    rc = _load().oracle_pca(_ptr(indptr, ctypes.c_uint32), _ptr(indices, ctypes.c_uint32),
                            _ptr(values, ctypes.c_float), n, n_cols, n_components, int(zero_center),
                            seed, _ptr(emb, ctypes.c_float), _ptr(comp, ctypes.c_float),
                            _ptr(ev, ctypes.c_float), _ptr(evr, ctypes.c_float))
    if rc == -1:
        raise ValueError(f"invalid arguments: n={n} n_cols={n_cols} n_components={n_components}")
    if rc != 0:
        raise RuntimeError("Jacobi eigenvalue iteration did not converge")
    return {"embedding": emb, "components": comp, "explained_variance": ev,
            "explained_variance_ratio": evr}


def pca_sketch(rows: int, cols: int, seed: int) -> np.ndarray:
    out = np.empty((rows, cols), dtype=np.float32)
    _load().oracle_pca_sketch(rows, cols, seed, _ptr(out, ctypes.c_float))
    return out


def jacobi_eigen(a: np.ndarray):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    n = a.shape[0]
    vals = np.empty(n)
    vecs = np.empty((n, n))
    p = ctypes.POINTER(ctypes.c_double)
    rc = _load().oracle_jacobi_eigen(a.ctypes.data_as(p), n, vals.ctypes.data_as(p), vecs.ctypes.data_as(p))
    if rc != 0:
        raise RuntimeError("Jacobi eigenvalue iteration did not converge")
    return vals, vecs
