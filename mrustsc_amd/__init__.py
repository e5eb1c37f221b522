"""scrust_hip: the exact k-nearest-neighbour graph behind scrust's ``pp.neighbors`` on
AMD Instinct MI355X (gfx950).

Python mirror of the drop-in boundary. ``knn(embedding, k, device)`` keeps the
signature, argument checks and outputs of the reference's ``_scrust.knn``
(crates/scrust-py/src/embedding.rs:60-77): a 2-D float32 numpy array in, a pair of
(uint32 ids, float32 distances) arrays of shape (n, k) out, nearest first, ties to the
smaller id. Work runs in libscrust_hip.so (hand-written HIP kernels); there is no CPU
fallback in this package.
"""

from __future__ import annotations

import ctypes
import threading

import numpy as np

from . import _native
from ._native import KnnOptions, KnnStats, ScrustHipError

__all__ = [
    "Context",
    "KnnStats",
    "ScrustHipError",
    "gpu_available",
    "knn",
    "knn_hip",
    "knn_device",
    "knn_device_noalloc",
    "NOALLOC_SCRATCH_BYTES",
    "connectivities",
    "connectivities_device",
    "neighbors",
    "neighbors_device",
    "tsne_squared_distances",
    "tsne_squared_distances_device",
    "pca",
    "pca_device",
    "tune",
    "tune_device",
    "tune_save",
    "tune_load",
    "tune_clear",
    "tune_size",
    "max_supported_k",
    "Batch",
    "Server",
]

DEVICES = ("auto", "gpu", "hip")


def gpu_available() -> bool:
    """True when a gfx950 device is usable (crates/scrust-py/src/lib.rs:37-40)."""
    return bool(_native.load().scrust_hip_gpu_available())


class Context:
    """A device context: one HIP stream plus a grow-only device workspace.

    Mirrors ``MetalContext`` (crates/scrust-gpu/src/context.rs:12-90): construction fails
    with ``ScrustHipError`` (status NO_GPU) when no usable device exists. Use one context
    from one thread at a time.
    """

    def __init__(self, device: int = 0):
        lib = _native.load()
        handle = ctypes.c_void_p()
        _native.raise_for(lib.scrust_hip_context_create(int(device), ctypes.byref(handle)))
        self._handle = handle
        self.device = int(device)

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._handle is None:
            raise ScrustHipError("context is closed")
        return self._handle

    def close(self) -> None:
        if getattr(self, "_handle", None) is not None:
            _native.load().scrust_hip_context_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def max_supported_k(self) -> int:
        return int(_native.load().scrust_hip_max_supported_k(self.handle))


_local = threading.local()


def _thread_context() -> Context:
    """One context per thread, built on first use (crates/scrust-py/src/embedding.rs:22-32)."""
    ctx = getattr(_local, "context", None)
    if ctx is None:
        ctx = Context(0)
        _local.context = ctx
    return ctx


def _handle_or_none(context: Context | None):
    if context is not None:
        return context.handle
    try:
        return _thread_context().handle
    except ScrustHipError:
        return None


def max_supported_k(context: Context | None = None) -> int:
    """Largest k any path serves (crates/scrust-gpu/src/kernels/knn.rs:21-29)."""
    return int(_native.load().scrust_hip_max_supported_k(
        None if context is None else context.handle))


def _embedding(embedding) -> np.ndarray:
    """Dtype and rank checks of convert.rs:70-93; a strided view is copied."""
    if not isinstance(embedding, np.ndarray) or embedding.ndim != 2 or embedding.dtype != np.float32:
        raise ValueError("embedding must be a 2-D numpy array of float32")
    return np.ascontiguousarray(embedding)


ORDERS = {"auto": 0, "input": 1, "grouped": 2}


def _options(metric: str, precision: str, tile: int = 0, order: str = "auto") -> KnnOptions:
    if metric not in _native.METRICS:
        raise ValueError(f"metric must be one of {sorted(_native.METRICS)}, got {metric!r}")
    if precision not in _native.PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(_native.PRECISIONS)}, got {precision!r}")
    if order not in ORDERS:
        raise ValueError(f"order must be one of {sorted(ORDERS)}, got {order!r}")
    return KnnOptions(_native.METRICS[metric], _native.PRECISIONS[precision], int(tile), ORDERS[order])


def knn_hip(embedding, k: int, *, metric: str = "euclidean", precision: str = "bf16",
            context: Context | None = None, stats: bool = False, tile: int = 0, order: str = "auto"):
    """kNN on the device from host memory. Returns (ids, distances) or, with
    ``stats=True``, (ids, distances, stats dict)."""
    x = _embedding(embedding)
    if isinstance(k, bool) or int(k) != k or k < 0:
        raise ValueError(f"k must be a non-negative integer, got {k!r}")
    k = int(k)
    opts = _options(metric, precision, tile, order)
    handle = _handle_or_none(context)
    n, d = x.shape
    ids = np.empty((n, k), dtype=np.uint32)
    dist = np.empty((n, k), dtype=np.float32)
    st = KnnStats()
    # Without a device the library still checks the arguments first, so argument
    # errors are reported as such and only valid calls report NO_GPU.
    status = _native.load().scrust_hip_knn_ex(
        handle, x.ctypes.data, n, d, k, ctypes.byref(opts), ids.ctypes.data,
        dist.ctypes.data, ctypes.byref(st))
    _native.raise_for(status)
    if stats:
        return ids, dist, st.as_dict()
    return ids, dist


def knn(embedding, k: int, device: str = "auto"):
    """Drop-in for ``_scrust.knn(embedding, k, device)`` (embedding.rs:60-77)."""
    if device not in DEVICES:
        if device in ("cpu", "metal"):
            raise ValueError(
                f"device {device!r} is served by scrust's own core; this package is its HIP path")
        raise ValueError(f"device must be one of auto, gpu, hip, got {device}")
    return knn_hip(embedding, k)


def knn_device(embedding_ptr: int, n: int, d: int, k: int, indices_ptr: int, distances_ptr: int,
               *, metric: str = "euclidean", precision: str = "bf16",
               context: Context | None = None, tile: int = 0, order: str = "auto") -> dict:
    """kNN on device-resident buffers (raw device pointers, e.g. ``tensor.data_ptr()``).
    Returns the call's stats."""
    opts = _options(metric, precision, tile, order)
    ctx = context if context is not None else _thread_context()
    st = KnnStats()
    status = _native.load().scrust_hip_knn_device(
        ctx.handle, ctypes.c_void_p(embedding_ptr), n, d, k, ctypes.byref(opts),
        ctypes.c_void_p(indices_ptr), ctypes.c_void_p(distances_ptr), ctypes.byref(st))
    _native.raise_for(status)
    return st.as_dict()


NOALLOC_SCRATCH_BYTES = 256 << 20  # SCRUST_HIP_NOALLOC_SCRATCH_BYTES


def knn_device_noalloc(embedding_ptr: int, n: int, d: int, k: int, indices_ptr: int,
                       distances_ptr: int, *, metric: str = "euclidean", precision: str = "bf16",
                       context: Context | None = None, tile: int = 0, order: str = "auto") -> dict:
    """As ``knn_device``, but the call allocates no device memory: its workspace is the
    library's per-device scratch array of ``NOALLOC_SCRATCH_BYTES`` (a build that does not
    fit raises ValueError). Calls that overlap on one device race. Returns the stats."""
    opts = _options(metric, precision, tile, order)
    ctx = context if context is not None else _thread_context()
    st = KnnStats()
    status = _native.load().scrust_hip_knn_device_noalloc(
        ctx.handle, ctypes.c_void_p(embedding_ptr), n, d, k, ctypes.byref(opts),
        ctypes.c_void_p(indices_ptr), ctypes.c_void_p(distances_ptr), ctypes.byref(st))
    _native.raise_for(status)
    return st.as_dict()


def connectivities(indices, distances, *, context: Context | None = None):
    """UMAP's fuzzy simplicial set of a kNN graph, mirroring ``_scrust.connectivities``
    (crates/scrust-py/src/embedding.rs:79-100): ``indices`` uint32 and ``distances``
    float32 arrays of equal 2-D shape (n, k) in, ``(indptr, indices, values, n_cols)`` CSR
    out (uint32, uint32, float32). Computed on the device."""
    idx = _graph_array(indices, np.uint32, "indices")
    dist = _graph_array(distances, np.float32, "distances")
    if idx.shape != dist.shape:
        raise ValueError(f"shape mismatch: expected distances shaped {idx.shape}, got {dist.shape}")
    n, k = idx.shape
    handle = _handle_or_none(context)
    cap = max(2 * n * k, 1)
    indptr = np.empty(n + 1, dtype=np.uint32)
    cols = np.empty(cap, dtype=np.uint32)
    vals = np.empty(cap, dtype=np.float32)
    nnz = ctypes.c_size_t()
    status = _native.load().scrust_hip_connectivities(
        handle, idx.ctypes.data_as(ctypes.c_void_p), dist.ctypes.data_as(ctypes.c_void_p), n, k,
        indptr.ctypes.data_as(ctypes.c_void_p), cols.ctypes.data_as(ctypes.c_void_p),
        vals.ctypes.data_as(ctypes.c_void_p), cap, ctypes.byref(nnz))
    _native.raise_for(status)
    return indptr, cols[: nnz.value].copy(), vals[: nnz.value].copy(), n


def connectivities_device(indices_ptr: int, distances_ptr: int, n: int, k: int, indptr_ptr: int,
                          out_indices_ptr: int, out_values_ptr: int, capacity: int, *,
                          context: Context | None = None) -> int:
    """Connectivities of a device-resident kNN graph into device CSR buffers; returns nnz."""
    ctx = context if context is not None else _thread_context()
    nnz = ctypes.c_size_t()
    status = _native.load().scrust_hip_connectivities_device(
        ctx.handle, ctypes.c_void_p(indices_ptr), ctypes.c_void_p(distances_ptr), n, k,
        ctypes.c_void_p(indptr_ptr), ctypes.c_void_p(out_indices_ptr), ctypes.c_void_p(out_values_ptr),
        capacity, ctypes.byref(nnz))
    _native.raise_for(status)
    return nnz.value


def neighbors(embedding, *, n_neighbors: int = 15, use_rep: str = "X_pca", metric: str = "euclidean",
              precision: str = "bf16", context: Context | None = None) -> dict:
    """What ``sr.pp.neighbors`` stores (python/scrust/pp/_basics.py:186-215), from an
    embedding array: ``obsp`` entries "distances" and "connectivities" as
    (indptr, indices, values, shape) CSR parts and ``uns["neighbors"]``. The kNN graph and
    its connectivities are built in one device call (``scrust_hip_neighbors``); scanpy's
    n_neighbors counts the cell itself, so the graph has n_neighbors - 1 columns."""
    if n_neighbors < 2:
        raise ValueError(f"n_neighbors must be at least 2, got {n_neighbors}")
    x = _embedding(embedding)
    n, d = x.shape
    k = int(n_neighbors) - 1
    opts = _options(metric, precision)
    handle = _handle_or_none(context)
    ids = np.empty((n, k), dtype=np.uint32)
    dist = np.empty((n, k), dtype=np.float32)
    cap = max(2 * n * k, 1)
    indptr = np.empty(n + 1, dtype=np.uint32)
    cols = np.empty(cap, dtype=np.uint32)
    vals = np.empty(cap, dtype=np.float32)
    nnz = ctypes.c_size_t()
    status = _native.load().scrust_hip_neighbors(
        handle, x.ctypes.data, n, d, k, ctypes.byref(opts), ids.ctypes.data, dist.ctypes.data,
        indptr.ctypes.data, cols.ctypes.data, vals.ctypes.data, cap, ctypes.byref(nnz))
    _native.raise_for(status)
    # knn returns one fixed-width row per cell, so the CSR offsets are the row starts.
    dist_indptr = np.arange(0, n * k + 1, k, dtype=np.uint32)
    return {
        "obsp": {
            "distances": (dist_indptr, ids.ravel(), dist.ravel(), (n, n)),
            "connectivities": (indptr, cols[: nnz.value].copy(), vals[: nnz.value].copy(), (n, n)),
        },
        "uns": {"neighbors": {"connectivities_key": "connectivities", "distances_key": "distances",
                              "params": {"n_neighbors": int(n_neighbors), "method": "umap",
                                         "use_rep": use_rep}}},
    }


def neighbors_device(embedding_ptr: int, n: int, d: int, n_neighbors: int, out_indices_ptr: int,
                     out_distances_ptr: int, conn_indptr_ptr: int, conn_indices_ptr: int,
                     conn_values_ptr: int, capacity: int, *, metric: str = "euclidean",
                     precision: str = "bf16", context: Context | None = None) -> int:
    """The device half of ``pp.neighbors`` with everything on the device: an (n, d) float32
    embedding already there (the PCA output kept on the device) in; the kNN graph
    (n x (n_neighbors - 1) uint32 ids and float32 distances, the rows of obsp["distances"])
    and the connectivities CSR (n + 1 offsets, ``capacity`` column ids and weights;
    2 n (n_neighbors - 1) always suffice) out, as device pointers. Returns nnz."""
    if n_neighbors < 2:
        raise ValueError(f"n_neighbors must be at least 2, got {n_neighbors}")
    opts = _options(metric, precision)
    ctx = context if context is not None else _thread_context()
    nnz = ctypes.c_size_t()
    status = _native.load().scrust_hip_neighbors_device(
        ctx.handle, ctypes.c_void_p(embedding_ptr), n, d, int(n_neighbors) - 1, ctypes.byref(opts),
        ctypes.c_void_p(out_indices_ptr), ctypes.c_void_p(out_distances_ptr),
        ctypes.c_void_p(conn_indptr_ptr), ctypes.c_void_p(conn_indices_ptr),
        ctypes.c_void_p(conn_values_ptr), capacity, ctypes.byref(nnz))
    _native.raise_for(status)
    return nnz.value


def tsne_squared_distances(points, *, context: Context | None = None) -> np.ndarray:
    """The pairwise squared distances exact t-SNE starts from,
    ``squared_distances(&centred(&points))`` (crates/scrust-core/src/tsne.rs:86, :173-188):
    a float32 (n, d) array in, the float32 (n, n) matrix out (diagonal included), computed on
    the device. Entries agree with the reference's to the expansion's rounding bound
    (2 d + 8) 2^-24 (|a|^2 + |b|^2). At most 20,000 cells (tsne.rs:146-152)."""
    if not isinstance(points, np.ndarray) or points.ndim != 2 or points.dtype != np.float32:
        raise ValueError("points must be a 2-D numpy array of float32")
    x = np.ascontiguousarray(points)
    n, d = x.shape
    handle = _handle_or_none(context)
    out = np.empty((n, n), dtype=np.float32)
    status = _native.load().scrust_hip_tsne_squared_distances(handle, x.ctypes.data, n, d, out.ctypes.data)
    _native.raise_for(status)
    return out


def tsne_squared_distances_device(points_ptr: int, n: int, d: int, out_ptr: int, *,
                                  context: Context | None = None) -> None:
    """Device-resident variant: an (n, d) float32 input and an (n, n) float32 output, both
    device pointers on the context's device."""
    ctx = context if context is not None else _thread_context()
    status = _native.load().scrust_hip_tsne_squared_distances_device(
        ctx.handle, ctypes.c_void_p(points_ptr), n, d, ctypes.c_void_p(out_ptr))
    _native.raise_for(status)


def _csr_array(a, dtype, name):
    if not isinstance(a, np.ndarray) or a.ndim != 1:
        raise ValueError(f"{name} must be a 1-D numpy array")
    return np.ascontiguousarray(a, dtype=dtype)


def pca(indptr, indices, values, n_cols: int, n_components: int, zero_center: bool = True,
        seed: int = 0, device: str = "auto", *, context: Context | None = None) -> dict:
    """Mirror of ``_scrust.pca`` (crates/scrust-py/src/preprocess.rs:154-185): truncated PCA
    of a CSR matrix (cells x genes) by randomised SVD on the device. Returns the dict the
    reference returns: ``embedding`` (n x n_components), ``components`` (n_components x
    n_cols), ``explained_variance``, ``explained_variance_ratio``."""
    if device not in DEVICES:
        raise ValueError(f"device must be one of auto, gpu, hip, got {device}")
    ip = _csr_array(indptr, np.uint32, "indptr")
    ix = _csr_array(indices, np.uint32, "indices")
    vs = _csr_array(values, np.float32, "values")
    n = ip.size - 1
    if n < 0 or ix.size != vs.size or (n >= 0 and ip.size and int(ip[-1]) != ix.size):
        raise ValueError(f"shape mismatch: expected indptr[-1] == len(indices) == len(values), got "
                         f"{int(ip[-1]) if ip.size else None}, {ix.size}, {vs.size}")
    k = int(n_components)
    emb = np.empty((max(n, 0), k), dtype=np.float32)
    comp = np.empty((k, int(n_cols)), dtype=np.float32)
    ev = np.empty(k, dtype=np.float32)
    evr = np.empty(k, dtype=np.float32)
    handle = _handle_or_none(context)
    status = _native.load().scrust_hip_pca(
        handle, ip.ctypes.data, ix.ctypes.data, vs.ctypes.data, max(n, 0), int(n_cols), k, int(bool(zero_center)),
        int(seed), emb.ctypes.data, comp.ctypes.data, ev.ctypes.data, evr.ctypes.data)
    _native.raise_for(status)
    return {"embedding": emb, "components": comp, "explained_variance": ev, "explained_variance_ratio": evr}


def pca_device(indptr_ptr: int, indices_ptr: int, values_ptr: int, n_rows: int, n_cols: int,
               n_components: int, out_embedding_ptr: int, *, zero_center: bool = True, seed: int = 0,
               context: Context | None = None) -> dict:
    """PCA of a device-resident CSR matrix; the embedding (n_rows x n_components float32) is
    written to device memory at ``out_embedding_ptr``, ready for ``knn_device`` or
    ``neighbors_device``. Returns the components and variances (host arrays)."""
    ctx = context if context is not None else _thread_context()
    k = int(n_components)
    comp = np.empty((k, int(n_cols)), dtype=np.float32)
    ev = np.empty(k, dtype=np.float32)
    evr = np.empty(k, dtype=np.float32)
    status = _native.load().scrust_hip_pca_device(
        ctx.handle, ctypes.c_void_p(indptr_ptr), ctypes.c_void_p(indices_ptr), ctypes.c_void_p(values_ptr),
        int(n_rows), int(n_cols), k, int(bool(zero_center)), int(seed), ctypes.c_void_p(out_embedding_ptr),
        comp.ctypes.data, ev.ctypes.data, evr.ctypes.data)
    _native.raise_for(status)
    return {"components": comp, "explained_variance": ev, "explained_variance_ratio": evr}


def tune(embedding, k: int, *, metric: str = "euclidean", precision: str = "bf16",
         skip_known: bool = False, context: Context | None = None) -> dict:
    """Times every work shape of the candidate pass (buffer 128/256/512 slots, reference
    ranges per query block, row order) on this embedding (host memory), checks that all of
    them return the same graph, and records the fastest for its (n, d, k) bucket in the
    process-wide table that later calls with ``tile=0`` use. Returns the tuning result."""
    x = _embedding(embedding)
    if isinstance(k, bool) or int(k) != k or k < 0:
        raise ValueError(f"k must be a non-negative integer, got {k!r}")
    opts = _options(metric, precision)
    handle = _handle_or_none(context)
    n, d = x.shape
    res = _native.TuneResult()
    status = _native.load().scrust_hip_tune(
        handle, x.ctypes.data, n, d, int(k), ctypes.byref(opts),
        _native.TUNE_SKIP_KNOWN if skip_known else 0, ctypes.byref(res))
    _native.raise_for(status)
    return res.as_dict()


def tune_device(embedding_ptr: int, n: int, d: int, k: int, *, metric: str = "euclidean",
                precision: str = "bf16", skip_known: bool = False,
                context: Context | None = None) -> dict:
    """``tune`` on a device-resident (n, d) float32 embedding (a raw device pointer)."""
    opts = _options(metric, precision)
    ctx = context if context is not None else _thread_context()
    res = _native.TuneResult()
    status = _native.load().scrust_hip_tune_device(
        ctx.handle, ctypes.c_void_p(embedding_ptr), n, d, k, ctypes.byref(opts),
        _native.TUNE_SKIP_KNOWN if skip_known else 0, ctypes.byref(res))
    _native.raise_for(status)
    return res.as_dict()


def tune_save(path) -> None:
    """Writes the tuning table to a text file (one bucket a line)."""
    _native.raise_for(_native.load().scrust_hip_tune_save(str(path).encode()))


def tune_load(path) -> None:
    """Merges a table written by ``tune_save``; a malformed file raises ValueError and
    leaves the table unchanged."""
    _native.raise_for(_native.load().scrust_hip_tune_load(str(path).encode()))


def tune_clear() -> None:
    """Empties the tuning table (automatic shapes return to the plan's defaults)."""
    _native.load().scrust_hip_tune_clear()


def tune_size() -> int:
    """Buckets in the tuning table."""
    return int(_native.load().scrust_hip_tune_size())


def _graph_array(a, dtype, name):
    """A 2-D array of `dtype` (convert.rs:70-93: a wrong dtype or rank is a ValueError)."""
    if not isinstance(a, np.ndarray):
        raise ValueError(f"{name} must be a numpy array")
    if a.dtype != dtype:
        raise ValueError(f"{name} must have dtype {np.dtype(dtype).name}, got {a.dtype}")
    if a.ndim != 2:
        raise ValueError(f"{name} must be 2-dimensional, got {a.ndim} dimensions")
    return np.ascontiguousarray(a)


class Batch:
    """Many embeddings in one persistent MFMA launch, captured once in a HIP graph.

    ``datasets`` is a sequence of (embedding_ptr, n, d, k, indices_ptr, distances_ptr)
    device pointers (e.g. from torch tensors). The pointers are bound into the graph:
    their contents may change between ``run`` calls, the pointers may not. No
    counterpart in the reference, whose kNN takes one embedding per call.
    """

    def __init__(self, datasets, *, metric: str = "euclidean", precision: str = "bf16",
                 context: Context | None = None, tile: int = 0):
        self._lib = _native.load()
        self._context = context if context is not None else _thread_context()
        arr = (_native.Dataset * len(datasets))()
        for i, (x, n, d, k, ids, dist) in enumerate(datasets):
            arr[i] = _native.Dataset(ctypes.c_void_p(x), n, d, k, ctypes.c_void_p(ids),
                                     ctypes.c_void_p(dist))
        opts = _options(metric, precision, tile)
        handle = ctypes.c_void_p()
        _native.raise_for(self._lib.scrust_hip_batch_create(
            self._context.handle, arr, len(datasets), ctypes.byref(opts), ctypes.byref(handle)))
        self._handle = handle

    def run(self) -> dict:
        st = _native.BatchStats()
        _native.raise_for(self._lib.scrust_hip_batch_run(self._handle, ctypes.byref(st)))
        return st.as_dict()

    def tune(self) -> dict:
        """Times the batch under every work shape (longest reference range per work item,
        candidate buffer), checks that every dataset's result stays the same, and keeps the
        fastest graph. Overwrites the outputs (the graph is replayed four times per shape)."""
        res = _native.BatchTuneResult()
        _native.raise_for(self._lib.scrust_hip_batch_tune(self._handle, ctypes.byref(res)))
        return res.as_dict()

    def close(self) -> None:
        if getattr(self, "_handle", None) is not None:
            self._lib.scrust_hip_batch_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """A kNN request loop on the device: builds are admitted in arrival order as slots
    and device memory free up, and admitted builds run concurrently (one HIP stream per
    slot). ``submit`` returns a ticket at once; ``wait`` returns the result.

    No counterpart in the reference (its kNN is one synchronous call).
    """

    def __init__(self, device: int = 0, *, slots: int = 4, memory_bytes: int = 0):
        self._lib = _native.load()
        cfg = _native.ServerConfig(int(slots), 0, int(memory_bytes))
        handle = ctypes.c_void_p()
        _native.raise_for(self._lib.scrust_hip_server_create(int(device), ctypes.byref(cfg),
                                                             ctypes.byref(handle)))
        self._handle = handle
        self._pending = {}

    def submit(self, embedding, k: int, *, metric: str = "euclidean",
               precision: str = "bf16") -> int:
        x = _embedding(embedding)
        n, d = x.shape
        ids = np.empty((n, int(k)), dtype=np.uint32)
        dist = np.empty((n, int(k)), dtype=np.float32)
        opts = _options(metric, precision)
        ticket = ctypes.c_uint64()
        _native.raise_for(self._lib.scrust_hip_server_submit(
            self._handle, x.ctypes.data, n, d, int(k), ctypes.byref(opts), ids.ctypes.data,
            dist.ctypes.data, ctypes.byref(ticket)))
        # Keep the buffers alive until the build is done.
        self._pending[ticket.value] = (x, ids, dist)
        return ticket.value

    def submit_device(self, embedding_ptr: int, n: int, d: int, k: int, *, metric: str = "euclidean",
                      precision: str = "bf16", keep_alive=None) -> int:
        """Submits a device-resident embedding (raw device pointer on the server's device);
        results come back to host arrays. ``keep_alive`` (e.g. the owning tensor) is held
        until the ticket is waited for."""
        ids = np.empty((int(n), int(k)), dtype=np.uint32)
        dist = np.empty((int(n), int(k)), dtype=np.float32)
        opts = _options(metric, precision)
        ticket = ctypes.c_uint64()
        _native.raise_for(self._lib.scrust_hip_server_submit_device(
            self._handle, ctypes.c_void_p(embedding_ptr), int(n), int(d), int(k), ctypes.byref(opts),
            ids.ctypes.data, dist.ctypes.data, ctypes.byref(ticket)))
        self._pending[ticket.value] = (keep_alive, ids, dist)
        return ticket.value

    def wait(self, ticket: int, timeout_ms: int = -1):
        """(ids, distances, info) of a ticket; info holds the request's timestamps."""
        info = _native.RequestInfo()
        status = self._lib.scrust_hip_server_wait(self._handle, int(ticket), int(timeout_ms),
                                                  ctypes.byref(info))
        if status != _native.OK and "timed out" in _native.last_error():
            _native.raise_for(status)
        _, ids, dist = self._pending.pop(int(ticket))
        _native.raise_for(status)
        return ids, dist, info.as_dict()

    def stats(self) -> dict:
        st = _native.ServerStats()
        _native.raise_for(self._lib.scrust_hip_server_stats_get(self._handle, ctypes.byref(st)))
        return st.as_dict()

    def close(self) -> None:
        if getattr(self, "_handle", None) is not None:
            self._lib.scrust_hip_server_destroy(self._handle)
            self._handle = None
            self._pending.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
