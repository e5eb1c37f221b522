"""ctypes binding of libscrust_hip.so (the C ABI in include/scrust_hip.h).

The library is built in-tree (``mrustsc_amd/lib/libscrust_hip.so``); importing this
module fails loudly when it is missing -- there is no fallback implementation.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libscrust_hip.so"

OK = 0
SHAPE = 1
INVALID_PARAMETER = 2
NO_GPU = 3
KERNEL = 4
OUT_OF_MEMORY = 5

EUCLIDEAN = 0
COSINE = 1
METRICS = {"euclidean": EUCLIDEAN, "cosine": COSINE}

BF16 = 0
FP32 = 1
EXACT = 2
FP16 = 3
PRECISIONS = {"bf16": BF16, "fp32": FP32, "exact": EXACT, "fp16": FP16}
PRECISION_NAMES = {v: k for k, v in PRECISIONS.items()}

# Every entry point include/scrust_hip.h declares.
EXPORTS = (
    "scrust_hip_last_error",
    "scrust_hip_gpu_available",
    "scrust_hip_context_create",
    "scrust_hip_context_destroy",
    "scrust_hip_max_supported_k",
    "scrust_hip_knn",
    "scrust_hip_knn_ex",
    "scrust_hip_knn_device",
    "scrust_hip_knn_device_noalloc",
    "scrust_hip_connectivities",
    "scrust_hip_connectivities_device",
    "scrust_hip_neighbors",
    "scrust_hip_neighbors_device",
    "scrust_hip_pca",
    "scrust_hip_pca_device",
    "scrust_hip_tsne_squared_distances",
    "scrust_hip_tsne_squared_distances_device",
    "scrust_hip_tune",
    "scrust_hip_tune_device",
    "scrust_hip_tune_save",
    "scrust_hip_tune_load",
    "scrust_hip_tune_clear",
    "scrust_hip_tune_size",
    "scrust_hip_batch_create",
    "scrust_hip_batch_run",
    "scrust_hip_batch_tune",
    "scrust_hip_batch_destroy",
    "scrust_hip_server_create",
    "scrust_hip_server_submit",
    "scrust_hip_server_submit_device",
    "scrust_hip_server_wait",
    "scrust_hip_server_stats_get",
    "scrust_hip_server_destroy",
)


class KnnOptions(ctypes.Structure):
    _fields_ = [
        ("metric", ctypes.c_uint32),
        ("precision", ctypes.c_uint32),
        ("tile", ctypes.c_uint32),
        ("order", ctypes.c_uint32),
    ]


class KnnStats(ctypes.Structure):
    _fields_ = [
        ("prepare_ms", ctypes.c_double),
        ("pass_ms", ctypes.c_double),
        ("rerank_ms", ctypes.c_double),
        ("fallback_ms", ctypes.c_double),
        ("device_ms", ctypes.c_double),
        ("fallback_rows", ctypes.c_uint64),
        ("candidates", ctypes.c_uint64),
        ("pass_flops", ctypes.c_uint64),
        ("path", ctypes.c_uint32),
        ("tile", ctypes.c_uint32),
        ("splits", ctypes.c_uint32),
        ("order", ctypes.c_uint32),
        ("tuned", ctypes.c_uint32),
        ("tail_blocks", ctypes.c_uint32),
        ("tail_splits", ctypes.c_uint32),
        ("workspace_bytes", ctypes.c_uint64),
        ("persistent", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("pairs_scored", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        out = {name: getattr(self, name) for name, _ in self._fields_}
        out["path"] = PRECISION_NAMES.get(out["path"], out["path"])
        return out


TUNE_SKIP_KNOWN = 1


class TuneResult(ctypes.Structure):
    _fields_ = [
        ("tile", ctypes.c_uint32),
        ("splits", ctypes.c_uint32),
        ("order", ctypes.c_uint32),
        ("default_tile", ctypes.c_uint32),
        ("default_splits", ctypes.c_uint32),
        ("default_order", ctypes.c_uint32),
        ("shapes", ctypes.c_uint32),
        ("reused", ctypes.c_uint32),
        ("best_ms", ctypes.c_double),
        ("default_ms", ctypes.c_double),
        ("worst_ms", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Dataset(ctypes.Structure):
    _fields_ = [
        ("embedding", ctypes.c_void_p),
        ("n", ctypes.c_size_t),
        ("d", ctypes.c_size_t),
        ("k", ctypes.c_size_t),
        ("out_indices", ctypes.c_void_p),
        ("out_distances", ctypes.c_void_p),
    ]


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("device_ms", ctypes.c_double),
        ("datasets", ctypes.c_uint64),
        ("pass_items", ctypes.c_uint64),
        ("exact_rows", ctypes.c_uint64),
        ("pass_flops", ctypes.c_uint64),
        ("cells", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchTuneResult(ctypes.Structure):
    _fields_ = [
        ("cut", ctypes.c_uint32),
        ("tile", ctypes.c_uint32),
        ("default_cut", ctypes.c_uint32),
        ("default_tile", ctypes.c_uint32),
        ("shapes", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("best_ms", ctypes.c_double),
        ("default_ms", ctypes.c_double),
        ("worst_ms", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("slots", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("memory_bytes", ctypes.c_uint64),
    ]


class RequestInfo(ctypes.Structure):
    _fields_ = [
        ("submit_ns", ctypes.c_int64),
        ("admit_ns", ctypes.c_int64),
        ("done_ns", ctypes.c_int64),
        ("fallback_rows", ctypes.c_uint64),
        ("slot", ctypes.c_uint32),
        ("status", ctypes.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("rejected", ctypes.c_uint64),
        ("peak_in_flight", ctypes.c_uint64),
        ("memory_bytes", ctypes.c_uint64),
        ("peak_memory_bytes", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class ScrustHipError(RuntimeError):
    """A failure reported by the library that is not an argument error."""


_lib = None


def _share_hip_runtime_with_torch() -> None:
    """PyTorch ships its own libamdhip64. Loading it before this library makes the
    process use one HIP runtime, so device pointers can pass between torch tensors and
    this library; loaded afterwards, torch would find the device already claimed by a
    second runtime."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    _share_hip_runtime_with_torch()
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C mrustsc_amd/csrc`"
        )
    lib = ctypes.CDLL(str(LIB_PATH))
    vp = ctypes.c_void_p
    sz = ctypes.c_size_t
    lib.scrust_hip_last_error.argtypes = []
    lib.scrust_hip_last_error.restype = ctypes.c_char_p
    lib.scrust_hip_gpu_available.argtypes = []
    lib.scrust_hip_gpu_available.restype = ctypes.c_int
    lib.scrust_hip_context_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    lib.scrust_hip_context_create.restype = ctypes.c_int
    lib.scrust_hip_context_destroy.argtypes = [vp]
    lib.scrust_hip_context_destroy.restype = None
    lib.scrust_hip_max_supported_k.argtypes = [vp]
    lib.scrust_hip_max_supported_k.restype = sz
    lib.scrust_hip_knn.argtypes = [vp, vp, sz, sz, sz, vp, vp]
    lib.scrust_hip_knn.restype = ctypes.c_int
    opts = ctypes.POINTER(KnnOptions)
    stats = ctypes.POINTER(KnnStats)
    lib.scrust_hip_knn_ex.argtypes = [vp, vp, sz, sz, sz, opts, vp, vp, stats]
    lib.scrust_hip_knn_ex.restype = ctypes.c_int
    lib.scrust_hip_knn_device.argtypes = [vp, vp, sz, sz, sz, opts, vp, vp, stats]
    lib.scrust_hip_knn_device.restype = ctypes.c_int
    lib.scrust_hip_pca.argtypes = [vp, vp, vp, vp, sz, sz, sz, ctypes.c_int, ctypes.c_uint64, vp, vp, vp, vp]
    lib.scrust_hip_pca.restype = ctypes.c_int
    lib.scrust_hip_pca_device.argtypes = [vp, vp, vp, vp, sz, sz, sz, ctypes.c_int, ctypes.c_uint64, vp, vp, vp, vp]
    lib.scrust_hip_pca_device.restype = ctypes.c_int
    lib.scrust_hip_knn_device_noalloc.argtypes = [vp, vp, sz, sz, sz, opts, vp, vp, stats]
    lib.scrust_hip_knn_device_noalloc.restype = ctypes.c_int
    lib.scrust_hip_connectivities.argtypes = [vp, vp, vp, sz, sz, vp, vp, vp, sz, ctypes.POINTER(sz)]
    lib.scrust_hip_connectivities.restype = ctypes.c_int
    lib.scrust_hip_connectivities_device.argtypes = [vp, vp, vp, sz, sz, vp, vp, vp, sz, ctypes.POINTER(sz)]
    lib.scrust_hip_connectivities_device.restype = ctypes.c_int
    lib.scrust_hip_tsne_squared_distances.argtypes = [vp, vp, sz, sz, vp]
    lib.scrust_hip_tsne_squared_distances.restype = ctypes.c_int
    lib.scrust_hip_tsne_squared_distances_device.argtypes = [vp, vp, sz, sz, vp]
    lib.scrust_hip_tsne_squared_distances_device.restype = ctypes.c_int
    lib.scrust_hip_neighbors.argtypes = [vp, vp, sz, sz, sz, opts, vp, vp, vp, vp, vp, sz, ctypes.POINTER(sz)]
    lib.scrust_hip_neighbors.restype = ctypes.c_int
    lib.scrust_hip_neighbors_device.argtypes = [vp, vp, sz, sz, sz, opts, vp, vp, vp, vp, vp, sz,
                                                ctypes.POINTER(sz)]
    lib.scrust_hip_neighbors_device.restype = ctypes.c_int
    tune = ctypes.POINTER(TuneResult)
    lib.scrust_hip_tune.argtypes = [vp, vp, sz, sz, sz, opts, ctypes.c_uint32, tune]
    lib.scrust_hip_tune.restype = ctypes.c_int
    lib.scrust_hip_tune_device.argtypes = [vp, vp, sz, sz, sz, opts, ctypes.c_uint32, tune]
    lib.scrust_hip_tune_device.restype = ctypes.c_int
    lib.scrust_hip_tune_save.argtypes = [ctypes.c_char_p]
    lib.scrust_hip_tune_save.restype = ctypes.c_int
    lib.scrust_hip_tune_load.argtypes = [ctypes.c_char_p]
    lib.scrust_hip_tune_load.restype = ctypes.c_int
    lib.scrust_hip_tune_clear.argtypes = []
    lib.scrust_hip_tune_clear.restype = None
    lib.scrust_hip_tune_size.argtypes = []
    lib.scrust_hip_tune_size.restype = sz
    lib.scrust_hip_batch_create.argtypes = [vp, ctypes.POINTER(Dataset), sz, opts, ctypes.POINTER(vp)]
    lib.scrust_hip_batch_create.restype = ctypes.c_int
    lib.scrust_hip_batch_run.argtypes = [vp, ctypes.POINTER(BatchStats)]
    lib.scrust_hip_batch_run.restype = ctypes.c_int
    lib.scrust_hip_batch_tune.argtypes = [vp, ctypes.POINTER(BatchTuneResult)]
    lib.scrust_hip_batch_tune.restype = ctypes.c_int
    lib.scrust_hip_batch_destroy.argtypes = [vp]
    lib.scrust_hip_batch_destroy.restype = None
    lib.scrust_hip_server_create.argtypes = [ctypes.c_int, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
    lib.scrust_hip_server_create.restype = ctypes.c_int
    lib.scrust_hip_server_submit.argtypes = [vp, vp, sz, sz, sz, opts, vp, vp, ctypes.POINTER(ctypes.c_uint64)]
    lib.scrust_hip_server_submit.restype = ctypes.c_int
    lib.scrust_hip_server_submit_device.argtypes = [vp, vp, sz, sz, sz, opts, vp, vp, ctypes.POINTER(ctypes.c_uint64)]
    lib.scrust_hip_server_submit_device.restype = ctypes.c_int
    lib.scrust_hip_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.c_int64, ctypes.POINTER(RequestInfo)]
    lib.scrust_hip_server_wait.restype = ctypes.c_int
    lib.scrust_hip_server_stats_get.argtypes = [vp, ctypes.POINTER(ServerStats)]
    lib.scrust_hip_server_stats_get.restype = ctypes.c_int
    lib.scrust_hip_server_destroy.argtypes = [vp]
    lib.scrust_hip_server_destroy.restype = None
    _lib = lib
    return lib


def last_error() -> str:
    return load().scrust_hip_last_error().decode("utf-8", "replace")


def raise_for(status: int) -> None:
    """Map a status onto the exception the reference's bindings raise
    (crates/scrust-py/src/lib.rs:24-35): Shape / InvalidParameter -> ValueError,
    everything else -> RuntimeError."""
    if status == OK:
        return
    message = last_error()
    if status in (SHAPE, INVALID_PARAMETER):
        raise ValueError(message)
    raise ScrustHipError(f"[status {status}] {message}")
