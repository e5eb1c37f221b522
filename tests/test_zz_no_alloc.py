"""The allocation-free entry point (scrust_hip_knn_device_noalloc): through the C ABI its
results equal scrust_hip_knn_device's bit for bit, on calls in a row whose second is smaller
(the scratch array keeps the first call's contents), and it allocates no device memory.
The first build's default work shape does not fit the scratch array; its compact shape does."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import mrustsc_amd  # noqa: E402
from mrustsc_amd import _native, synthetic  # noqa: E402

pytestmark = pytest.mark.gpu


def _call(fn, ctx, dx, n, d, k, precision, metric="euclidean"):
    ids = torch.full((n, k), -1, dtype=torch.int32, device="cuda")
    dist = torch.full((n, k), -1.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    opts = _native.KnnOptions(_native.METRICS[metric], _native.PRECISIONS[precision], 0, 0)
    st = _native.KnnStats()
    status = fn(ctx.handle, ctypes.c_void_p(dx.data_ptr()), n, d, k, ctypes.byref(opts),
                ctypes.c_void_p(ids.data_ptr()), ctypes.c_void_p(dist.data_ptr()), ctypes.byref(st))
    _native.raise_for(status)
    return ids.cpu().numpy().view(np.uint32), dist.cpu().numpy(), st.as_dict()


@pytest.mark.parametrize("precision,metric", [("bf16", "euclidean"), ("fp32", "euclidean"),
                                              ("bf16", "cosine"), ("exact", "euclidean")])
def test_noalloc_equals_the_allocating_entry_point_on_two_calls_in_a_row(precision, metric):
    lib = _native.load()
    with mrustsc_amd.Context(0) as ctx:
        # The first build needs the compact work shape to fit; the second is smaller.
        for n, d, k, seed in ((60000, 50, 15, 61), (3000, 30, 30, 62)):
            print(f"[case] noalloc n={n} d={d} k={k} {precision} {metric}", flush=True)
            x = synthetic.clustered(n, d, seed=seed, n_clusters=8)
            dx = torch.from_numpy(x).to("cuda")
            got = _call(lib.scrust_hip_knn_device_noalloc, ctx, dx, n, d, k, precision, metric)
            want = _call(lib.scrust_hip_knn_device, ctx, dx, n, d, k, precision, metric)
            np.testing.assert_array_equal(got[0], want[0])
            np.testing.assert_array_equal(got[1], want[1])
            assert got[2]["path"] == want[2]["path"] and got[2]["fallback_rows"] == want[2]["fallback_rows"]


def test_noalloc_allocates_no_device_memory():
    lib = _native.load()
    x = synthetic.clustered(16000, 50, seed=63)
    dx = torch.from_numpy(x).to("cuda")
    with mrustsc_amd.Context(0) as ctx:
        _call(lib.scrust_hip_knn_device_noalloc, ctx, dx, 16000, 50, 15, "bf16")  # warm-up
        ids = torch.empty((16000, 15), dtype=torch.int32, device="cuda")
        dist = torch.empty((16000, 15), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info()
        for _ in range(3):
            mrustsc_amd.knn_device_noalloc(dx.data_ptr(), 16000, 50, 15, ids.data_ptr(), dist.data_ptr(),
                                           context=ctx)
        free1, _ = torch.cuda.mem_get_info()
        # The allocating entry point's workspace for this build is tens of MB.
        assert free0 - free1 < (2 << 20), (free0, free1)
        w_ids, w_dist = mrustsc_amd.knn_hip(x, 15, context=ctx)
        np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32), w_ids)
        np.testing.assert_array_equal(dist.cpu().numpy(), w_dist)


def test_noalloc_refuses_a_build_larger_than_its_scratch():
    x = synthetic.clustered(200000, 50, seed=64)
    dx = torch.from_numpy(x).to("cuda")
    ids = torch.empty((200000, 15), dtype=torch.int32, device="cuda")
    dist = torch.empty((200000, 15), dtype=torch.float32, device="cuda")
    with mrustsc_amd.Context(0) as ctx:
        with pytest.raises(ValueError, match="scratch array"):
            mrustsc_amd.knn_device_noalloc(dx.data_ptr(), 200000, 50, 15, ids.data_ptr(), dist.data_ptr(),
                                           context=ctx)
        # The context still serves calls that fit.
        small = torch.from_numpy(synthetic.clustered(1000, 50, seed=65)).to("cuda")
        mrustsc_amd.knn_device_noalloc(small.data_ptr(), 1000, 50, 15, ids.data_ptr(), dist.data_ptr(),
                                       context=ctx)
