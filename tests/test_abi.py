"""The C ABI on the CPU: the library loads, exports every declared entry point, and
checks arguments (mirroring knn_metal's validation, crates/scrust-gpu/src/kernels/knn.rs:37-59)
before it needs a device. No compute calls are made here."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import mrustsc_amd
from mrustsc_amd import _native, synthetic

HEADER = Path(__file__).resolve().parents[1] / "include" / "scrust_hip.h"


def declared_functions() -> set[str]:
    text = HEADER.read_text()
    return set(re.findall(r"\b(scrust_hip_[a-z0-9_]+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    declared = declared_functions()
    assert declared, "no declarations found"
    assert declared == set(_native.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name


def test_max_supported_k_is_reported_without_a_device():
    assert mrustsc_amd.max_supported_k() == 256


def _call(n, d, k, *, embedding=True, outputs=True, metric=0, precision=0):
    lib = _native.load()
    x = np.zeros((max(n, 1), max(d, 1)), dtype=np.float32)
    ids = np.zeros(max(n * k, 1), dtype=np.uint32)
    dist = np.zeros(max(n * k, 1), dtype=np.float32)
    opts = _native.KnnOptions(metric, precision, 0, 0)
    return lib.scrust_hip_knn_ex(None, x.ctypes.data if embedding else None, n, d, k,
                                 ctypes.byref(opts), ids.ctypes.data if outputs else None,
                                 dist.ctypes.data if outputs else None, None)


@pytest.mark.parametrize(("n", "d", "k"), [(8, 3, 0), (8, 3, 8), (8, 3, 9), (8, 0, 3), (0, 4, 3),
                                           (4096, 2, 257)])
def test_degenerate_arguments_are_invalid_parameters(n, d, k):
    assert _call(n, d, k) == _native.INVALID_PARAMETER
    assert " must be " in _native.last_error()


def test_null_pointers_and_unknown_options_are_rejected():
    assert _call(8, 3, 2, embedding=False) == _native.INVALID_PARAMETER
    assert _call(8, 3, 2, outputs=False) == _native.INVALID_PARAMETER
    assert _call(8, 3, 2, metric=7) == _native.INVALID_PARAMETER
    assert _call(8, 3, 2, precision=9) == _native.INVALID_PARAMETER


def test_valid_arguments_without_a_context_report_no_gpu():
    assert _call(8, 3, 2) == _native.NO_GPU
    assert _native.last_error()


@pytest.mark.skipif(mrustsc_amd.gpu_available(), reason="checks the no-device path")
def test_context_creation_without_a_device_fails_with_no_gpu():
    with pytest.raises(mrustsc_amd.ScrustHipError, match="status 3"):
        mrustsc_amd.Context(0)
    assert mrustsc_amd.gpu_available() is False


def test_python_boundary_checks_dtype_rank_and_device():
    x = synthetic.random_embedding(10, 3, 1)
    with pytest.raises(ValueError):
        mrustsc_amd.knn(x.astype(np.float64), 3)
    with pytest.raises(ValueError):
        mrustsc_amd.knn(x[:, 0], 3)
    with pytest.raises(ValueError):
        mrustsc_amd.knn(x, 3, device="cuda")
    with pytest.raises(ValueError):
        mrustsc_amd.knn(x, 3, device="cpu")
    with pytest.raises(ValueError):
        mrustsc_amd.knn(x, 0)
    with pytest.raises(ValueError):
        mrustsc_amd.knn_hip(x, 3, metric="manhattan")


def test_status_codes_map_to_the_reference_exceptions():
    with pytest.raises(ValueError):
        _native.raise_for(_native.SHAPE)
    with pytest.raises(ValueError):
        _native.raise_for(_native.INVALID_PARAMETER)
    for status in (_native.NO_GPU, _native.KERNEL, _native.OUT_OF_MEMORY):
        with pytest.raises(RuntimeError):
            _native.raise_for(status)


# --- generators restated from the reference's tests ------------------------------------------


def splitmix64_stream(seed: int, count: int) -> list[int]:
    mask = (1 << 64) - 1
    state, out = seed, []
    for _ in range(count):
        state = (state + 0x9E3779B97F4A7C15) & mask
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        out.append(z ^ (z >> 31))
    return out


def test_splitmix64_matches_its_published_vectors():
    # Published outputs of splitmix64 seeded with 1234567.
    assert splitmix64_stream(1234567, 3) == [6457827717110365317, 3203168211198807973,
                                             9817491932198370423]


def test_reference_embedding_is_the_splitmix64_stream():
    x = synthetic.reference_embedding(7, 5, 1234567)
    want = np.array([np.float32(z >> 40) / np.float32(1 << 24)
                     for z in splitmix64_stream(1234567, 35)], dtype=np.float32).reshape(7, 5)
    np.testing.assert_array_equal(x, want)


def test_random_embedding_is_the_lcg_stream():
    x = synthetic.random_embedding(3, 4, 9)
    state, mask, want = 19, (1 << 64) - 1, []
    for _ in range(12):
        state = (state * 6364136223846793005 + 1442695040888963407) & mask
        want.append(np.float32(state >> 40) / np.float32(1 << 24))
    np.testing.assert_array_equal(x.ravel(), np.array(want, dtype=np.float32))


def test_mixtures_are_deterministic_and_shaped():
    a = synthetic.clustered(1000, 50, seed=1)
    b = synthetic.clustered(1000, 50, seed=1)
    np.testing.assert_array_equal(a, b)
    assert a.shape == (1000, 50) and a.dtype == np.float32
    assert len(np.unique(a, axis=0)) <= 1000 - 4  # 0.5% duplicate rows (some may collide)
    c1 = synthetic.pbmc_like()
    assert c1.shape == (2638, 50)
    specs = synthetic.batch_specs()
    assert len(specs) == 64
    assert all(500 <= s["n"] <= 50000 and s["d"] in (30, 50, 100) and s["k"] in (15, 30)
               for s in specs)


def test_a_missing_library_fails_loudly_with_no_fallback(tmp_path):
    # A fresh interpreter whose loader points at a path with no library: every entry point
    # raises ImportError naming the build step, instead of falling back to any other code.
    import subprocess
    import sys

    root = Path(__file__).resolve().parents[1]
    script = (
        "import sys; sys.path.insert(0, %r)\n"
        "from pathlib import Path\n"
        "import mrustsc_amd\n"
        "from mrustsc_amd import _native\n"
        "_native.LIB_PATH = Path(%r)\n"
        "_native._lib = None\n"
        "import numpy as np\n"
        "try:\n"
        "    mrustsc_amd.knn(np.zeros((8, 3), dtype=np.float32), 2)\n"
        "except ImportError as e:\n"
        "    print('raised:', e)\n"
        "    sys.exit(0)\n"
        "sys.exit(1)\n"
    ) % (str(root), str(tmp_path / "absent.so"))
    done = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, timeout=300)
    assert done.returncode == 0, done.stdout + done.stderr
    assert "is missing" in done.stdout and "build" in done.stdout


# C struct of include/scrust_hip.h -> its ctypes mirror in mrustsc_amd/_native.py.
STRUCTS = {
    "scrust_hip_knn_options": "KnnOptions",
    "scrust_hip_knn_stats": "KnnStats",
    "scrust_hip_tune_result": "TuneResult",
    "scrust_hip_dataset": "Dataset",
    "scrust_hip_batch_stats": "BatchStats",
    "scrust_hip_batch_tune_result": "BatchTuneResult",
    "scrust_hip_server_config": "ServerConfig",
    "scrust_hip_request_info": "RequestInfo",
    "scrust_hip_server_stats": "ServerStats",
}


def test_ctypes_structs_match_the_header_layout(tmp_path):
    # A C program built against the header prints every struct's size and each field's
    # offset; the ctypes mirrors must agree field by field (same names, same places).
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler on this machine")
    lines = ['#include <stddef.h>', '#include <stdio.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        cls = getattr(_native, pyname)
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-o", str(exe), str(src)], check=True, capture_output=True, text=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, field, value = line.split()
        got[(cname, field)] = int(value)
    for cname, pyname in STRUCTS.items():
        cls = getattr(_native, pyname)
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert got[(cname, field)] == getattr(cls, field).offset, f"{cname}.{field}"


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_header_compiles_as_c_and_cpp(tmp_path, lang):
    import shutil
    import subprocess

    cc = shutil.which("gcc" if lang == "c" else "g++")
    if cc is None:
        pytest.skip(f"no {lang} compiler on this machine")
    std = "-std=c99" if lang == "c" else "-std=c++11"
    done = subprocess.run([cc, "-fsyntax-only", "-Wall", "-Wextra", "-Werror", std, "-x", lang, str(HEADER)],
                          capture_output=True, text=True)
    assert done.returncode == 0, done.stderr


def test_bench_refuses_more_than_one_gpu():
    import json
    import subprocess
    import sys

    root = Path(__file__).resolve().parents[1]
    done = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True,
                          text=True, timeout=300, cwd=str(root))
    assert done.returncode == 2
    line = json.loads(done.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]
