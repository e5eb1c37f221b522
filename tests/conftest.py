"""Shared test setup: the `gpu` marker and the device context fixture.

GPU tests FAIL (never skip) when no usable gfx950 device is present: a run that asked
for `-m gpu` must not read a missing device as a pass.
"""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

try:  # one HIP runtime for torch and libscrust_hip.so (torch ships its own)
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct MI355X (gfx950)")


@pytest.fixture(scope="session")
def context():
    import mrustsc_amd

    if not mrustsc_amd.gpu_available():
        pytest.fail("a gfx950 GPU was required for this test but none is usable")
    ctx = mrustsc_amd.Context(0)
    yield ctx
    ctx.close()
