"""The C1 workload's graph as a committed fixture (tests/golden/c1_oracle_fixture.json, written
by tests/golden/make_c1_fixture.py): the generator, the oracle (CPU) and the HIP path (GPU) must
each reproduce it bit for bit. Beyond regression, the duplicate rows pin the snapping: each of
the 20 copied rows and its source are mutual neighbours at distance exactly 0."""

from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import mrustsc_amd
import oracle
from mrustsc_amd import synthetic

FIXTURE = json.loads((Path(__file__).parent / "golden" / "c1_oracle_fixture.json").read_text())


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def check_graph(ids: np.ndarray, dist: np.ndarray) -> None:
    for r, row in FIXTURE["rows"].items():
        np.testing.assert_array_equal(ids[int(r)], np.array(row["ids"], dtype=np.uint32), err_msg=f"row {r}")
        bits = np.array([int(b, 16) for b in row["dist_bits"]], dtype=np.uint32)
        np.testing.assert_array_equal(dist[int(r)].view(np.uint32), bits, err_msg=f"row {r}")
    assert digest(ids.astype(np.uint32)) == FIXTURE["ids_sha256"]
    assert digest(dist.astype(np.float32)) == FIXTURE["dist_sha256"]


def test_generator_reproduces_the_c1_embedding():
    x = synthetic.pbmc_like(0)
    assert list(x.shape) == FIXTURE["shape"]
    assert digest(x) == FIXTURE["x_sha256"]


def test_oracle_reproduces_the_fixture_and_duplicates_sit_at_zero():
    x = synthetic.pbmc_like(0)
    ids, dist = oracle.knn(x, FIXTURE["k"])
    check_graph(ids, dist)
    # The 20 duplicated rows: find them from the data, not from the graph.
    _, inverse, counts = np.unique(x, axis=0, return_inverse=True, return_counts=True)
    dup_rows = np.nonzero(counts[inverse.ravel()] > 1)[0]
    assert len(dup_rows) == 40 and (counts == 2).sum() == 20 and counts.max() == 2
    for r in dup_rows:
        twin = [q for q in dup_rows if q != r and np.array_equal(x[q], x[r])][0]
        assert ids[r][0] == twin and dist[r][0] == 0.0
    assert (dist == 0).sum() == FIXTURE["zero_distances"] == 40


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["bf16", "fp32", "exact"])
def test_hip_path_reproduces_the_fixture(context, precision):
    print(f"[case] C1 fixture {precision}", flush=True)
    x = synthetic.pbmc_like(0)
    ids, dist = mrustsc_amd.knn_hip(x, FIXTURE["k"], precision=precision, context=context)
    check_graph(ids, dist)
