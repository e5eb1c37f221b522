"""Writes c1_oracle_fixture.json: the C1 workload (2,638 x 50, 8 isotropic Gaussians with
centres ~ N(0, 25 I), spread 1, 20 exact duplicate rows; mrustsc_amd.synthetic.pbmc_like(0))
and the oracle's k = 15 graph on it, as SHA-256 digests of the raw arrays plus a few rows in
full (distances as float32 bit patterns).

    python tests/golden/make_c1_fixture.py
"""

from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from mrustsc_amd import synthetic  # noqa: E402

K = 15
ROWS = [0, 1, 2, 1000, 2637]


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main() -> None:
    x = synthetic.pbmc_like(0)
    ids, dist = oracle.knn(x, K)
    fixture = {
        "_source": "oracle k=15 graph of mrustsc_amd.synthetic.pbmc_like(0) (the C1 workload); "
                   "written by tests/golden/make_c1_fixture.py",
        "shape": list(x.shape),
        "k": K,
        "x_sha256": digest(x.astype(np.float32)),
        "ids_sha256": digest(ids.astype(np.uint32)),
        "dist_sha256": digest(dist.astype(np.float32)),
        "zero_distances": int((dist == 0).sum()),
        "rows": {str(r): {"ids": ids[r].astype(int).tolist(),
                          "dist_bits": [f"{b:08x}" for b in dist[r].astype(np.float32).view(np.uint32)]}
                 for r in ROWS},
    }
    out = Path(__file__).parent / "c1_oracle_fixture.json"
    out.write_text(json.dumps(fixture, indent=1) + "\n")
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
