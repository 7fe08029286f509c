"""Writes tests/golden/cloud128_seed1.json: the checksum and summary figures of
the 128^3 procedural cloud (Perlin fBm seed 1, 5 octaves, base frequency 4,
lacunarity 2, gain 0.5, threshold 0.4), the small version of the 512^3
benchmark cloud.  The remap and falloff are this project's own choices; the
fixture fixes them, so a change shows up in the tests.

Run from the repository root:  python tests/golden/make_cloud_fixture.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def summary(grid):
    g = np.ascontiguousarray(grid, np.float32)
    idx = [(0, 0, 0), (64, 64, 64), (40, 70, 90), (100, 30, 60), (64, 20, 64), (17, 83, 51)]
    return {
        "sha256": hashlib.sha256(g.tobytes()).hexdigest(),
        "sum": float(g.astype(np.float64).sum()),
        "max": float(g.max()),
        "nonzero": int(np.count_nonzero(g)),
        "samples": {"%d,%d,%d" % i: float(g[i]) for i in idx},
    }


def main():
    import oracle
    rec = {"about": "oracle.generate_cloud(128, seed=1) with the benchmark cloud's defaults "
                    "(octaves 5, base_freq 4, lacunarity 2, gain 0.5, threshold 0.4); FP32, z fastest",
           "res": 128, "seed": 1}
    rec.update(summary(oracle.generate_cloud(128, 1)))
    path = os.path.join(ROOT, "tests", "golden", "cloud128_seed1.json")
    with open(path, "w") as f:
        json.dump(rec, f, indent=1)
    print("wrote", path, rec["sha256"])


if __name__ == "__main__":
    main()
