"""Writes tests/golden/golden.json: counts and SHA-256 digests of the oracle's
output for fixed workloads (regression anchors for the CPU oracle and for the
HIP path; the reference ships no golden vectors for this path).

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402


def digest(r):
    h = hashlib.sha256()
    h.update(r["quads"].astype(np.int32).tobytes())
    h.update(r["vertices"].astype(np.float32).view(np.uint32).tobytes())
    return h.hexdigest()


def cases():
    yield "sphere64", oracle.field_sphere((64, 64, 64), (32.0, 32.0, 32.0), 20.0), 0.0
    yield "blobs_96x80x72_s1", oracle.field_blobs((96, 80, 72), 1, 12, sigma=(3.0, 12.0)), 0.1
    yield "gyroid_77x50x41_s5", oracle.field_gyroid((77, 50, 41), 8.0, 5, 0.8), 0.15
    yield "blobs_260x33x20_s9", oracle.field_blobs((260, 33, 20), 9, 10, sigma=(2.0, 9.0)), 0.05


def main():
    out = {}
    for name, f, iso in cases():
        r = oracle.run(f, iso)
        out[name] = {"iso": iso, "dims": [int(f.shape[2]), int(f.shape[1]), int(f.shape[0])],
                     "num_active": r["num_active"], "num_vertices": r["num_vertices"], "num_quads": r["num_quads"],
                     "num_lut2": int(((r["info"] >> 7) & 1).sum()), "num_tie_edges": r["num_tie_edges"],
                     "sha256": digest(r)}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print(path)


if __name__ == "__main__":
    main()
