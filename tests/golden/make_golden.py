"""Writes the exact-rational golden vertex sets under tests/golden/ (run once, commit the JSON).

Each fixture holds the generators (float.hex, exact), the box [0,1]^d with the
reference's plane order, and every vertex of the bounded Voronoi tessellation as
(sorted 1-based key, exact rational coordinates) from oracle/exact_vertices.py.
Cell volumes come from Qhull (scipy ConvexHull of each cell's vertices) and are
stored for the integrator tests.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "highvoronoi.jl_amd"))

from exact_vertices import exact_vertices  # noqa: E402
from hvgpu import synth  # noqa: E402

CASES = [
    # name, d, n, seed
    ("d2_n5", 2, 5, 101),
    ("d2_n30", 2, 30, 11),
    ("d3_n5", 3, 5, 102),
    ("d3_n16", 3, 16, 12),
    ("d4_n6", 4, 6, 103),
    ("d4_n12", 4, 12, 13),
    ("d5_n7", 5, 7, 104),
    ("d5_n9", 5, 9, 14),
    ("d6_n8", 6, 8, 105),
]


def cell_volumes(n, d, keys, coords):
    from scipy.spatial import ConvexHull
    vols = []
    for i in range(1, n + 1):
        pts = np.array([c for k, c in zip(keys, coords) if i in k])
        vols.append(float(ConvexHull(pts).volume))
    return vols


def main(only=None):
    for name, d, n, seed in CASES:
        if only and name not in only:
            continue
        xs = synth.uniform(d, n, seed)
        t0 = time.time()
        keys, coords = exact_vertices(xs)
        fc = [[float(v) for v in c] for c in coords]
        vols = cell_volumes(n, d, keys, fc)
        out = {
            "d": d, "n": n, "seed": seed,
            "generator": f"hvgpu.synth.uniform({d}, {n}, {seed})",
            "points_hex": [[float(v).hex() for v in row] for row in xs],
            "boundary": "cuboid(d, periodic=[]): plane 2k-1 is x_k = 1, plane 2k is x_k = 0 (1-based)",
            "keys": [list(k) for k in keys],
            "coords_exact": [[f"{v.numerator}/{v.denominator}" for v in c] for c in coords],
            "coords": fc,
            "volumes_qhull": vols,
        }
        with open(os.path.join(HERE, f"{name}.json"), "w") as f:
            json.dump(out, f)
        print(f"{name}: {len(keys)} vertices, sum(vol)={sum(vols):.15f}, {time.time() - t0:.1f}s")


if __name__ == "__main__":
    main(sys.argv[1:])
