"""CPU pinning of the periodic test reference: the oracle's walk over the originals' cells
of a tiled point set (tests/helpers.periodic_oracle) against Qhull (scipy.spatial.Voronoi) on a
larger tiling, cell by cell, and against the torus volume.

The device path meshes periodic boxes with images inside a margin and labels each generator
(original, shift); the GPU tests compare those labels and coordinates with periodic_oracle, so
this file checks that periodic_oracle itself gives the periodic tessellation.
"""
import numpy as np
import pytest
from scipy.spatial import ConvexHull, Voronoi

from helpers import periodic_oracle, periodic_tiling
from hvgpu import synth


@pytest.mark.parametrize("d,n,seed,per", [(2, 60, 1, (0, 1)), (2, 50, 2, (0,)), (3, 30, 3, (0, 1, 2))])
def test_periodic_oracle_matches_qhull_cells(d, n, seed, per):
    xs = synth.uniform(d, n, seed)
    ref = periodic_oracle(xs, per)
    # Qhull on images up to 2 lengths away, plus the reflections of all of them across the
    # Dirichlet faces (a reflected point never cuts a cell inside the box, and a point's own
    # reflection cuts its cell exactly at the face), so the originals' Qhull regions are the
    # clipped cells
    tiled, _, _, _, _ = periodic_tiling(xs, per, copies=2)
    pts = [tiled]
    for k in range(d):
        if k in per:
            continue
        for face in (0.0, 1.0):
            r = tiled.copy()
            r[:, k] = 2 * face - r[:, k]
            pts.append(r)
    vq = Voronoi(np.vstack(pts))
    vol = 0.0
    for i in range(n):
        region = vq.regions[vq.point_region[i]]
        assert -1 not in region
        q = vq.vertices[region]
        mine = np.array([c for lab, c in ref.items() if (0, i) + (0,) * d in lab])
        vol += ConvexHull(q).volume
        assert len(mine) == len(q)
        dist = np.min(np.linalg.norm(mine[:, None, :] - q[None, :, :], axis=2), axis=1)
        assert np.max(dist) < 1e-9
    assert abs(vol - 1.0) < 1e-10


def test_periodic_oracle_labels_are_translation_consistent():
    # each vertex of the periodic tessellation appears once per original cell it belongs to,
    # translated: shifting every label of a vertex by -s (s the shift of one of its generators)
    # gives another listed vertex, at the coordinates moved by -s
    xs = synth.uniform(2, 80, 7)
    ref = periodic_oracle(xs, (0, 1))
    for lab, c in ref.items():
        for g in lab:
            s = np.array(g[2:])
            if not s.any():
                continue
            moved = tuple(sorted((0, h[1]) + tuple(int(v) for v in np.array(h[2:]) - s) for h in lab))
            assert moved in ref
            assert np.max(np.abs(ref[moved] - (c - s))) < 1e-9
