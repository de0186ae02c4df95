"""CPU tests: the oracle pinned against exact-rational fixtures, Qhull and invariants.

The reference ships no golden vectors for this path (its tests only check invariants,
test/rcmethods.jl:3-13), so the oracle is pinned by
  * exact-rational brute force on tiny inputs (tests/golden/*.json),
  * scipy's Qhull Delaunay triangulation (interior vertices = in-box circumcentres),
  * graph closure (every walkable edge has exactly two end points),
  * the published vertex counts of docs/src/index.md:93,96 (statistically).
"""
import numpy as np
import pytest

from helpers import edge_closure_defects, golden_cases, key_set, load_golden
from hvgpu import synth
from oracle_py import Oracle
from exact_vertices import exact_vertices


@pytest.mark.parametrize("name", golden_cases())
def test_oracle_matches_exact_fixture(name):
    g, xs, keys, exact = load_golden(name)
    d = g["d"]
    b, nrm = synth.cube_planes(d)
    ok, oc, st = Oracle(xs, b, nrm).voronoi()
    assert key_set(ok) == key_set(keys)
    ex = {tuple(k): np.array([float(v) for v in c]) for k, c in zip(keys.tolist(), exact)}
    err = max(np.max(np.abs(c - ex[tuple(k)])) for k, c in zip(ok.tolist(), oc))
    assert err < 1e-12
    assert st["irreparable"] == 0 and st["duplicates"] == 0


def test_exact_enumerator_small_closed_form():
    # three generators in the unit square; corners and the circumcentre are closed forms
    from fractions import Fraction
    xs = np.array([[0.25, 0.5], [0.75, 0.5], [0.5, 0.9]])
    keys, coords = exact_vertices(xs)
    pts = dict(zip(keys, coords))
    # corner (0,0): generator 1 with planes x1=0 (plane 2 -> key 3+2) and x2=0 (plane 4 -> key 3+4)
    assert pts[(1, 5, 7)] == (0, 0)
    assert pts[(2, 4, 7)] == (1, 0)
    # |c-g1| = |c-g3| with c = (1/2, y) and a = the double 0.9:
    #   1/16 + (y-1/2)^2 = (y-a)^2  ->  y = (a^2 - 5/16) / (2a - 1)   (= 199/320 for a = 9/10)
    a = Fraction(0.9)
    assert pts[(1, 2, 3)] == (Fraction(1, 2), (a * a - Fraction(5, 16)) / (2 * a - 1))
    # 4 corners, 3 crossings of a bisector with the boundary, 1 interior vertex
    assert len(keys) == 8


def test_fixture_check_detects_a_dropped_vertex():
    # the comparison used above must fail when the oracle misses one vertex
    g, xs, keys, _ = load_golden("d3_n16")
    b, nrm = synth.cube_planes(3)
    ok, _, _ = Oracle(xs, b, nrm).voronoi()
    assert key_set(ok[1:]) != key_set(keys)


def _delaunay_interior(xs):
    from scipy.spatial import Delaunay
    tri = Delaunay(xs)
    P = xs[tri.simplices]
    A = 2 * (P[:, 1:, :] - P[:, :1, :])
    b = (P[:, 1:, :] ** 2).sum(-1) - (P[:, :1, :] ** 2).sum(-1)
    c = np.linalg.solve(A, b[..., None])[..., 0]
    inside = np.all((c > 0) & (c < 1), axis=1)
    return key_set(np.sort(tri.simplices[inside], axis=1) + 1)


@pytest.mark.parametrize("d,n,seed", [(2, 1000, 0), (3, 2000, 5), (4, 1500, 1), (5, 400, 9)])
def test_oracle_interior_matches_qhull(d, n, seed):
    xs = synth.uniform(d, n, seed)
    b, nrm = synth.cube_planes(d)
    ok, _, _ = Oracle(xs, b, nrm).voronoi()
    inner = ok[np.all(ok <= n, axis=1)]
    assert key_set(inner) == _delaunay_interior(xs)


@pytest.mark.parametrize("d,n,seed", [(2, 500, 3), (3, 800, 4), (4, 600, 6), (5, 300, 7), (6, 120, 8)])
def test_oracle_edge_closure(d, n, seed):
    xs = synth.uniform(d, n, seed)
    b, nrm = synth.cube_planes(d)
    ok, _, _ = Oracle(xs, b, nrm).voronoi()
    assert ok.shape[1] == d + 1
    assert edge_closure_defects(ok, n) == 0


def test_oracle_counts_match_published_statistics():
    # docs/src/index.md:93: 4D N=1000 i.i.d. uniform -> 23 467 vertices, 6 507 on the boundary
    xs = synth.uniform(4, 1000, 21)
    b, nrm = synth.cube_planes(4)
    ok, _, st = Oracle(xs, b, nrm).voronoi()
    nb = int(np.sum(np.any(ok > 1000, axis=1)))
    assert abs(len(ok) - 23467) / 23467 < 0.03
    assert abs(nb - 6507) / 6507 < 0.06
    # raycasts ~ vertices (published 552 452 of 553 108 at N=20 000)
    assert st["walkrays"] + st["descents"] >= len(ok) * 0.99


def test_oracle_rejects_bad_input():
    b, nrm = synth.cube_planes(2)
    with pytest.raises(ValueError):
        Oracle(np.array([[0.5, 0.5], [1.5, 0.5], [0.2, 0.2]]), b, nrm)
    with pytest.raises(ValueError):
        Oracle(np.array([[0.5, 0.5], [0.2, 0.5]]), b, nrm)


def test_oracle_single_walkray_matches_its_mesh():
    xs = synth.uniform(3, 200, 17)
    b, nrm = synth.cube_planes(3)
    o = Oracle(xs, b, nrm)
    ok, oc, _ = o.voronoi()
    ks = key_set(ok)
    rng = np.random.default_rng(0)
    done = 0
    for i in rng.permutation(len(ok))[:200]:
        key = ok[i]
        for drop in range(4):
            e = np.delete(key, drop)
            if e[0] > 200:
                continue
            res = o.walkray(key, drop, oc[i], int(e[0]))
            assert res is not None
            assert tuple(res[0].tolist()) in ks
            done += 1
    assert done > 300
