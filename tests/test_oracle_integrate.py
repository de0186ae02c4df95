"""CPU tests of the oracle integrators (oracle/hv_integrate.c), pinned independently.

Exact polytope integrator against: the Qhull cell volumes stored with the golden
fixtures (exact vertex sets); exact rational polygon areas (shoelace) in 2D; the unit-box
tiling (sum of volumes = 1); per-cell Qhull volumes on random meshes (d <= 4); per-facet
Qhull (d-1)-volumes of every interface (the facet's vertices projected into its hyperplane,
d = 3 and 4); the two measurements of each interface (one from each cell) agreeing; and a
planted mistake (a wrong cone height) that these checks catch.  Monte-Carlo integrator: statistical agreement
with the exact volumes and areas, and the fixed direction stream (directions differ).
"""
from fractions import Fraction

import numpy as np
import pytest

from helpers import golden_cases, load_golden
from hvgpu import synth
from oracle_py import Oracle, adjacency, mc_direction, mc_dirs, polytope


def integrate(xs, keys, coords):
    b, nrm = synth.cube_planes(xs.shape[1])
    return polytope(xs, b, nrm, keys, coords)


@pytest.mark.parametrize("name", golden_cases())
def test_volumes_match_qhull_on_exact_fixtures(name):
    g, xs, keys, exact = load_golden(name)
    coords = np.array([[float(v) for v in row] for row in exact])
    vol, ptr, nbr, area = integrate(xs, keys, coords)
    want = np.array(g["volumes_qhull"])
    assert np.max(np.abs(vol - want) / want) < 1e-12
    assert abs(vol.sum() - 1.0) < 1e-13


def test_2d_areas_and_volumes_are_exact_shoelace():
    g, xs, keys, exact = load_golden("d2_n30")
    coords = np.array([[float(v) for v in row] for row in exact])
    vol, ptr, nbr, area = integrate(xs, keys, coords)
    n = len(xs)
    for i in range(n):
        rows = [r for r in range(len(keys)) if i + 1 in keys[r]]
        pts = [exact[r] for r in rows]
        cx = sum(p[0] for p in pts) / len(pts)
        cy = sum(p[1] for p in pts) / len(pts)
        pts.sort(key=lambda p: np.arctan2(float(p[1] - cy), float(p[0] - cx)))
        A = Fraction(0)
        for a, b in zip(pts, pts[1:] + pts[:1]):
            A += a[0] * b[1] - a[1] * b[0]
        assert abs(vol[i] - float(abs(A) / 2)) < 1e-14
        for t in range(ptr[i], ptr[i + 1]):
            ends = [exact[r] for r in rows if nbr[t] in keys[r]]
            assert len(ends) == 2
            L2 = (ends[0][0] - ends[1][0]) ** 2 + (ends[0][1] - ends[1][1]) ** 2
            assert abs(area[t] - float(L2) ** 0.5) < 1e-14


@pytest.mark.parametrize("d,n,seed", [(2, 300, 1), (3, 250, 2), (4, 150, 3)])
def test_random_meshes_match_qhull_and_tile_the_box(d, n, seed):
    from scipy.spatial import ConvexHull
    xs = synth.uniform(d, n, seed)
    b, nrm = synth.cube_planes(d)
    keys, coords, _ = Oracle(xs, b, nrm).voronoi()
    vol, ptr, nbr, area = polytope(xs, b, nrm, keys, coords)
    cp, cv, _, _ = adjacency(keys, n)
    qh = np.array([ConvexHull(coords[cv[cp[i]:cp[i + 1]]]).volume for i in range(n)])
    assert np.max(np.abs(vol - qh) / qh) < 1e-11
    assert abs(vol.sum() - 1.0) < 1e-12
    # interface areas from the two sides agree
    for i in range(n):
        for t in range(ptr[i], ptr[i + 1]):
            k = nbr[t]
            if k <= n:
                tk = ptr[k - 1] + np.searchsorted(nbr[ptr[k - 1]:ptr[k]], i + 1)
                assert abs(area[t] - area[tk]) <= 1e-10 * max(area[t], 1e-6 * area.max())


@pytest.mark.parametrize("d,n,seed", [(3, 250, 4), (4, 120, 5)])
def test_facet_areas_match_qhull(d, n, seed):
    """Every interface area of the oracle's integrator equals the (d-1)-volume of the convex
    hull of the facet's vertices (those whose keys hold both ids) in an orthonormal basis of
    the facet's hyperplane: the bisector of two generators, or the boundary plane."""
    from scipy.spatial import ConvexHull
    xs = synth.uniform(d, n, seed)
    b, nrm = synth.cube_planes(d)
    keys, coords, _ = Oracle(xs, b, nrm).voronoi()
    vol, ptr, nbr, area = polytope(xs, b, nrm, keys, coords)
    rows_of = {}
    for r, key in enumerate(keys.tolist()):
        for g in key:
            rows_of.setdefault(g, set()).add(r)
    checked = 0
    for i in range(n):
        for t in range(ptr[i], ptr[i + 1]):
            k = int(nbr[t])
            rows = sorted(rows_of[i + 1] & rows_of[k])
            normal = xs[k - 1] - xs[i] if k <= n else nrm[k - n - 1]
            normal = normal / np.linalg.norm(normal)
            # orthonormal basis of the hyperplane: the last d - 1 columns of a QR of [normal | I]
            q, _ = np.linalg.qr(np.column_stack([normal, np.eye(d)]))
            basis = q[:, 1:d]
            pts = (coords[rows] - coords[rows[0]]) @ basis
            if len(rows) < d:
                assert area[t] < 1e-12
                continue
            try:
                a = ConvexHull(pts).volume
            except Exception:   # flat facet (all vertices in a (d-2)-plane): no (d-1)-volume
                a = 0.0
            assert abs(area[t] - a) <= 1e-9 * max(a, 1e-6 * area.max())
            checked += 1
    assert checked > n * d


def test_planted_height_mistake_is_caught():
    # the checks above must fail for a plausible slip: cone heights |x_i - x_k| instead of
    # |x_i - x_k| / 2 on generator interfaces (volumes rebuilt from the oracle's areas)
    g, xs, keys, exact = load_golden("d3_n16")
    coords = np.array([[float(v) for v in row] for row in exact])
    vol, ptr, nbr, area = integrate(xs, keys, coords)
    n, d = xs.shape
    b, nrm = synth.cube_planes(d)
    right = np.zeros(n)
    wrong = np.zeros(n)
    for i in range(n):
        for t in range(ptr[i], ptr[i + 1]):
            k = nbr[t]
            if k > n:
                h = float(nrm[k - n - 1] @ b[k - n - 1] - nrm[k - n - 1] @ xs[i])
                hw = h
            else:
                h = 0.5 * float(np.linalg.norm(xs[i] - xs[k - 1]))
                hw = 2.0 * h
            right[i] += area[t] * h / d
            wrong[i] += area[t] * hw / d
    qh = np.array(g["volumes_qhull"])
    assert np.max(np.abs(right - qh) / qh) < 1e-12
    assert abs(wrong.sum() - 1.0) > 0.1
    assert np.max(np.abs(wrong - qh) / qh) > 0.1


def test_montecarlo_is_unbiased_against_the_exact_integrator():
    d, n = 3, 120
    xs = synth.uniform(d, n, 5)
    b, nrm = synth.cube_planes(d)
    keys, coords, _ = Oracle(xs, b, nrm).voronoi()
    vol, ptr, nbr, area = polytope(xs, b, nrm, keys, coords)
    vm, am = mc_dirs(xs, b, nrm, ptr, nbr, 4000, 3)
    err = (vm - vol) / vol
    assert abs(np.median(err)) < 0.01
    assert np.median(np.abs(err)) < 0.05
    assert abs(vm.sum() - 1.0) < 0.01
    big = area > 0.2 * area.max()
    assert np.median(np.abs(am[big] - area[big]) / area[big]) < 0.1


def test_montecarlo_directions_are_independent_unit_vectors():
    us = np.array([mc_direction(4, 9, 3, 50, q) for q in range(50)])
    assert np.allclose(np.linalg.norm(us, axis=1), 1.0, atol=1e-15)
    # the reference's stream repeats its first direction; this one does not
    assert len({tuple(np.round(u, 12)) for u in us}) == 50
    assert abs(us.mean(axis=0)).max() < 0.5
