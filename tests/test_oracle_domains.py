"""CPU tests: the oracle on domains other than cuboids, pinned by Qhull and closed forms.

Boundary(planes...) of any unit normals bounding a convex polytope (src/boundary.jl:109-131):
inside a convex domain a mirror is never closer than its own generator, so the oracle's
vertices made of generators only are the Delaunay circumcentres (scipy's Qhull) that lie
inside the domain, and the exact integrator's volumes tile the polytope (closed-form volumes:
the cut cube 1 - 2^-d / d!, the simplex 1 / d!, the regular pentagon).  Boundary() without
planes (src/boundary.jl:133-135): every Delaunay circumcentre is a vertex and every facet of
the convex hull leaves one unbounded ray (the reference's t = Inf, src/sysvoronoi.jl:504-511).
"""
import math

import numpy as np
import pytest
from scipy.spatial import ConvexHull, Delaunay

from helpers import key_set
from hvgpu import synth
from oracle_py import Oracle, polytope


def circumcentres(xs):
    tri = Delaunay(xs)
    P = xs[tri.simplices]
    A = 2 * (P[:, 1:, :] - P[:, :1, :])
    b = (P[:, 1:, :] ** 2).sum(-1) - (P[:, :1, :] ** 2).sum(-1)
    return np.sort(tri.simplices, axis=1) + 1, np.linalg.solve(A, b[..., None])[..., 0]


def cut_cube(d):
    b, n = synth.cube_planes(d)
    return np.vstack([b, np.full(d, (d - 0.5) / d)]), np.vstack([n, np.ones(d) / np.sqrt(d)])


def simplex(d):
    base = np.zeros((d + 1, d))
    normal = -np.eye(d + 1, d)
    normal[d] = np.ones(d) / np.sqrt(d)
    base[d] = np.full(d, 1.0 / d)
    return base, normal


def pentagon(r=0.45):
    ang = np.pi / 2 + 2 * np.pi * np.arange(5) / 5 + np.pi / 5
    normal = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    return 0.5 + r * np.cos(np.pi / 5) * normal, normal


def inside(x, base, normal, margin=0.0):
    return np.all(x @ normal.T < np.sum(base * normal, axis=1) - margin, axis=-1)


def sample(d, n, seed, base, normal):
    xs = synth.uniform(d, 4 * n, seed)
    return np.ascontiguousarray(xs[inside(xs, base, normal, 1e-6)][:n])


CASES = [
    ("cut cube 2D", *cut_cube(2), 2, 600, 1, 1 - 0.25 / 2),
    ("cut cube 3D", *cut_cube(3), 3, 500, 2, 1 - 0.125 / 6),
    ("cut cube 4D", *cut_cube(4), 4, 250, 3, 1 - 0.0625 / 24),
    ("pentagon", *pentagon(), 2, 400, 4, 2.5 * (0.45 ** 2) * math.sin(2 * math.pi / 5)),
]


@pytest.mark.parametrize("name,base,normal,d,n,seed,vol", CASES, ids=[c[0] for c in CASES])
def test_polytope_vertices_match_qhull_and_volumes_tile(name, base, normal, d, n, seed, vol):
    xs = sample(d, n, seed, base, normal)
    ok, oc, _ = Oracle(xs, base, normal).voronoi()
    n = len(xs)
    inner = np.all(ok <= n, axis=1)
    keys, cen = circumcentres(xs)
    # circumcentres within 1e-9 of a plane would make the test depend on rounding
    h = np.sum(base * normal, axis=1)
    assert np.min(np.abs(cen @ normal.T - h)) > 1e-9
    assert key_set(ok[inner]) == key_set(keys[inside(cen, base, normal)])
    v, _, _, _ = polytope(xs, base, normal, ok, oc)
    assert abs(v.sum() - vol) < 1e-12
    assert np.all(v > 0)


def test_simplex_6d_volumes_tile():
    d = 6
    base, normal = simplex(d)
    e = np.random.default_rng(91).exponential(size=(60, d + 1))
    xs = np.ascontiguousarray((e / e.sum(axis=1, keepdims=True))[:, :d])
    ok, oc, _ = Oracle(xs, base, normal).voronoi()
    v, _, _, _ = polytope(xs, base, normal, ok, oc)
    assert abs(v.sum() - 1.0 / math.factorial(d)) < 1e-13


@pytest.mark.parametrize("d,n,seed", [(2, 1500, 11), (3, 800, 12), (4, 300, 13)])
def test_unbounded_vertices_are_all_circumcentres(d, n, seed):
    xs = synth.uniform(d, n, seed)
    ok, oc, st = Oracle(xs).voronoi()
    keys, cen = circumcentres(xs)
    assert key_set(ok) == key_set(keys)
    row = {k: i for i, k in enumerate(map(tuple, keys.tolist()))}
    j = np.array([row[k] for k in map(tuple, ok.tolist())])
    assert float(np.max(np.abs(oc - cen[j]) / np.maximum(1.0, np.abs(cen[j])))) < 1e-9
    assert st["unbounded"] == len(ConvexHull(xs).simplices)


def test_planted_missing_plane_is_caught():
    # the cut cube's vertex set without its cut plane is a different set: the pin above
    # separates a domain from one with a plane dropped
    d = 3
    base, normal = cut_cube(d)
    xs = sample(d, 500, 2, base, normal)
    ok, _, _ = Oracle(xs, base[:-1], normal[:-1]).voronoi()
    keys, cen = circumcentres(xs)
    inner = np.all(ok <= len(xs), axis=1)
    assert key_set(ok[inner]) != key_set(keys[inside(cen, base, normal)])


def test_pyramid_apex_is_out_of_general_position():
    # four slanted planes through the apex of a square pyramid: the vertex of the apex's cell
    # there is equidistant from the generator and its four mirrors (d + 2 members), which the
    # oracle's walk reports instead of continuing (its FastEdgeIterator path is not restated)
    s = 1.0 / np.sqrt(2.0)
    base = np.array([[0.5, 0.5, 0.0]] + [[0.5, 0.5, 1.0]] * 4)
    normal = np.array([[0, 0, -1.0], [s, 0, s], [-s, 0, s], [0, s, s], [0, -s, s]])
    xs = sample(3, 300, 92, base, normal)
    with pytest.raises(RuntimeError, match="non-general"):
        Oracle(xs, base, normal).voronoi()


@pytest.mark.parametrize("d,n,seed", [(2, 40, 21), (3, 25, 22)])
def test_integrand_rule_exact_for_affine_f_on_the_cut_cube(d, n, seed):
    # the reference's quadrature is exact for affine f up to d = 3: the cells' integrals of
    # x_k add up to the cut cube's, 1/2 - (c - c / (2 (d + 1))) with c = 2^-d / d! the volume
    # of the corner cut off (y = 1 - x ranges over a simplex of edge 1/2, centroid 1/(2(d+1)))
    from oracle_py import polytope_integrand
    base, normal = cut_cube(d)
    xs = sample(d, n, seed, base, normal)
    keys, coords, _ = Oracle(xs, base, normal).voronoi()
    vol, _, _, _ = polytope(xs, base, normal, keys, coords)
    f = lambda p: np.concatenate([np.ones((len(p), 1)), p], axis=1)
    bulk, _ = polytope_integrand(xs, base, normal, keys, coords, f, d + 1)
    assert np.allclose(bulk[:, 0], vol, rtol=1e-11, atol=0)
    c = 0.5 ** d / math.factorial(d)
    assert np.allclose(bulk[:, 1:].sum(axis=0), 0.5 - (c - c / (2 * (d + 1))), atol=1e-12)
