"""GPU parity of the integrators (libhv.so through the C ABI) against the oracle.

Exact polytope integrator: neighbour lists identical to the oracle's adjacency (integer
work, bit-exact); volumes within 1e-10 relative and interface areas within 1e-10 relative
to max(|A|, 1e-6 max A) of the oracle's FP64 restatement of src/polyintegrator.jl; plus the properties the geometry fixes: the volumes
tile the unit box (sum = 1 to 1e-12), both sides of an interface measure the same area,
and 2D areas equal the exact edge lengths.  Monte-Carlo integrator: the same directions as
the oracle (counter-based stream), so volumes and areas agree with the oracle's FP64 run to
1e-9, and statistically with the exact values.
"""
import numpy as np
import pytest

from hvgpu import synth
from oracle_py import Oracle, adjacency, mc_dirs, polytope

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def ctx():
    from hvgpu._lib import Context
    return Context(0)


def mesh_of(ctx, xs, precision="fp64"):
    b, nrm = synth.cube_planes(xs.shape[1])
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    return ctx.build_mesh(precision=precision)


def rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def rel_area(a, b):
    """Relative error of interface areas with a floor of 1e-6 of the largest interface:
    slivers (areas ~1e-20 next to ~1e-4) are ill-conditioned in any FP64 evaluation (the
    oracle's two measurements of one such facet, from either cell, differ by ~1e-8)."""
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-6 * np.max(np.abs(b)))))


@pytest.mark.parametrize("d,n,seed", [(2, 400, 3), (3, 600, 4), (4, 500, 5), (5, 200, 6), (6, 80, 7)])
def test_polytope_matches_oracle(ctx, d, n, seed):
    xs = synth.uniform(d, n, seed)
    m = mesh_of(ctx, xs)
    keys, coords = m.export()
    I = ctx.integrate(m, "polytope")
    m.free()
    b, nrm = synth.cube_planes(d)
    vol, ptr, nbr, area = polytope(xs, b, nrm, keys, coords)
    assert np.array_equal(I["nbr_ptr"], ptr)
    assert np.array_equal(I["neighbors"], nbr)
    assert rel(I["volumes"], vol) < TOL
    assert rel_area(I["areas"], area) < TOL
    assert abs(I["volumes"].sum() - 1.0) < 1e-12
    # each interface measured from both sides
    for i in range(n):
        for t in range(ptr[i], ptr[i + 1]):
            k = nbr[t]
            if k <= n:
                tk = ptr[k - 1] + np.searchsorted(nbr[ptr[k - 1]:ptr[k]], i + 1)
                assert abs(I["areas"][t] - I["areas"][tk]) <= 1e-11 * I["areas"][t]


def test_polytope_2d_areas_are_edge_lengths(ctx):
    xs = synth.uniform(2, 300, 11)
    m = mesh_of(ctx, xs)
    keys, coords = m.export()
    I = ctx.integrate(m, "polytope")
    m.free()
    n = 300
    # in 2D the interface (i, k) is the Voronoi edge whose two end points carry both ids
    ptr, nbr = I["nbr_ptr"], I["neighbors"]
    for i in range(0, n, 7):
        for t in range(ptr[i], ptr[i + 1]):
            k = nbr[t]
            rows = np.where(np.any(keys == i + 1, axis=1) & np.any(keys == k, axis=1))[0]
            assert len(rows) == 2
            assert abs(np.linalg.norm(coords[rows[0]] - coords[rows[1]]) - I["areas"][t]) < 1e-13


def test_polytope_4d_20k(ctx):
    # 4D N=20 000 uniform, cuboid(4, periodic=[]): exact volumes and areas vs the oracle
    xs = synth.uniform(4, 20000, 1)
    m = mesh_of(ctx, xs)
    keys, coords = m.export()
    I = ctx.integrate(m, "polytope")
    m.free()
    b, nrm = synth.cube_planes(4)
    vol, ptr, nbr, area = polytope(xs, b, nrm, keys, coords)
    assert np.array_equal(I["neighbors"], nbr)
    assert rel(I["volumes"], vol) < TOL
    assert rel_area(I["areas"], area) < TOL
    assert abs(I["volumes"].sum() - 1.0) < 1e-12


def test_polytope_on_bf16_mesh_equals_fp64_mesh(ctx):
    xs = synth.uniform(5, 400, 21)
    m = mesh_of(ctx, xs, "fp64")
    a = ctx.integrate(m, "polytope")
    m.free()
    m = mesh_of(ctx, xs, "bf16")
    c = ctx.integrate(m, "polytope")
    m.free()
    assert np.array_equal(a["neighbors"], c["neighbors"])
    assert np.array_equal(a["volumes"], c["volumes"])
    assert np.array_equal(a["areas"], c["areas"])


@pytest.mark.parametrize("d,n,seed", [(2, 200, 3), (3, 300, 4), (5, 120, 6)])
def test_montecarlo_matches_oracle(ctx, d, n, seed):
    xs = synth.uniform(d, n, seed)
    m = mesh_of(ctx, xs)
    keys, coords = m.export()
    I = ctx.integrate(m, "montecarlo", ndir=1500, seed=99)
    E = ctx.integrate(m, "polytope")
    m.free()
    b, nrm = synth.cube_planes(d)
    _, _, ptr, nbr = adjacency(keys, n)
    vol, area = mc_dirs(xs, b, nrm, ptr, nbr, 1500, 99)
    assert np.array_equal(I["neighbors"], nbr)
    assert rel(I["volumes"], vol) < 1e-9
    assert float(np.max(np.abs(I["areas"] - area))) < 1e-9 * float(np.max(area))
    # statistical agreement with the exact integrator: unbiased per cell, sum near 1
    err = (I["volumes"] - E["volumes"]) / E["volumes"]
    assert abs(np.median(err)) < 0.01
    assert abs(I["volumes"].sum() - 1.0) < 0.02


def test_voronoi_geometry_and_data_mirror(ctx):
    from hvgpu import VI_POLYGON, VoronoiData, VoronoiGeometry, cuboid
    xs = synth.uniform(3, 300, 41)
    VG = VoronoiGeometry(xs, cuboid(3, periodic=[]), integrator=VI_POLYGON, search_settings={"precision": "bf16"})
    data = VoronoiData(VG)
    keys, coords = VG.vertex_table()
    b, nrm = synth.cube_planes(3)
    ok, oc, _ = Oracle(xs, b, nrm).voronoi()
    assert np.array_equal(keys, ok[np.lexsort(ok.T[::-1])])
    vol, ptr, nbr, area = polytope(xs, b, nrm, keys, coords)
    assert rel(data.volume, vol) < TOL
    for i in (0, 17, 299):
        assert np.array_equal(data.neighbors[i], nbr[ptr[i]:ptr[i + 1]])
        assert rel_area(data.area[i], area[ptr[i]:ptr[i + 1]]) < TOL
        for key, r in data.vertices[i].items():
            assert i + 1 in key and len(key) == 4
    # the reference's default cuboid is periodic in every direction: public ids name
    # originals only (images reduced to their original, no boundary planes), the cells tile
    # the torus, and cell i's vertices all hold i + 1
    VG = VoronoiGeometry(xs, cuboid(3), integrator=VI_POLYGON, search_settings={"precision": "bf16"})
    data = VoronoiData(VG)
    assert abs(data.volume.sum() - 1.0) < 1e-12
    for i in (0, 17, 299):
        assert np.all((data.neighbors[i] >= 1) & (data.neighbors[i] <= 300)) and i + 1 not in data.neighbors[i]
        assert len(data.vertices[i]) >= 4
        for key in data.vertices[i]:
            assert i + 1 in key and max(key) <= 300
