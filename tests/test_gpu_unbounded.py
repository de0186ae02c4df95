"""GPU parity on domains without a bounded boundary: Boundary() (src/boundary.jl:133-135, no
planes, as test/fraud.jl:7 builds it) and plane sets that leave a direction open (a wedge, a
slab).  A ray whose start candidates all lie behind it searches the half-space ahead of it in
the tree; when that holds no generator the ray is unbounded (the reference's t = Inf,
src/sysvoronoi.jl:504-511) and ends without a vertex.

Bars: the vertex keys equal the oracle's walk bit for bit, coordinates within 1e-10 relative
to max(1, |x|) (far vertices reach |x| ~ 1e4); the unbounded rays are the oracle's, one per
facet of the convex hull (scipy's Qhull); the BF16 and FP32 filters give the FP64 build's keys
and coordinates; in 2D the vertex count is 2n - 2 - h (Euler, h hull points) at the size of
the reference's test/fraud.jl (N = 300 000); the integrators refuse such meshes.  Cell subsets,
refinement and batches work on such domains as on bounded ones (a subset equals the full
build's vertices of its cells, a refined mesh a fresh build, a batch mesh its single build).
"""
import numpy as np
import pytest
from scipy.spatial import ConvexHull

from helpers import rows_view
from hvgpu import synth
from hvgpu._lib import Context, HVError
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

NONE = (np.zeros((0, 2)), np.zeros((0, 2)))


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def build(ctx, xs, base=None, normal=None, prec="fp64"):
    d = xs.shape[1]
    ctx.set_points(xs)
    ctx.set_boundary(np.zeros((0, d)) if base is None else base, np.zeros((0, d)) if normal is None else normal)
    return ctx.build_mesh(precision=prec)


def check_against_oracle(keys, coords, ok, oc):
    assert np.array_equal(rows_view(keys), rows_view(ok))
    row = {k: i for i, k in enumerate(map(tuple, keys.tolist()))}
    j = np.array([row[k] for k in map(tuple, ok.tolist())])
    assert float(np.max(np.abs(coords[j] - oc) / np.maximum(1.0, np.abs(oc)))) < 1e-10


@pytest.mark.parametrize("d,n,seed", [(2, 2000, 1), (3, 1000, 2), (4, 300, 3), (5, 120, 4)])
def test_no_planes_matches_oracle(ctx, d, n, seed):
    xs = synth.uniform(d, n, seed)
    m = build(ctx, xs)
    keys, coords = m.export()
    st = m.stats()
    m.free()
    ok, oc, ost = Oracle(xs).voronoi()
    check_against_oracle(keys, coords, ok, oc)
    hull = len(ConvexHull(xs).simplices)
    assert ost["unbounded"] == hull and st["unbounded"] == hull


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
@pytest.mark.parametrize("d,n,seed", [(3, 5000, 11), (5, 2000, 12)])
def test_filters_equal_fp64_without_planes(ctx, d, n, seed, prec):
    xs = synth.uniform(d, n, seed)
    m0 = build(ctx, xs)
    k0, c0 = m0.export()
    m0.free()
    m1 = build(ctx, xs, prec=prec)
    k1, c1 = m1.export()
    m1.free()
    assert np.array_equal(k0, k1) and np.array_equal(c0, c1)


def test_wedge_and_slab_match_oracle(ctx):
    xs = synth.uniform(2, 500, 7) * 0.5 + 0.25
    for base, normal in ((np.array([[0.0, 0.0], [0.0, 0.0]]), np.array([[-1.0, 0.0], [0.0, -1.0]])),
                         (np.array([[0.0, 1.0], [0.0, 0.0]]), np.array([[0.0, 1.0], [0.0, -1.0]]))):
        m = build(ctx, xs, base, normal, "bf16")
        keys, coords = m.export()
        st = m.stats()
        m.free()
        ok, oc, ost = Oracle(xs, base, normal).voronoi()
        check_against_oracle(keys, coords, ok, oc)
        assert st["unbounded"] == ost["unbounded"] > 0


def test_reference_fraud_size_2d(ctx):
    # test/fraud.jl:7: VoronoiGeometry(VoronoiNodes(rand(2, 300000)), Boundary(), integrate=false)
    n = 300_000
    xs = synth.uniform(2, n, 8)
    m = build(ctx, xs, prec="bf16")
    keys, _ = m.export()
    st = m.stats()
    m.free()
    h = len(ConvexHull(xs).vertices)
    assert len(keys) == 2 * n - 2 - h
    assert st["unbounded"] == h
    # every edge of the triangulation's dual closes twice, except the h unbounded ones
    e = np.concatenate([np.delete(keys, p, axis=1) for p in range(3)])
    _, cnt = np.unique(rows_view(e), return_counts=True)
    assert np.sum(cnt == 1) == h and np.all(cnt <= 2)


def test_integrators_refuse_unbounded_meshes(ctx):
    xs = synth.uniform(3, 300, 21)
    m = build(ctx, xs)
    for method in ("polytope", "montecarlo", "points"):
        with pytest.raises(HVError) as e:
            ctx.integrate(m, method)
        assert e.value.code == -7
    m.free()


def test_voronoi_geometry_without_planes(ctx):
    from hvgpu import Boundary, VI_GEOMETRY, VoronoiGeometry
    xs = synth.uniform(2, 800, 31)
    VG = VoronoiGeometry(xs, Boundary(), integrator=VI_GEOMETRY)
    keys, _ = VG.vertex_table()
    ok, _, _ = Oracle(xs).voronoi()
    assert np.array_equal(rows_view(keys), rows_view(ok))


def test_cells_refine_and_batch_without_planes(ctx):
    # a cell subset is the full build's vertices that hold a listed cell; a refined mesh equals
    # a fresh build of all points; a batch of unbounded meshes equals their single builds
    from hvgpu._lib import build_mesh_batch
    xs = synth.uniform(3, 900, 41)
    n0 = 800
    m = build(ctx, xs[:n0])
    kf, _ = m.export()
    cells = [1, 5, 17, 400]
    sub = ctx.build_mesh_cells(cells)
    ks, _ = sub.export()
    assert np.array_equal(rows_view(ks), rows_view(kf[np.isin(kf, cells).any(axis=1)]))
    sub.free()
    ref, removed = ctx.refine(m, xs[n0:])
    kr, cr = ref.export()
    assert removed > 0
    fresh = build(ctx, xs)
    kx, cx = fresh.export()
    assert np.array_equal(rows_view(kr), rows_view(kx))
    row = {k: i for i, k in enumerate(map(tuple, kx.tolist()))}
    j = np.array([row[k] for k in map(tuple, kr.tolist())])
    assert float(np.max(np.abs(cr - cx[j]) / np.maximum(1.0, np.abs(cx[j])))) < 1e-10
    for mm in (m, ref, fresh):
        mm.free()
    cases = [synth.uniform(d, n, 50 + d) for d, n in ((2, 400), (3, 300), (4, 150))]
    out, status = build_mesh_batch(cases, [(np.zeros((0, x.shape[1])), np.zeros((0, x.shape[1]))) for x in cases],
                                   precision="bf16")
    assert list(status) == [0] * len(cases)
    for mb, x in zip(out, cases):
        single = build(ctx, x, prec="bf16")
        k1, c1 = single.export()
        k, c = mb.export()
        assert np.array_equal(k, k1) and np.array_equal(c, c1)
        single.free()
        mb.free()
