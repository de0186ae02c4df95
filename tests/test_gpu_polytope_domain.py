"""GPU parity on domains other than cuboids: Boundary(planes...) (src/boundary.jl:109-131) with
Dirichlet / Neumann planes of any unit normal that bound a convex polytope.  The mirrors are
the reference's reflect (src/boundary.jl:206-211) for every plane, as for the cuboid.

Domains: the unit cube with the corner at (1, ..., 1) cut off by sum(x) <= d - 1/2 (volume
1 - 2^-d / d!), a regular pentagon and a triangle in 2D, the unit cube turned by a random
rotation (3D, 4D: no plane normal along an axis) and the 6D simplex x >= 0, sum(x) <= 1
(volume 1/720).  Every vertex of these polytopes has exactly d planes, so the tessellations
are in general position.  The apex of a square pyramid has four planes: the vertex of the cell
there has d + 2 elements, which the oracle reports as a non-general vertex and the device as
HV_ERR_DEGENERATE.

Bars: the vertex keys equal the oracle's walk (src/sysvoronoi.jl) bit for bit, coordinates
within 1e-10; the BF16 and FP32 filters give the FP64 build's keys and coordinates; the exact
integrals equal the oracle's polytope restatement (volumes 1e-10 relative, the same
neighbour rows) and tile the domain (sum of volumes = the polytope's volume to 1e-12); a mesh
of some cells and a refined mesh equal their full and fresh builds; integrands (f = 1, x)
equal the oracle's rule to 1e-10 and the cut cube's closed-form first moments to 1e-12.  Normals of other than
unit length are refused with HV_ERR_ARG, point samples (drawn in a box) with
HV_ERR_UNSUPPORTED.  (Planes that bound no domain: tests/test_gpu_unbounded.py.)
"""
import math

import numpy as np
import pytest

from helpers import rows_view
from hvgpu import synth
from hvgpu._lib import Context, HVError
from oracle_py import Oracle, polytope

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def cut_cube(d, cut=None):
    """Planes of the unit cube (cuboid order) and the cut sum(x) <= cut."""
    cut = d - 0.5 if cut is None else cut
    b, n = synth.cube_planes(d)
    nn = np.ones(d) / np.sqrt(d)
    return np.vstack([b, np.full(d, cut / d)]), np.vstack([n, nn])


def pentagon(r=0.45, c=(0.5, 0.5)):
    """A regular pentagon: planes through its edge midpoints with outer unit normals."""
    ang = np.pi / 2 + 2 * np.pi * np.arange(5) / 5 + np.pi / 5
    normal = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    base = np.asarray(c) + r * np.cos(np.pi / 5) * normal
    return base, normal


def inside(xs, base, normal, margin=1e-6):
    h = np.sum(base * normal, axis=1)
    return np.all(xs @ normal.T < h - margin, axis=1)


def points(d, n, seed, base, normal):
    xs = synth.uniform(d, 4 * n, seed)
    return np.ascontiguousarray(xs[inside(xs, base, normal)][:n])


def build(ctx, xs, base, normal, prec="fp64"):
    ctx.set_points(xs)
    ctx.set_boundary(base, normal)
    return ctx.build_mesh(precision=prec)


def check_against_oracle(keys, coords, ok, oc):
    assert np.array_equal(rows_view(keys), rows_view(ok))
    row = {k: i for i, k in enumerate(map(tuple, keys.tolist()))}
    j = np.array([row[k] for k in map(tuple, ok.tolist())])
    assert float(np.max(np.abs(coords[j] - oc))) < 1e-10


@pytest.mark.parametrize("d,n,seed", [(2, 600, 1), (3, 500, 2), (4, 300, 3), (5, 150, 4)])
def test_cut_cube_matches_oracle(ctx, d, n, seed):
    base, normal = cut_cube(d)
    xs = points(d, n, seed, base, normal)
    m = build(ctx, xs, base, normal)
    keys, coords = m.export()
    ok, oc, _ = Oracle(xs, base, normal).voronoi()
    check_against_oracle(keys, coords, ok, oc)
    # the cut plane is id n + 2d + 1: some vertices lie on it
    assert np.any(keys == len(xs) + 2 * d + 1)
    I = ctx.integrate(m, "polytope")
    m.free()
    vol, ptr, nbr, area = polytope(xs, base, normal, keys, coords)
    assert np.array_equal(I["nbr_ptr"], ptr) and np.array_equal(I["neighbors"], nbr)
    assert float(np.max(np.abs(I["volumes"] - vol) / vol)) < 1e-10
    assert abs(I["volumes"].sum() - (1.0 - 0.5 ** d / math.factorial(d))) < 1e-12


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
@pytest.mark.parametrize("d,n,seed", [(3, 3000, 11), (5, 2000, 12)])
def test_cut_cube_filters_equal_fp64(ctx, d, n, seed, prec):
    base, normal = cut_cube(d)
    xs = points(d, n, seed, base, normal)
    m0 = build(ctx, xs, base, normal, "fp64")
    k0, c0 = m0.export()
    m0.free()
    m1 = build(ctx, xs, base, normal, prec)
    k1, c1 = m1.export()
    m1.free()
    assert np.array_equal(k0, k1) and np.array_equal(c0, c1)


def test_pentagon_matches_oracle(ctx):
    base, normal = pentagon()
    xs = points(2, 800, 21, base, normal)
    m = build(ctx, xs, base, normal, "bf16")
    keys, coords = m.export()
    ok, oc, _ = Oracle(xs, base, normal).voronoi()
    check_against_oracle(keys, coords, ok, oc)
    I = ctx.integrate(m, "polytope")
    # area of the regular pentagon with circumradius r: (5/2) r^2 sin(72 deg)
    assert abs(I["volumes"].sum() - 2.5 * 0.45 ** 2 * np.sin(2 * np.pi / 5)) < 1e-12
    D = ctx.integrate(m, "montecarlo", ndir=400, seed=2)
    assert np.array_equal(D["nbr_ptr"], I["nbr_ptr"]) and np.array_equal(D["neighbors"], I["neighbors"])
    assert float(np.sum(np.abs(D["volumes"] - I["volumes"])) / np.sum(I["volumes"])) < 0.05
    with pytest.raises(HVError) as e:
        ctx.integrate(m, "points", nsamples=1_000_000)
    assert e.value.code == -7
    m.free()


def test_cells_and_refine_on_a_cut_cube(ctx):
    d = 3
    base, normal = cut_cube(d)
    xs = points(d, 900, 31, base, normal)
    n0 = 800
    m = build(ctx, xs[:n0], base, normal)
    kf, cf = m.export()
    sub = ctx.build_mesh_cells([1, 5, 17, 400])
    ks, cs = sub.export()
    sel = np.isin(kf, [1, 5, 17, 400]).any(axis=1)
    assert np.array_equal(rows_view(ks), rows_view(kf[sel]))
    sub.free()
    ref, removed = ctx.refine(m, xs[n0:])
    kr, cr = ref.export()
    assert removed > 0
    fresh = build(ctx, xs, base, normal)
    kx, cx = fresh.export()
    assert np.array_equal(rows_view(kr), rows_view(kx))
    for mm in (m, ref, fresh):
        mm.free()


def test_normals_of_other_than_unit_length_are_refused(ctx):
    ctx.set_points(synth.uniform(2, 50, 41) * 0.5 + 0.25)
    with pytest.raises(HVError) as e:
        ctx.set_boundary(np.array([[1.0, 0.0]]), np.array([[2.0, 0.0]]))
    assert e.value.code == -1


def test_triangle_matches_oracle(ctx):
    # x >= 0, y >= 0, x + y <= 1
    tb = np.array([[0.0, 0.0], [0.0, 0.0], [0.5, 0.5]])
    tn = np.array([[-1.0, 0.0], [0.0, -1.0], [1.0 / np.sqrt(2), 1.0 / np.sqrt(2)]])
    xt = points(2, 300, 42, tb, tn)
    m = build(ctx, xt, tb, tn)
    keys, coords = m.export()
    ok, oc, _ = Oracle(xt, tb, tn).voronoi()
    check_against_oracle(keys, coords, ok, oc)
    assert abs(ctx.integrate(m, "polytope")["volumes"].sum() - 0.5) < 1e-12
    m.free()


def test_periodic_flags_need_a_cuboid(ctx):
    d = 2
    base, normal = cut_cube(d)
    bc = np.array([2, 1, 0, 0, 0], dtype=np.int16)   # the x pair periodic, plus a cut
    xs = points(d, 100, 51, base, normal)
    ctx.set_points(xs)
    ctx.set_boundary(base, normal, bc)
    with pytest.raises(HVError) as e:
        ctx.build_mesh()
    assert e.value.code == -7


def test_batch_and_voronoi_geometry_on_polytopes(ctx):
    # non-cuboid planes go to the batch's host workers; each mesh equals its single build
    from hvgpu import BC_Dirichlet, BC_Neumann, VI_POLYGON, VoronoiGeometry, boundary
    from hvgpu._lib import build_mesh_batch
    cases = []
    for d, n, seed in ((2, 300, 61), (3, 400, 62), (4, 200, 63)):
        base, normal = cut_cube(d)
        cases.append((points(d, n, seed, base, normal), base, normal))
    pb, pn = pentagon()
    cases.append((points(2, 250, 64, pb, pn), pb, pn))
    out, status = build_mesh_batch([x for x, _, _ in cases], [(b, nr) for _, b, nr in cases], precision="bf16")
    assert list(status) == [0] * len(cases)
    for m, (x, b, nr) in zip(out, cases):
        single = build(ctx, x, b, nr, "bf16")
        k1, c1 = single.export()
        k, c = m.export()
        assert np.array_equal(k, k1) and np.array_equal(c, c1)
        single.free()
        m.free()
    # VoronoiGeometry with Boundary(planes...): Dirichlet and Neumann planes act alike on the mesh
    d = 3
    base, normal = cut_cube(d)
    xs = points(d, 500, 65, base, normal)
    planes = [BC_Dirichlet(b, nr) if i % 2 else BC_Neumann(b, nr) for i, (b, nr) in enumerate(zip(base, normal))]
    VG = VoronoiGeometry(xs, boundary(*planes), integrator=VI_POLYGON)
    assert abs(VG.integral["volumes"].sum() - (1.0 - 0.5 ** d / 6.0)) < 1e-12
    keys, _ = VG.vertex_table()
    ok, _, _ = Oracle(xs, base, normal).voronoi()
    assert np.array_equal(rows_view(keys), rows_view(ok))


def rotation(d, seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).normal(size=(d, d)))
    return q * np.sign(np.diag(r))


@pytest.mark.parametrize("d,n,seed", [(3, 400, 81), (4, 200, 82)])
def test_rotated_cube_matches_oracle(ctx, d, n, seed):
    # the unit cube about (1/2, ..., 1/2) turned by a random rotation: every plane has a
    # non-axis normal; keys equal to the oracle, volumes summing to 1
    Q = rotation(d, seed)
    c = np.full(d, 0.5)
    b0, n0 = synth.cube_planes(d)
    normal = n0 @ Q.T
    base = (b0 - c) @ Q.T + c
    xs = points(d, n, seed, base, normal)
    m = build(ctx, xs, base, normal, "bf16")
    keys, coords = m.export()
    ok, oc, _ = Oracle(xs, base, normal).voronoi()
    check_against_oracle(keys, coords, ok, oc)
    assert abs(ctx.integrate(m, "polytope")["volumes"].sum() - 1.0) < 1e-12
    m.free()


def test_simplex_6d_matches_oracle(ctx):
    # x_k >= 0 and sum x <= 1: volume 1/6!
    d = 6
    base = np.zeros((d + 1, d))
    normal = -np.eye(d + 1, d)
    normal[d] = np.ones(d) / np.sqrt(d)
    base[d] = np.full(d, 1.0 / d)
    e = np.random.default_rng(91).exponential(size=(60, d + 1))
    xs = np.ascontiguousarray((e / e.sum(axis=1, keepdims=True))[:, :d])
    assert inside(xs, base, normal).all()
    m = build(ctx, xs, base, normal)
    keys, coords = m.export()
    ok, oc, _ = Oracle(xs, base, normal).voronoi()
    check_against_oracle(keys, coords, ok, oc)
    assert abs(ctx.integrate(m, "polytope")["volumes"].sum() - 1.0 / 720.0) < 1e-13
    m.free()


def test_pyramid_apex_is_out_of_general_position(ctx):
    # four slanted planes meet at the apex of a square pyramid: the vertex there holds the
    # apex's cell and four planes, d + 2 elements (the reference's non-general path); the
    # device reports it as HV_ERR_DEGENERATE
    s = 1.0 / np.sqrt(2.0)
    base = np.array([[0.5, 0.5, 0.0], [0.5, 0.5, 1.0], [0.5, 0.5, 1.0], [0.5, 0.5, 1.0], [0.5, 0.5, 1.0]])
    normal = np.array([[0, 0, -1.0], [s, 0, s], [-s, 0, s], [0, s, s], [0, -s, s]])
    base[1:] = np.array([0.5, 0.5, 1.0])
    xs = points(3, 300, 92, base, normal)
    ctx.set_points(xs)
    ctx.set_boundary(base, normal)
    with pytest.raises(HVError) as e:
        ctx.build_mesh()
    assert e.value.code == -4


@pytest.mark.parametrize("d,n,seed", [(2, 40, 21), (3, 25, 22)])
def test_integrand_on_a_cut_cube_matches_oracle_and_closed_form(ctx, d, n, seed):
    # f = (1, x): the device's bulk integrals equal the oracle's restatement of the reference
    # rule (1e-10 relative), and the first moments add up to the cut cube's closed form
    from oracle_py import polytope_integrand
    base, normal = cut_cube(d)
    xs = points(d, n, seed, base, normal)
    m = build(ctx, xs, base, normal)
    f = lambda p: np.concatenate([np.ones((len(p), 1)), p], axis=1)
    r = ctx.integrate(m, "polytope", integrand=f, nf=d + 1)
    m.free()
    keys, coords, _ = Oracle(xs, base, normal).voronoi()
    ob, _ = polytope_integrand(xs, base, normal, keys, coords, f, d + 1)
    assert np.allclose(r["bulk"], ob, rtol=1e-10, atol=1e-14)
    c = 0.5 ** d / math.factorial(d)
    assert np.allclose(r["bulk"][:, 1:].sum(axis=0), 0.5 - (c - c / (2 * (d + 1))), atol=1e-12)
