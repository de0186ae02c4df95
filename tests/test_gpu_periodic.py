"""GPU parity of periodic boxes (cuboid(d; periodic=...), src/boundary.jl:510-534; the
reference's periodic node copies, src/domain.jl:139-390).

Bar: the device mesh of a periodic box holds exactly the vertices of the originals' cells of
the periodic tessellation.  Generators are compared through (original, shift) labels, keys
bit-exact, coordinates within 1e-10 of the oracle's walk over the originals' cells of a tiled
point set (helpers.periodic_oracle, itself checked against Qhull in test_oracle_periodic.py).
Integrals: the originals' cells tile the torus (volumes sum to the box volume), every interface
is seen from both sides with the same area, and the tessellation is invariant under
translations modulo the box.
"""
import numpy as np
import pytest

from helpers import periodic_mesh_labels, periodic_oracle
from hvgpu import synth
from hvgpu.geometry import cuboid
from hvgpu._lib import Context

pytestmark = pytest.mark.gpu

COORD_TOL = 1e-10


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def build(ctx, xs, periodic, precision="fp64"):
    base, normal, bc = cuboid(xs.shape[1], periodic=[k + 1 for k in periodic]).arrays()
    ctx.set_points(xs)
    ctx.set_boundary(base, normal, bc)
    return ctx.build_mesh(precision=precision)


def compare(ref, mine):
    assert set(ref) == set(mine)
    err = max(float(np.max(np.abs(ref[k] - mine[k]))) for k in ref)
    assert err < COORD_TOL


@pytest.mark.parametrize("prec", ["fp64", "bf16", "fp32"])
@pytest.mark.parametrize("d,n,seed,per", [(2, 300, 1, (0, 1)), (2, 200, 2, (1,)), (3, 200, 3, (0, 1, 2)),
                                          (3, 150, 4, (0, 2)), (4, 60, 5, (0, 1, 2, 3))])
def test_periodic_mesh_matches_tiled_oracle(ctx, d, n, seed, per, prec):
    xs = synth.uniform(d, n, seed)
    m = build(ctx, xs, per, prec)
    mine = periodic_mesh_labels(m)
    m.free()
    compare(periodic_oracle(xs, per), mine)


@pytest.mark.parametrize("d,n,seed", [(2, 6, 81), (3, 9, 82), (2, 3, 83)])
def test_periodic_few_points(ctx, d, n, seed):
    # cells about as large as the box: images from up to two lengths away
    xs = synth.uniform(d, n, seed)
    m = build(ctx, xs, tuple(range(d)))
    mine = periodic_mesh_labels(m)
    assert np.abs(m.images()[1]).max() >= 1
    I = ctx.integrate(m, "polytope")
    m.free()
    compare(periodic_oracle(xs, tuple(range(d)), copies=2), mine)
    assert abs(I["volumes"].sum() - 1.0) < 1e-12


def test_periodic_box_with_offset_and_lengths(ctx):
    # cuboid(3; dimensions=[2, 1, 0.5], offset=[-1, 0.5, 0], periodic=[1, 3]): volume 1
    dims, off = np.array([2.0, 1.0, 0.5]), np.array([-1.0, 0.5, 0.0])
    xs = off + synth.uniform(3, 250, 91) * dims
    base, normal, bc = cuboid(3, dimensions=dims, offset=off, periodic=[1, 3]).arrays()
    ctx.set_points(xs)
    ctx.set_boundary(base, normal, bc)
    m = ctx.build_mesh(precision="bf16")
    mine = periodic_mesh_labels(m)
    I = ctx.integrate(m, "polytope")
    m.free()
    compare(periodic_oracle(xs, (0, 2), lo=off, hi=off + dims), mine)
    assert abs(I["volumes"].sum() - 1.0) < 1e-12


def test_periodic_images_table(ctx):
    xs = synth.uniform(3, 400, 11)
    m = build(ctx, xs, (0, 1, 2))
    orig, shift = m.images()
    n = len(xs)
    assert np.array_equal(orig[:n], np.arange(1, n + 1)) and not shift[:n].any()
    assert len(orig) > n and np.all(np.abs(shift[n:]).sum(axis=1) > 0)
    # every image is the original moved by whole box lengths, inside the margin
    k, _ = m.export()
    assert k.min() >= 1 and k[:, 0].max() <= n   # each vertex has an original generator
    m.free()


def volumes(ctx, m, method="polytope", **kw):
    I = ctx.integrate(m, method, **kw)
    return I


@pytest.mark.parametrize("d,n,seed,per", [(2, 500, 21, (0, 1)), (3, 400, 22, (0, 1, 2)), (4, 300, 23, (0, 1, 2, 3)),
                                          (5, 200, 24, (0, 1, 2, 3, 4)), (3, 300, 25, (1,)), (6, 150, 26, (0, 1, 2, 3, 4, 5))])
def test_periodic_cells_tile_the_torus(ctx, d, n, seed, per):
    xs = synth.uniform(d, n, seed)
    m = build(ctx, xs, per)
    orig, shift = m.images()
    I = ctx.integrate(m, "polytope")
    vol, ptr, nbr, area = I["volumes"], I["nbr_ptr"], I["neighbors"], I["areas"]
    assert len(vol) == n and np.all(vol > 0)
    assert abs(vol.sum() - 1.0) < 1e-12
    # interface (i, image j) seen from j: (j, image of i shifted the other way), same area
    next_ = len(orig)
    lab = {}
    for i in range(n):
        for t in range(ptr[i], ptr[i + 1]):
            j = nbr[t]
            if j > next_:
                continue   # a Dirichlet plane
            lab[(i, int(orig[j - 1]) - 1, tuple(shift[j - 1]))] = area[t]
    for (i, j, s), a in lab.items():
        back = (j, i, tuple(-v for v in s))
        assert back in lab
        assert abs(lab[back] - a) <= 1e-9 * max(a, 1e-12)
    m.free()


@pytest.mark.parametrize("d,n,seed", [(2, 400, 31), (3, 300, 32), (5, 150, 33)])
def test_periodic_translation_invariance(ctx, d, n, seed):
    xs = synth.uniform(d, n, seed)
    t = synth.uniform(d, 1, seed + 100)[0]
    m0 = build(ctx, xs, tuple(range(d)))
    v0 = ctx.integrate(m0, "polytope")["volumes"]
    m0.free()
    m1 = build(ctx, np.mod(xs + t, 1.0), tuple(range(d)))
    v1 = ctx.integrate(m1, "polytope")["volumes"]
    m1.free()
    assert np.max(np.abs(v0 - v1)) < 1e-10


def test_reference_periodic_5d_volume_sum(ctx):
    # test/jld.jl:16-20: 5D N=100, periodic=[1], sum of volumes 1 +- 0.01 (here to 1e-12)
    xs = synth.uniform(5, 100, 41)
    m = build(ctx, xs, (0,))
    I = ctx.integrate(m, "polytope")
    assert abs(I["volumes"].sum() - 1.0) < 1e-12
    m.free()


@pytest.mark.parametrize("d,n,seed", [(3, 2000, 51), (4, 1000, 52)])
def test_periodic_point_sample_mc(ctx, d, n, seed):
    xs = synth.uniform(d, n, seed)
    m = build(ctx, xs, tuple(range(d)))
    ex = ctx.integrate(m, "polytope")
    mc = ctx.integrate(m, "points", nsamples=20_000_000, seed=3)
    # every sample is counted in exactly one original cell
    assert abs(mc["volumes"].sum() - 1.0) < 1e-12
    rel = np.abs(mc["volumes"] - ex["volumes"]).sum() / ex["volumes"].sum()
    assert rel < 0.02
    assert np.array_equal(mc["nbr_ptr"], ex["nbr_ptr"]) and np.array_equal(mc["neighbors"], ex["neighbors"])
    tot_ex, tot_mc = ex["areas"].sum(), mc["areas"].sum()
    assert abs(tot_mc - tot_ex) < 0.02 * tot_ex
    m.free()


def test_partly_periodic_point_sample_mc(ctx):
    # periodic in x only: the x planes bound nothing, the y and z planes are boundary rows
    xs = synth.uniform(3, 1500, 53)
    m = build(ctx, xs, (0,))
    ex = ctx.integrate(m, "polytope")
    mc = ctx.integrate(m, "points", nsamples=20_000_000, seed=4)
    assert abs(mc["volumes"].sum() - 1.0) < 1e-12
    assert np.array_equal(mc["neighbors"], ex["neighbors"])
    n_ext = len(m.images()[0])
    planes = mc["neighbors"] > n_ext
    assert set(np.unique(mc["neighbors"][planes] - n_ext)) == {3, 4, 5, 6}
    # the boundary faces (y, z sides: total area 4) and all interfaces, within sampling noise
    assert abs(mc["areas"][planes].sum() - 4.0) < 0.02 * 4.0
    assert abs(ex["areas"][planes].sum() - 4.0) < 1e-10
    assert np.abs(mc["volumes"] - ex["volumes"]).sum() < 0.02
    m.free()


def test_periodic_directional_mc(ctx):
    xs = synth.uniform(3, 500, 61)
    m = build(ctx, xs, (0, 1, 2))
    ex = ctx.integrate(m, "polytope")
    mc = ctx.integrate(m, "montecarlo", ndir=2000, seed=5)
    assert np.array_equal(mc["neighbors"], ex["neighbors"])
    rel = np.abs(mc["volumes"] - ex["volumes"]).sum() / ex["volumes"].sum()
    assert rel < 0.05
    m.free()


def test_periodic_then_dirichlet_on_one_context(ctx):
    # the working set goes back to the originals when the boundary changes
    xs = synth.uniform(3, 300, 71)
    m = build(ctx, xs, (0, 1, 2))
    m.free()
    base, normal, bc = cuboid(3, periodic=[]).arrays()
    ctx.set_boundary(base, normal, bc)
    k, c = ctx.build_mesh(precision="fp64").export()
    ctx2 = Context(0)
    ctx2.set_points(xs)
    ctx2.set_boundary(base, normal, bc)
    k2, c2 = ctx2.build_mesh(precision="fp64").export()
    assert np.array_equal(k, k2) and np.array_equal(c, c2)
