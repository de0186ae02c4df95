"""GPU parity of hv_refine_mesh: refine!(VG, xs) (src/geometry.jl:404-438) through
systematic_refine! (src/meshrefine.jl:183-217), and of the DatabaseVertexStorage export
(hv_mesh_export_db, src/hvdatabase.jl:34-110, src/vdbdatabaseref.jl:88-145).

Bars: a refined mesh has exactly the keys of a fresh build of the union (old generators, then
the new ones) and of the oracle's walk over the union, coordinates within 1e-12 of the fresh
build (every coordinate comes from its key) and 1e-10 of the oracle; refining twice equals one
fresh build; the old mesh is untouched; a periodic mesh refined equals the periodic mesh of the union (by
generator labels) and drops exactly the old vertices whose circumsphere holds a new point or an
image of one; the exact integrals of the refined mesh equal those of
the fresh build (volumes and areas to 1e-12 relative, the same neighbour rows); the database
layout decodes back to the export and lists every (generator, vertex) incidence once, negative
for the key's first generator.
"""
import numpy as np
import pytest

from helpers import rows_view
from hvgpu import synth
from hvgpu.geometry import VI_POLYGON, VoronoiGeometry, cuboid
from hvgpu._lib import Context, HVError
from oracle_py import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def setup(ctx, xs, bc=None):
    base, normal = synth.cube_planes(xs.shape[1])
    ctx.set_points(xs)
    ctx.set_boundary(base, normal, bc)
    return base, normal


def fresh(xs, prec):
    c = Context(0)
    setup(c, xs)
    m = c.build_mesh(precision=prec)
    k, x = m.export()
    return c, m, k, x


def same_keys_and_coords(k1, x1, k2, x2, tol):
    assert np.array_equal(rows_view(k1), rows_view(k2))
    assert float(np.max(np.abs(x1 - x2))) <= tol


@pytest.mark.parametrize("prec", ["fp64", "bf16"])
@pytest.mark.parametrize("d,n,m,seed", [(2, 3000, 300, 11), (3, 2000, 40, 12), (4, 1500, 150, 13),
                                        (5, 600, 7, 14), (6, 300, 30, 15)])
def test_refine_equals_fresh_build_of_the_union(ctx, d, n, m, seed, prec):
    xs = synth.uniform(d, n + m, seed)
    setup(ctx, xs[:n])
    old = ctx.build_mesh(precision=prec)
    k0, x0 = old.export()
    ref, removed = ctx.refine(old, xs[n:], precision=prec)
    k1, x1 = ref.export()
    assert removed > 0 and ctx.n == n + m
    _, fm, kf, xf = fresh(xs, prec)
    same_keys_and_coords(k1, x1, kf, xf, 1e-12)
    # the old mesh is unchanged: its keys and coordinates, and plane ids n + p
    k0b, x0b = old.export()
    assert np.array_equal(k0, k0b) and np.array_equal(x0, x0b)
    # the kept old vertices are the refined rows without a new id, with plane ids moved up by m
    kept = k1[~((k1 > n) & (k1 <= n + m)).any(axis=1)]
    assert len(kept) == len(k0) - removed
    back, old_rows = rows_view(np.where(kept > n + m, kept - m, kept)), rows_view(k0)
    pos = np.minimum(np.searchsorted(old_rows, back), len(old_rows) - 1)
    assert np.all(old_rows[pos] == back)
    old.free(); ref.free(); fm.free()


def test_refine_matches_the_oracle_and_refines_twice(ctx):
    d, n = 3, 800
    xs = synth.uniform(d, n + 120, 21)
    base, normal = setup(ctx, xs[:n])
    m0 = ctx.build_mesh()
    m1, _ = ctx.refine(m0, xs[n:n + 60])
    m2, _ = ctx.refine(m1, xs[n + 60:])
    k2, x2 = m2.export()
    ok, oc, _ = Oracle(xs, base, normal).voronoi()
    assert np.array_equal(rows_view(k2), rows_view(ok))
    row = {k: i for i, k in enumerate(map(tuple, k2.tolist()))}
    j = np.array([row[k] for k in map(tuple, ok.tolist())])
    assert float(np.max(np.abs(x2[j] - oc))) < 1e-10
    _, fm, kf, xf = fresh(xs, "fp64")
    same_keys_and_coords(k2, x2, kf, xf, 1e-12)
    for mm in (m0, m1, m2, fm):
        mm.free()


def test_refined_integrals_equal_fresh_integrals(ctx):
    d, n, m = 4, 2000, 200
    xs = synth.uniform(d, n + m, 31)
    setup(ctx, xs[:n])
    m0 = ctx.build_mesh(precision="bf16")
    m1, _ = ctx.refine(m0, xs[n:], precision="bf16")
    I = ctx.integrate(m1, "polytope")
    c2, fm, _, _ = fresh(xs, "bf16")
    J = c2.integrate(fm, "polytope")
    assert np.array_equal(I["nbr_ptr"], J["nbr_ptr"]) and np.array_equal(I["neighbors"], J["neighbors"])
    assert np.allclose(I["volumes"], J["volumes"], rtol=1e-12, atol=0)
    assert np.allclose(I["areas"], J["areas"], rtol=1e-12, atol=1e-14)
    assert abs(I["volumes"].sum() - 1.0) < 1e-12
    for mm in (m0, m1, fm):
        mm.free()


def test_voronoi_geometry_refine(ctx):
    d, n, m = 3, 1500, 100
    xs = synth.uniform(d, n + m, 41)
    box = cuboid(d, periodic=[])
    VG = VoronoiGeometry(xs[:n], box, integrator=VI_POLYGON, search_settings={"precision": "bf16"})
    VG.refine(xs[n:])
    ref = VoronoiGeometry(xs, box, integrator=VI_POLYGON, search_settings={"precision": "bf16"})
    k1, _ = VG.vertex_table()
    k2, _ = ref.vertex_table()
    assert np.array_equal(k1, k2)
    assert np.allclose(VG.integral["volumes"], ref.integral["volumes"], rtol=1e-12, atol=0)
    assert VG.nodes.shape == (n + m, d)


@pytest.mark.parametrize("d,n,m,per,seed", [(2, 600, 40, None, 71), (3, 400, 25, [1], 72), (4, 200, 10, [1, 3], 73)])
def test_refine_periodic_equals_fresh_build(ctx, d, n, m, per, seed):
    # a periodic mesh refined equals the periodic mesh of the union (labels and coordinates),
    # and `removed` counts the old vertices (by label) the refined mesh no longer holds
    from helpers import periodic_mesh_labels
    xs = synth.uniform(d, n + m, seed)
    box = cuboid(d) if per is None else cuboid(d, periodic=per)
    base, normal, bc = box.arrays()
    ctx.set_points(xs[:n])
    ctx.set_boundary(base, normal, bc)
    old = ctx.build_mesh(precision="bf16")
    lab_old = periodic_mesh_labels(old)
    ref, removed = ctx.refine(old, xs[n:], precision="bf16")
    assert ctx.n == n + m
    lab_ref = periodic_mesh_labels(ref)
    c2 = Context(0)
    c2.set_points(xs)
    c2.set_boundary(base, normal, bc)
    fresh = c2.build_mesh(precision="bf16")
    lab_fresh = periodic_mesh_labels(fresh)
    assert set(lab_ref) == set(lab_fresh)
    assert max(float(np.max(np.abs(lab_ref[k] - lab_fresh[k]))) for k in lab_ref) == 0.0
    assert removed == len(set(lab_old) - set(lab_ref)) > 0
    # the kept old vertices are those whose circumsphere holds no new point nor an image of one
    # (clean_affected!, src/meshrefine.jl:126-150): count them from the old coordinates
    L = 1.0
    shifts = np.array(np.meshgrid(*[[-1, 0, 1] if (per is None or k + 1 in per) else [0] for k in range(d)],
                                  indexing="ij")).reshape(d, -1).T * L
    new_imgs = (xs[n:][:, None, :] + shifts[None, :, :]).reshape(-1, d)
    keys_old = list(lab_old)
    kept = 0
    for k in keys_old:
        r = lab_old[k]
        g = k[0]   # a generator label (0, orig, *shift) sorts first
        x0 = xs[g[1]] + np.array(g[2:]) * L
        R = np.linalg.norm(r - x0)
        dmin = np.min(np.linalg.norm(new_imgs - r, axis=1))
        kept += R <= (1 + 1e-7) * dmin
    assert kept == len(keys_old) - removed
    for mm in (old, ref, fresh):
        mm.free()


def test_refine_refusals(ctx):
    d = 3
    xs = synth.uniform(d, 400, 51)
    setup(ctx, xs)
    sub = ctx.build_mesh_cells([1, 2, 3])
    with pytest.raises(HVError) as e:
        ctx.refine(sub, synth.uniform(d, 5, 53))
    assert e.value.code == -7
    sub.free()
    setup(ctx, xs)
    full = ctx.build_mesh()
    with pytest.raises(HVError) as e:
        ctx.refine(full, np.array([[0.5, 0.5, 1.5]]))
    assert e.value.code == -2
    full.free()


@pytest.mark.parametrize("d,n,seed", [(2, 500, 61), (4, 300, 62)])
def test_database_layout(ctx, d, n, seed):
    xs = synth.uniform(d, n, seed)
    setup(ctx, xs)
    m = ctx.build_mesh()
    keys, coords = m.export()
    heap, ptr, refs = m.export_db()
    m.free()
    W = 2 * d + 2
    nv = len(keys)
    H = heap.reshape(nv, W)
    assert np.all(H[:, 0] == d + 1)
    assert np.array_equal(H[:, 1:d + 2], keys)
    assert np.array_equal(H[:, d + 2:].view(np.float64), coords)
    assert len(ptr) == n + 1 and ptr[-1] == len(refs) == int(np.sum(keys <= n))
    pos = np.arange(nv, dtype=np.int64) * W + 1
    for i in range(1, n + 1):
        got = refs[ptr[i - 1]:ptr[i]]
        rows = np.nonzero((keys == i).any(axis=1))[0]
        want = np.where(keys[rows, 0] == i, -pos[rows], pos[rows])
        assert np.array_equal(np.sort(got), np.sort(want))
