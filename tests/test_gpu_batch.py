"""GPU parity of hv_build_mesh_batch (many independent meshes, one call, several streams).

Bar: every mesh of a batch equals the mesh that a single context builds from the same points
(keys and coordinates bit-exact), small meshes equal the oracle's vertex set, and a mesh that
fails reports its own status without stopping the others.
"""
import numpy as np
import pytest

from helpers import rows_view
from hvgpu import synth
from hvgpu._lib import Context, build_mesh_batch
from oracle_py import Oracle

pytestmark = pytest.mark.gpu


def single(ctx, xs, precision):
    b, nrm = synth.cube_planes(xs.shape[1])
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    m = ctx.build_mesh(precision=precision)
    out = m.export()
    m.free()
    return out


@pytest.mark.parametrize("precision,nworkers", [("fp64", 4), ("bf16", 8), ("bf16", 1), ("fp32", 0), ("bf16", -4)])
def test_batch_equals_single_builds(precision, nworkers):
    meshes = synth.batched_meshes(count=24, seed=3, nmax=1500)
    meshes.append((2, synth.uniform(2, 500, 11)))
    pts = [x for _, x in meshes]
    planes = [synth.cube_planes(x.shape[1]) for x in pts]
    out, status = build_mesh_batch(pts, planes, precision=precision, nworkers=nworkers)
    assert list(status) == [0] * len(pts)
    ctx = Context(0)
    for m, x in zip(out, pts):
        k, c = m.export()
        k1, c1 = single(ctx, x, precision)
        assert np.array_equal(k, k1)
        assert np.array_equal(c, c1)
        m.free()


def test_batch_small_meshes_match_oracle():
    pts = [synth.uniform(2, 300, 21), synth.uniform(3, 200, 22), synth.uniform(4, 120, 23)]
    planes = [synth.cube_planes(x.shape[1]) for x in pts]
    out, status = build_mesh_batch(pts, planes, precision="bf16", nworkers=3)
    assert list(status) == [0, 0, 0]
    for m, x in zip(out, pts):
        b, nrm = synth.cube_planes(x.shape[1])
        ok, oc, _ = Oracle(x, b, nrm).voronoi()
        k, c = m.export()
        assert np.array_equal(rows_view(k), rows_view(ok[np.lexsort(ok.T[::-1])]))
        m.free()


def test_batch_reports_failures_per_mesh():
    good = synth.uniform(3, 200, 5)
    few = synth.uniform(3, 3, 6)                       # n <= d: HV_ERR_FEW_POINTS
    outside = synth.uniform(3, 200, 7) + 2.0           # outside the unit cube: HV_ERR_DOMAIN
    pts = [good, few, outside, good]
    planes = [synth.cube_planes(3)] * 4
    out, status = build_mesh_batch(pts, planes, precision="fp64", nworkers=2)
    assert list(status) == [0, -3, -2, 0]
    assert out[1] is None and out[2] is None
    k0, c0 = out[0].export()
    k3, c3 = out[3].export()
    assert np.array_equal(k0, k3) and np.array_equal(c0, c3)
    out[0].free()
    out[3].free()


def test_c4_batch_of_512_meshes():
    # the C4 workload: 512 meshes, d uniform in {3..6}, n log-uniform in [200, 5000], unit cubes;
    # one fused walk per dimension.  Every mesh has the vertex count of its single build, 64 of
    # them (16 per dimension) equal their single builds bit for bit, and a cell subset of the
    # largest mesh of each dimension equals the oracle's vertices of those cells.
    meshes = synth.batched_meshes(count=512, seed=3)
    pts = [x for _, x in meshes]
    planes = [synth.cube_planes(x.shape[1]) for x in pts]
    out, status = build_mesh_batch(pts, planes, precision="bf16")
    assert list(status) == [0] * len(pts)
    ctx = Context(0)
    per_d = {}
    for k, (m, x) in enumerate(zip(out, pts)):
        per_d.setdefault(x.shape[1], []).append(k)
    check = set()
    for dd, ks in per_d.items():
        check.update(ks[:16])
    for k in sorted(check):
        kk, cc = out[k].export()
        k1, c1 = single(ctx, pts[k], "bf16")
        assert np.array_equal(kk, k1) and np.array_equal(cc, c1), k
    counts = [m.counts()[0] for m in out]
    assert sum(counts) > 100_000_000
    for dd, ks in per_d.items():
        big = max(ks, key=lambda k: len(pts[k]))
        x = pts[big]
        b, nrm = synth.cube_planes(dd)
        kk, _ = out[big].export()
        cells = (0, 20)
        ok, _, _ = Oracle(x, b, nrm).voronoi(cells=cells)
        sel = np.isin(kk, np.arange(1, 21)).any(axis=1)
        assert np.array_equal(rows_view(kk[sel]), rows_view(ok))
    for m in out:
        m.free()


def test_empty_batch():
    out, status = build_mesh_batch([], [], precision="fp64")
    assert out == [] and len(status) == 0


def test_batch_with_periodic_boxes_equals_single_builds():
    # per-mesh conditions: periodic, partly periodic and Dirichlet boxes in one batch
    from hvgpu.geometry import cuboid
    cases = [(synth.uniform(3, 400, 31), cuboid(3)), (synth.uniform(2, 300, 32), cuboid(2, periodic=[2])),
             (synth.uniform(4, 200, 33), cuboid(4, periodic=[])), (synth.uniform(5, 150, 34), cuboid(5, periodic=[1, 3]))]
    pts = [x for x, _ in cases]
    planes = [b.arrays() for _, b in cases]
    out, status = build_mesh_batch(pts, planes, precision="bf16", nworkers=3)
    assert list(status) == [0] * len(cases)
    ctx = Context(0)
    for m, (x, b) in zip(out, cases):
        base, normal, bc = b.arrays()
        ctx.set_points(x)
        ctx.set_boundary(base, normal, bc)
        m1 = ctx.build_mesh(precision="bf16")
        k, c = m.export()
        k1, c1 = m1.export()
        assert np.array_equal(k, k1) and np.array_equal(c, c1)
        o, sh = m.images()
        o1, sh1 = m1.images()
        assert np.array_equal(o, o1) and np.array_equal(sh, sh1)
        m.free()
        m1.free()


def test_batch_meshes_integrate_through_any_context():
    # a mesh keeps the generators and boundary it was built on: a context that holds other
    # points (or none) integrates batch meshes exactly as the single builds integrate
    from hvgpu.geometry import cuboid
    cases = [(synth.uniform(3, 300, 51), cuboid(3)), (synth.uniform(2, 250, 52), cuboid(2, periodic=[1])),
             (synth.uniform(4, 150, 53), cuboid(4, periodic=[])), (synth.uniform(3, 200, 54), cuboid(3, periodic=[2, 3]))]
    pts = [x for x, _ in cases]
    planes = [b.arrays() for _, b in cases]
    out, status = build_mesh_batch(pts, planes, precision="bf16", nworkers=4)
    assert list(status) == [0] * len(cases)
    other = Context(0)
    other.set_points(synth.uniform(5, 1000, 55))          # unrelated points, other dimension
    other.set_boundary(*synth.cube_planes(5))
    ctx = Context(0)
    for m, (x, b) in zip(out, cases):
        base, normal, bc = b.arrays()
        ctx.set_points(x)
        ctx.set_boundary(base, normal, bc)
        m1 = ctx.build_mesh(precision="bf16")
        for method, kw in (("polytope", {}), ("points", {"nsamples": 200_000, "seed": 3})):
            i_batch = other.integrate(m, method, **kw)
            i_single = ctx.integrate(m1, method, **kw)
            for key in ("volumes", "nbr_ptr", "neighbors", "areas"):
                assert np.array_equal(i_batch[key], i_single[key]), (method, key)
        m1.free()
    # the other context still builds its own points
    k_other = other.build_mesh()
    assert k_other.counts()[0] > 0
    k_other.free()
    for m in out:
        m.free()


def test_mesh_survives_new_points_on_its_context():
    xs = synth.uniform(3, 400, 61)
    ctx = Context(0)
    b, nrm = synth.cube_planes(3)
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    m = ctx.build_mesh()
    want = ctx.integrate(m, "polytope")
    ctx.set_points(synth.uniform(3, 400, 62))           # same n and d, other points
    ctx.set_boundary(b, nrm)
    got = ctx.integrate(m, "polytope")
    assert np.array_equal(got["volumes"], want["volumes"])
    assert np.array_equal(got["areas"], want["areas"])
    m.free()


def test_batch_meshes_exported_and_freed_from_many_threads():
    # meshes of one batch share one allocation; distinct meshes may be exported and freed from
    # different host threads at once (include/hv.h), the allocation going with the last of them
    import threading
    meshes = synth.batched_meshes(count=64, seed=9)
    pts = [x for _, x in meshes]
    planes = [synth.cube_planes(x.shape[1]) for x in pts]
    out, status = build_mesh_batch(pts, planes, precision="bf16")
    assert list(status) == [0] * len(pts)
    counts = [m.counts()[0] for m in out]
    got = [None] * len(out)
    errors = []

    def work(w):
        try:
            for k in range(w, len(out), 8):
                keys, _ = out[k].export()
                got[k] = len(keys)
                out[k].free()
        except Exception as e:   # reported below
            errors.append(e)

    th = [threading.Thread(target=work, args=(w,)) for w in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors and got == counts
    # the device is fine afterwards: the same batch again
    out2, status2 = build_mesh_batch(pts, planes, precision="bf16")
    assert list(status2) == [0] * len(pts) and [m.counts()[0] for m in out2] == counts
    for m in out2:
        m.free()
