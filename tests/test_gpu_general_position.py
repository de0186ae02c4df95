"""GPU parity at the edge of general position (the reference's winner rule,
src/raycast.jl:877-950).

Bars: on a constructed near tie the device resolves the ray as the oracle's walkray does
(same key, status 2 = near tie resolved, coordinates within 1e-10), and reports no tie where
the oracle has none.  With the candidates swapped (the farther one has the larger <u, x - x0>)
the reference's result depends on the order of its range query: its single pass over the
candidates (src/raycast.jl:883-894) lowers the window's upper end to each t it meets, so a
farther candidate met after the nearer one is dropped (the oracle's order: no tie), and one
met before it stays and wins, making the nearer one, inside the winner's sphere, a further
member of the key (a vertex out of general position).  The device takes every candidate
within the window whatever the order and reports that vertex as out of general position
(status < 0).  Perturbed lattices that the oracle walks in general position give the
oracle's key set bit for bit in every precision; an exact lattice, where the reference's keys
grow past d + 1, fails with HV_ERR_DEGENERATE (-4) after every kernel has exited, and the
context builds the next mesh normally.
"""
import numpy as np
import pytest

from helpers import lattice, near_tie_points, rows_view
from hvgpu import synth
from hvgpu._lib import Context, HVError
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

PRECISIONS = ["fp64", "fp32", "bf16"]


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("delta,tie", [(1e-13, False), (3e-15, True), (1e-15, True)])
def test_constructed_near_tie(ctx, prec, delta, tie):
    xs, r = near_tie_points(delta, False)
    b, nrm = synth.cube_planes(2)
    o = Oracle(xs, b, nrm)
    key, r2, _ = o.walkray(np.array([1, 2, 3]), 2, r, 1)
    assert (o.walkray_stats()["near_ties"] == 1) == tie
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    out_sig, out_r, _, out_st = ctx.raycast_batch(np.array([[1, 2, 3]]), np.array([2]), r[None, :], precision=prec)
    assert out_sig[0].tolist() == key.tolist()
    assert int(out_st[0]) == (2 if tie else 0)
    assert np.max(np.abs(out_r[0] - r2)) < 1e-10


@pytest.mark.parametrize("prec", PRECISIONS)
def test_swapped_near_tie_is_not_general(ctx, prec):
    xs, r = near_tie_points(3e-15, True)
    b, nrm = synth.cube_planes(2)
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    _, _, _, out_st = ctx.raycast_batch(np.array([[1, 2, 3]]), np.array([2]), r[None, :], precision=prec)
    assert int(out_st[0]) < 0


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("d,k,eps", [(2, 30, 1e-9), (2, 30, 1e-10), (2, 30, 5e-11), (2, 30, 2e-11), (3, 8, 1e-5)])
def test_perturbed_lattice_equals_oracle(ctx, prec, d, k, eps):
    xs = lattice(d, k, eps)
    b, nrm = synth.cube_planes(d)
    ok, oc, _ = Oracle(xs, b, nrm).voronoi()
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    m = ctx.build_mesh(precision=prec)
    keys, coords = m.export()
    st = m.stats()
    m.free()
    assert np.array_equal(rows_view(keys), rows_view(ok))
    assert st["degenerate"] == 0


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("d,k", [(2, 30), (3, 8)])
def test_exact_lattice_is_degenerate(ctx, prec, d, k):
    xs = lattice(d, k)
    b, nrm = synth.cube_planes(d)
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    with pytest.raises(HVError) as e:
        ctx.build_mesh(precision=prec)
    assert e.value.code == -4
    # the context is usable afterwards
    ys = synth.uniform(d, 500, 7)
    ctx.set_points(ys)
    m = ctx.build_mesh(precision=prec)
    keys, _ = m.export()
    m.free()
    ok, _, _ = Oracle(ys, b, nrm).voronoi()
    assert np.array_equal(rows_view(keys), rows_view(ok))


@pytest.mark.parametrize("prec", ["fp64", "bf16"])
def test_clipped_mixture_is_degenerate(ctx, prec):
    # generators on a boundary plane (the clipped 5D N=100k mixture): not in general position
    xs = synth.gaussian_mixture_5d(clip=True)
    b, nrm = synth.cube_planes(5)
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    with pytest.raises(HVError) as e:
        ctx.build_mesh(precision=prec)
    assert e.value.code == -4


def test_batch_mesh_with_a_generator_on_a_face_is_degenerate():
    from hvgpu._lib import build_mesh_batch
    pts = [synth.uniform(3, 400, 71), synth.uniform(3, 300, 72), synth.uniform(4, 500, 73)]
    pts[1][5, 2] = 0.0   # on the face x_3 = 0
    planes = [synth.cube_planes(x.shape[1]) for x in pts]
    out, status = build_mesh_batch(pts, planes, precision="bf16")
    assert list(status) == [0, -4, 0]
    for k in (0, 2):
        b, nrm = planes[k]
        ok, _, _ = Oracle(pts[k], b, nrm).voronoi()
        keys, _ = out[k].export()
        assert np.array_equal(rows_view(keys), rows_view(ok))
    for m in out:
        if m is not None:
            m.free()


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("d,k", [(2, 8), (3, 4)])
def test_degenerate_inclusion_matches_oracle(ctx, prec, d, k):
    """hv_raycast_batch_keys on rays of an exact lattice that end at vertices with more than
    d + 1 generators (and planes) on their sphere: the oracle's walkray keys, status 3, the
    same vertex; the rays that stay general give hv_raycast_batch's answer."""
    from helpers import lattice_boundary_rays
    xs = lattice(d, k)
    b, nrm = synth.cube_planes(d)
    o = Oracle(xs, b, nrm)
    sig, drop, r, cell = lattice_boundary_rays(d, k)
    want = [o.walkray(sig[q], int(drop[q]), r[q], cell[q]) for q in range(len(sig))]
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    keys, lens, out_r, _, st = ctx.raycast_batch_keys(sig, drop, r, max_sig=16, precision=prec)
    for q in range(len(sig)):
        wk, wr, _ = want[q]
        assert int(st[q]) == 3 and int(lens[q]) == len(wk) > d + 1
        assert keys[q, :lens[q]].tolist() == wk.tolist() and not keys[q, lens[q]:].any()
        assert np.max(np.abs(out_r[q] - wr)) < 1e-12
    # general rays (from random points) are unchanged
    ys = synth.uniform(d, 300, 5)
    ctx.set_points(ys)
    m = ctx.build_mesh(precision=prec)
    kk, cc = m.export()
    m.free()
    rows = np.nonzero(kk[:, 1] <= 300)[0][::7][:60]   # edges (drop position 0) holding a generator
    gs = kk[rows]
    gd = np.zeros(len(rows), dtype=np.int32)
    a = ctx.raycast_batch(gs, gd, cc[rows], precision=prec)
    keys2, lens2, out_r2, _, st2 = ctx.raycast_batch_keys(gs, gd, cc[rows], max_sig=d + 1, precision=prec)
    assert np.array_equal(a[3], st2)
    ok = st2 != 1
    assert np.array_equal(a[0][ok], keys2[ok]) and np.all(lens2[ok] == d + 1)
