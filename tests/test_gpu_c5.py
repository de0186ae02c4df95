"""6D N=50 000 uniform in the periodic unit cube cuboid(6) (src/boundary.jl:510: periodic in every
direction): exact FP64 polytope integration and 1e8 point samples on the same mesh, and the
exact integrator on meshes of some cells.

The working set holds the 50 000 originals and about 3.0M periodic images.  The exact
integrator numbers the ids that occur in the mesh's keys densely and integrates a chunk of
cells at a time when one pass would not fit.

Bars:
  * exact: the originals' cells tile the torus (volumes sum to 1 within 1e-12), every interface
    is measured from both sides with the same area (relative 1e-9), every volume positive;
  * point samples (the same mesh): every sample counted once (volumes sum to 1 within 1e-12),
    every cell volume within 6 binomial standard deviations of the exact one, the L1 volume
    error within 1.1 times its expectation under binomial noise, the same neighbour rows, the
    total interface area within 3% of the exact total;
  * cells: the tiled oracle (tests/helpers.periodic_oracle restricted to a prefix of cells) at 5D
    and 6D equals the device's mesh of those cells, and the exact integrals of a mesh of some
    cells equal those of the full mesh for those cells, bit for bit.
"""
import numpy as np
import pytest

from helpers import periodic_mesh_labels, periodic_tiling, vertex_labels
from hvgpu import synth
from hvgpu.geometry import cuboid
from hvgpu._lib import Context
from oracle_py import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def periodic_setup(ctx, xs, periodic=None):
    d = xs.shape[1]
    b = cuboid(d) if periodic is None else cuboid(d, periodic=periodic)
    base, normal, bc = b.arrays()
    ctx.set_points(xs)
    ctx.set_boundary(base, normal, bc)


def shift_code(s):
    s = np.asarray(s, dtype=np.int64) + 8
    return np.sum(s << (4 * np.arange(s.shape[1], dtype=np.int64)), axis=1)


def test_c5_exact_and_point_samples_on_the_same_mesh(ctx):
    n, d = 50_000, 6
    xs = synth.uniform(d, n, 4)
    periodic_setup(ctx, xs)
    m = ctx.build_mesh(precision="bf16")
    orig, shift = m.images()
    n_ext = len(orig)
    assert n_ext > 2_000_000
    ex = ctx.integrate(m, "polytope")
    vol, ptr, nbr, area = ex["volumes"], ex["nbr_ptr"], ex["neighbors"], ex["areas"]
    assert len(vol) == n and np.all(vol > 0)
    assert abs(vol.sum() - 1.0) < 1e-12
    # interface (i, image of o at shift s) seen from o: (o, image of i at -s), the same area
    cell = np.repeat(np.arange(n, dtype=np.int64), np.diff(ptr))
    assert np.all(nbr <= n_ext)               # periodic in every direction: no plane rows
    o = orig[nbr - 1] - 1
    s = shift[nbr - 1].astype(np.int64)
    fwd = (cell << 41) | (o << 24) | shift_code(s)
    bwd = (o << 41) | (cell << 24) | shift_code(-s)
    order = np.argsort(fwd)
    pos = np.searchsorted(fwd[order], bwd)
    assert np.all(pos < len(fwd))
    j = order[pos]
    assert np.array_equal(fwd[j], bwd)
    assert np.all(np.abs(area[j] - area) <= 1e-9 * np.maximum(area, 1e-12))
    # 1e8 point samples on the same mesh
    ns = 100_000_000
    mc = ctx.integrate(m, "points", nsamples=ns, seed=5)
    m.free()
    assert abs(mc["volumes"].sum() - 1.0) < 1e-12
    assert np.array_equal(mc["nbr_ptr"], ptr) and np.array_equal(mc["neighbors"], nbr)
    sigma = np.sqrt(vol * (1.0 - vol) / ns)
    dv = np.abs(mc["volumes"] - vol)
    assert float(np.max(dv / sigma)) < 6.0
    assert dv.sum() < 1.1 * np.sqrt(2.0 / np.pi) * sigma.sum()
    assert abs(mc["areas"].sum() - area.sum()) < 0.03 * area.sum()


def prefix_oracle(xs, per, b):
    """{label: coordinates} of the cells 1..b of a periodic box, from the oracle's walk over
    Iter = 1:b of the tiled set (originals first, so these are the originals' cells)."""
    tiled, orig, shift, base, normal = periodic_tiling(xs, per)
    keys, coords, _ = Oracle(tiled, base, normal).voronoi(cells=(0, b))
    return dict(zip(vertex_labels(keys, len(tiled), orig, shift), coords))


@pytest.mark.parametrize("d,n,b,seed", [(5, 3000, 100, 81), (6, 1500, 20, 82)])
def test_periodic_cells_match_tiled_oracle(ctx, d, n, b, seed):
    xs = synth.uniform(d, n, seed)
    periodic_setup(ctx, xs)
    m = ctx.build_mesh_cells(np.arange(1, b + 1), precision="bf16")
    mine = periodic_mesh_labels(m)
    m.free()
    ref = prefix_oracle(xs, tuple(range(d)), b)
    assert set(mine) == set(ref)
    assert max(float(np.max(np.abs(mine[k] - ref[k]))) for k in ref) < 1e-10


@pytest.mark.parametrize("d,n,per,seed", [(4, 1500, [], 91), (3, 800, None, 92), (5, 600, [1, 4], 93)])
def test_exact_integrals_of_some_cells(ctx, d, n, per, seed):
    xs = synth.uniform(d, n, seed)
    periodic_setup(ctx, xs, per)
    full = ctx.build_mesh()
    I = ctx.integrate(full, "polytope")
    lab_full = full.images()
    full.free()
    rng = np.random.default_rng(seed)
    cells = np.sort(rng.choice(n, size=n // 10, replace=False)) + 1
    m = ctx.build_mesh_cells(cells)
    J = ctx.integrate(m, "polytope")
    lab_sub = m.images()
    m.free()
    listed = np.zeros(n, dtype=bool)
    listed[cells - 1] = True
    same_set = len(lab_full[0]) == len(lab_sub[0])   # the same periodic margin: the same ids
    if same_set:
        assert np.array_equal(J["volumes"][listed], I["volumes"][listed])
    else:
        assert np.allclose(J["volumes"][listed], I["volumes"][listed], rtol=1e-12, atol=0)
    assert np.all(J["volumes"][~listed] == 0.0)

    def rows(R, lab, i):
        orig, shift = lab
        a, b = R["nbr_ptr"][i], R["nbr_ptr"][i + 1]
        out = {}
        for j, ar in zip(R["neighbors"][a:b], R["areas"][a:b]):
            key = (int(orig[j - 1]), tuple(shift[j - 1])) if j <= len(orig) else ("plane", int(j - len(orig)))
            out[key] = ar
        return out

    for i in np.nonzero(listed)[0]:
        ri, rj = rows(I, lab_full, i), rows(J, lab_sub, i)
        assert set(ri) == set(rj)
        for k in ri:
            assert (ri[k] == rj[k]) if same_set else abs(ri[k] - rj[k]) <= 1e-12 * max(ri[k], 1e-12)
    assert J["nbr_ptr"][-1] == int(np.sum(np.diff(I["nbr_ptr"])[listed]))
