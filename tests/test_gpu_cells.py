"""GPU parity of hv_build_mesh_cells: the vertices of some cells only, the reference's
voronoi(mesh; Iter=cells) (src/sysvoronoi.jl:21-39) as systematic_refine! calls it
(src/meshrefine.jl:199,213,215).

Bar: the key set is bit-exact against the oracle.  A prefix of cells 1..b is exactly what the
oracle's own walk over Iter = 1:b returns (its walk order makes every cell of the prefix
complete); an arbitrary list is the oracle's full vertex set restricted to the keys that hold a
listed cell.  Coordinates are bit-identical to the full device build (every vertex's
coordinates come from its key) and within 1e-10 of the oracle.
"""
import numpy as np
import pytest

from helpers import periodic_mesh_labels, rows_view
from hvgpu import synth
from hvgpu.geometry import cuboid
from hvgpu._lib import Context, HVError
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

COORD_TOL = 1e-10


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def holding(keys, cells):
    """Rows of a 1-based key table that hold one of `cells` (1-based)."""
    return np.isin(keys, np.asarray(cells)).any(axis=1)


def setup(ctx, xs, bc=None):
    base, normal = synth.cube_planes(xs.shape[1])
    ctx.set_points(xs)
    ctx.set_boundary(base, normal, bc)
    return base, normal


def coords_match(k1, c1, k2, c2, tol):
    idx = {k: i for i, k in enumerate(map(tuple, k2.tolist()))}
    j = np.array([idx[k] for k in map(tuple, k1.tolist())], dtype=np.int64)
    return float(np.max(np.abs(c1 - c2[j]))) <= tol if len(j) else True


@pytest.mark.parametrize("prec", ["fp64", "bf16"])
@pytest.mark.parametrize("d,n,b,seed", [(2, 2000, 300, 1), (3, 3000, 500, 2), (4, 1500, 200, 3), (5, 800, 60, 4)])
def test_prefix_equals_oracle_walk_over_the_prefix(ctx, d, n, b, seed, prec):
    xs = synth.uniform(d, n, seed)
    base, normal = setup(ctx, xs)
    m = ctx.build_mesh_cells(np.arange(1, b + 1), precision=prec)
    k, c = m.export()
    m.free()
    ok, oc, _ = Oracle(xs, base, normal).voronoi(cells=(0, b))
    assert len(k) == len(ok) > 0
    assert np.array_equal(rows_view(k), rows_view(ok))
    assert coords_match(k, c, ok, oc, COORD_TOL)
    assert holding(k, np.arange(1, b + 1)).all()


@pytest.mark.parametrize("d,n,seed", [(2, 1500, 11), (3, 2000, 12), (4, 1000, 13), (5, 600, 14), (6, 200, 15)])
def test_scattered_cells_equal_the_restricted_full_mesh(ctx, d, n, seed):
    xs = synth.uniform(d, n, seed)
    base, normal = setup(ctx, xs)
    rng = np.random.default_rng(seed)
    # scattered cells, the cells nearest to the box's faces and corners (their fast descents
    # meet a plane first), repeats and an unsorted order
    near = np.argsort(np.min(np.minimum(xs, 1.0 - xs), axis=1))[:12] + 1
    cells = np.concatenate([rng.choice(n, size=min(40, n // 4), replace=False) + 1, near, near[:3]])
    rng.shuffle(cells)
    m = ctx.build_mesh_cells(cells, precision="bf16")
    k, c = m.export()
    m.free()
    full = ctx.build_mesh(precision="fp64")
    kf, cf = full.export()
    full.free()
    sel = holding(kf, cells)
    assert np.array_equal(k, kf[sel])          # both exports are in lexicographic key order
    assert np.array_equal(c, cf[sel])
    ok, _, _ = Oracle(xs, base, normal).voronoi()
    assert np.array_equal(rows_view(k), rows_view(ok[holding(ok, cells)]))


def test_single_boundary_cell(ctx):
    # one cell at a corner of the box: no interior seed, the exact descent supplies its vertex
    xs = synth.uniform(3, 2000, 21)
    corner = int(np.argmin(np.sum(xs, axis=1))) + 1
    setup(ctx, xs)
    m = ctx.build_mesh_cells([corner])
    k, _ = m.export()
    m.free()
    full = ctx.build_mesh()
    kf, _ = full.export()
    full.free()
    assert np.array_equal(k, kf[holding(kf, [corner])])
    assert np.any(k > 2000)                     # the corner cell has boundary vertices


def test_periodic_cells_equal_the_restricted_periodic_mesh(ctx):
    xs = synth.uniform(3, 300, 31)
    base, normal, bc = cuboid(3).arrays()
    ctx.set_points(xs)
    ctx.set_boundary(base, normal, bc)
    full = periodic_mesh_labels(ctx.build_mesh())
    cells = np.array([1, 7, 50, 151, 299])
    m = ctx.build_mesh_cells(cells)
    mine = periodic_mesh_labels(m)
    m.free()
    # vertices holding a listed original itself (shift 0): its cell in its own frame
    listed = set(cells.tolist())
    want = {lab: x for lab, x in full.items()
            if any(g[0] == 0 and g[1] + 1 in listed and not any(g[2:]) for g in lab)}
    assert set(mine) == set(want)
    assert max(float(np.max(np.abs(mine[k] - want[k]))) for k in want) < COORD_TOL


def test_empty_list_and_bad_ids(ctx):
    xs = synth.uniform(3, 200, 41)
    setup(ctx, xs)
    m = ctx.build_mesh_cells([])
    assert m.counts()[0] == 0
    m.free()
    for bad in ([0], [201], [-3]):
        with pytest.raises(HVError) as e:
            ctx.build_mesh_cells(bad)
        assert e.value.code == -1


def test_monte_carlo_integrators_refuse_a_mesh_of_some_cells(ctx):
    # the polytope integrators take the listed cells (tests/test_gpu_c5.py); the Monte-Carlo
    # ones need every cell's neighbours
    xs = synth.uniform(2, 300, 42)
    setup(ctx, xs)
    m = ctx.build_mesh_cells([1, 2, 3])
    for method in ("montecarlo", "points"):
        with pytest.raises(HVError) as e:
            ctx.integrate(m, method)
        assert e.value.code == -7
    m.free()
