"""GPU parity tests of the device walk (libhv.so through the C ABI) against the oracle.

Bar: the vertex key sets are identical (bit-exact integer work); vertex coordinates
agree to 1e-10 relative (the tolerance the path promises for the FP64 kernels).
"""
import numpy as np
import pytest

from helpers import (edge_closure_defects, golden_cases, key_digest, key_set, load_golden, rows_view)
from hvgpu import synth
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

COORD_TOL = 1e-10


@pytest.fixture(scope="module")
def ctx():
    from hvgpu._lib import Context
    return Context(0)


def build(ctx, xs, precision="fp64"):
    d = xs.shape[1]
    b, nrm = synth.cube_planes(d)
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    m = ctx.build_mesh(precision=precision)
    keys, coords = m.export()
    st = m.stats()
    m.free()
    return keys, coords, st


def max_rel_coord_err(k1, c1, k2, c2):
    idx = {k: i for i, k in enumerate(map(tuple, k2.tolist()))}
    j = np.array([idx[k] for k in map(tuple, k1.tolist())])
    return float(np.max(np.abs(c1 - c2[j]) / np.maximum(1.0, np.abs(c2[j]))))


@pytest.mark.parametrize("name", golden_cases())
def test_mesh_matches_exact_fixture(ctx, name):
    g, xs, keys, exact = load_golden(name)
    k, c, st = build(ctx, xs)
    assert key_set(k) == key_set(keys)
    ex = np.array([[float(v) for v in row] for row in exact])
    assert max_rel_coord_err(k, c, keys, ex) < 1e-12
    assert st["near_ties"] == 0


@pytest.mark.parametrize("d,n,seed", [(2, 1000, 0), (3, 3000, 5), (4, 2000, 1), (5, 1000, 9), (6, 300, 8)])
def test_mesh_matches_oracle(ctx, d, n, seed):
    xs = synth.uniform(d, n, seed)
    k, c, st = build(ctx, xs)
    b, nrm = synth.cube_planes(d)
    ok, oc, _ = Oracle(xs, b, nrm).voronoi()
    assert len(k) == len(ok)
    assert np.array_equal(rows_view(k), rows_view(ok))
    assert max_rel_coord_err(k, c, ok, oc) < COORD_TOL
    assert st["vertices"] == len(k)
    # every ray ended at a vertex: none unbounded in the box, no singular circumcentre
    assert st["unbounded"] == 0 and st["degenerate"] == 0


def test_config_4d_20k_matches_oracle(ctx):
    # 4D, N=20 000 uniform, cuboid(4, periodic=[]): the full vertex set, bit-exact
    xs = synth.uniform(4, 20000, 1)
    k, c, st = build(ctx, xs)
    b, nrm = synth.cube_planes(4)
    ok, oc, _ = Oracle(xs, b, nrm).voronoi()
    assert np.array_equal(rows_view(k), rows_view(ok))
    assert max_rel_coord_err(k, c, ok, oc) < COORD_TOL
    assert edge_closure_defects(k, 20000) == 0


def test_small_inputs_and_tiny_n(ctx):
    # n = d + 1 generators: the smallest valid input
    for d in (2, 3, 4, 5, 6):
        xs = synth.uniform(d, d + 1, 40 + d)
        k, c, _ = build(ctx, xs)
        b, nrm = synth.cube_planes(d)
        ok, _, _ = Oracle(xs, b, nrm).voronoi()
        assert key_set(k) == key_set(ok)


def test_rejects_points_outside_domain(ctx):
    from hvgpu._lib import HVError
    xs = synth.uniform(3, 100, 3)
    xs[7, 1] = 1.5
    with pytest.raises(HVError) as e:
        build(ctx, xs)
    assert e.value.code == -2


def test_rejects_too_few_points(ctx):
    from hvgpu._lib import HVError
    with pytest.raises(HVError) as e:
        ctx.set_points(synth.uniform(3, 3, 1))
    assert e.value.code == -3


def test_raycast_batch_matches_oracle_walkray(ctx):
    d, n = 4, 500
    xs = synth.uniform(d, n, 23)
    b, nrm = synth.cube_planes(d)
    o = Oracle(xs, b, nrm)
    ok, oc, _ = o.voronoi()
    rng = np.random.default_rng(1)
    sig, drop, r, want = [], [], [], []
    for i in rng.permutation(len(ok))[:400]:
        for p in range(d + 1):
            e = np.delete(ok[i], p)
            if e[0] > n:
                continue
            res = o.walkray(ok[i], p, oc[i], int(e[0]))
            sig.append(ok[i]); drop.append(p); r.append(oc[i]); want.append(res)
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    out_sig, out_r, out_t, out_st = ctx.raycast_batch(np.array(sig), np.array(drop), np.array(r))
    assert np.all(out_st == 0)
    for k in range(len(sig)):
        assert tuple(out_sig[k].tolist()) == tuple(want[k][0].tolist())
        assert np.max(np.abs(out_r[k] - want[k][1])) < 1e-10
    assert np.all(out_t > 0)


def test_nn_and_inrange_match_brute_force(ctx):
    d, n = 5, 3000
    xs = synth.uniform(d, n, 31)
    ctx.set_points(xs)
    rng = np.random.default_rng(2)
    q = rng.random((500, d))
    u = rng.normal(size=(500, d))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    c = np.einsum("ij,ij->i", q, u) + rng.normal(scale=0.05, size=500)
    c[:50] = -np.inf
    idx, dist = ctx.nn_halfspace(q, u, c)
    D = np.linalg.norm(xs[None, :, :] - q[:, None, :], axis=2)
    ok = (xs @ u.T).T > c[:, None]
    Dm = np.where(ok, D, np.inf)
    want = np.argmin(Dm, axis=1) + 1
    has = np.isfinite(Dm.min(axis=1))
    assert np.array_equal(idx[has], want[has])
    assert np.all(idx[~has] == 0)
    assert np.allclose(dist[has], Dm.min(axis=1)[has], rtol=0, atol=1e-14)
    rad = rng.uniform(0.05, 0.3, size=500)
    ptr, ids = ctx.inrange(q, rad)
    for k in range(500):
        got = ids[ptr[k]:ptr[k + 1]]
        exp = np.nonzero(D[k] <= rad[k])[0] + 1
        assert np.array_equal(got, exp)


def test_table_growth_paths_give_the_same_mesh(ctx):
    # a tiny capacity hint forces vertex-table growth, vertex-hash and edge-hash rehashes
    # (with re-lookup of stored edge slots) in the middle of the walk
    xs = synth.uniform(4, 3000, 91)
    b, nrm = synth.cube_planes(4)
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    m = ctx.build_mesh(vertex_capacity=500)
    k1, c1 = m.export()
    m.free()
    k2, c2, _ = build(ctx, xs)
    assert np.array_equal(k1, k2)
    assert np.array_equal(c1, c2)


@pytest.mark.parametrize("chunks", [1, 2, 7])
def test_chunking_does_not_change_the_mesh(ctx, chunks):
    xs = synth.uniform(5, 3000, 92)
    b, nrm = synth.cube_planes(5)
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    m = ctx.build_mesh(chunks=chunks)
    k1, c1 = m.export()
    m.free()
    ok, oc, _ = Oracle(xs, b, nrm).voronoi()
    assert np.array_equal(rows_view(k1), rows_view(ok))


def test_repeated_builds_are_identical(ctx):
    xs = synth.uniform(5, 2000, 77)
    k1, c1, _ = build(ctx, xs)
    k2, c2, _ = build(ctx, xs)
    assert key_digest(k1) == key_digest(k2)
    assert np.array_equal(k1, k2)
    assert np.array_equal(c1, c2)


def test_interior_seeds_keep_the_walk_shallow(ctx):
    # descent seeds every few generators: the breadth-first walk needs a handful of levels
    # instead of a number growing like N^(1/d) from the box corners alone
    xs = synth.uniform(2, 200_000, 3)
    k, c, st = build(ctx, xs)
    assert st["vertices"] == len(k) == 2 * 200_000 + 2
    assert st["levels"] <= 20
    assert st["unbounded"] == 0 and st["degenerate"] == 0


@pytest.mark.parametrize("ppw,floor,pmin", [(8, 64, 0), (16, 1024, 0), (32, 4096, 0), (64, 2048, 0), (64, 100000, 0),
                                            (64, 100000, 1), (64, 4096, 2)])
def test_scan_tiles_do_not_change_the_mesh(ctx, ppw, floor, pmin):
    xs = synth.uniform(5, 2500, 93)
    b, nrm = synth.cube_planes(5)
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    ref = ctx.build_mesh(precision="bf16")
    k0, c0 = ref.export()
    ref.free()
    m = ctx.build_mesh(precision="bf16", rays_per_block=ppw, packet_floor=floor, packet_min=pmin)
    k1, c1 = m.export()
    m.free()
    assert np.array_equal(k0, k1) and np.array_equal(c0, c1)


@pytest.mark.parametrize("d,n", [(3, 60000), (5, 2500)])
def test_descent_seed_changes_the_traversal_not_the_mesh(ctx, d, n):
    """hv_opts.seed draws the interior seeds' descent directions (the reference's random
    descent): every seed gives the same vertex set and coordinates."""
    xs = synth.uniform(d, n, 95)
    b, nrm = synth.cube_planes(d)
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    ref = ctx.build_mesh(precision="bf16")
    k0, c0 = ref.export()
    ref.free()
    for seed in (1, 12345, 2**63 + 7):
        m = ctx.build_mesh(precision="bf16", seed=seed)
        k1, c1 = m.export()
        m.free()
        assert np.array_equal(k0, k1) and np.array_equal(c0, c1)


def test_autotuner_picks_candidate_tiles_and_writes_its_cache(ctx, tmp_path):
    import json
    xs = synth.uniform(4, 4000, 94)
    b, nrm = synth.cube_planes(4)
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    path = tmp_path / "tune.jsonl"
    ppw, floor, ms = ctx.tune("bf16", str(path))
    assert ppw in (32, 64) and floor in (1024, 2048, 4096) and ms > 0
    rec = json.loads(path.read_text().splitlines()[-1])
    assert rec["d"] == 4 and rec["precision"] == 2 and rec["size_class"] == 11
    assert rec["rays_per_block"] == ppw and rec["packet_floor"] == floor
    # builds after tuning use the tuned tiles and give the same mesh
    k, c, _ = build(ctx, xs, "bf16")
    ok, _, _ = Oracle(xs, b, nrm).voronoi()
    assert np.array_equal(rows_view(k), rows_view(ok))
