"""GPU parity of the mixed-precision scans (BF16 MFMA filter, FP32 VALU filter; each with
FP64 re-verification of its survivors).

Bar: a mixed-precision build returns exactly the vertex set of the FP64 build and of the
oracle (bit-exact keys) with the same coordinates, because every survivor of the filter is
decided by the FP64 rule; at the 5D N=100k mixture the check is the whole key table.
"""
import numpy as np
import pytest

from helpers import golden_cases, key_digest, key_set, load_golden, rows_view
from hvgpu import synth
from hvgpu._lib import HVError
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

COORD_TOL = 1e-10


@pytest.fixture(scope="module")
def ctx():
    from hvgpu._lib import Context
    return Context(0)


def build(ctx, xs, precision):
    b, nrm = synth.cube_planes(xs.shape[1])
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    m = ctx.build_mesh(precision=precision)
    keys, coords = m.export()
    st = m.stats()
    m.free()
    return keys, coords, st


MIXED = ["bf16", "fp32"]


@pytest.mark.parametrize("prec", MIXED)
@pytest.mark.parametrize("name", golden_cases())
def test_mixed_matches_exact_fixture(ctx, name, prec):
    g, xs, keys, exact = load_golden(name)
    k, c, st = build(ctx, xs, prec)
    assert key_set(k) == key_set(keys)


@pytest.mark.parametrize("prec", MIXED)
@pytest.mark.parametrize("d,n,seed", [(2, 1000, 0), (3, 3000, 5), (4, 2000, 1), (5, 1000, 9), (6, 300, 8)])
def test_mixed_matches_oracle_and_fp64(ctx, d, n, seed, prec):
    xs = synth.uniform(d, n, seed)
    k, c, st = build(ctx, xs, prec)
    k64, c64, _ = build(ctx, xs, "fp64")
    b, nrm = synth.cube_planes(d)
    ok, oc, _ = Oracle(xs, b, nrm).voronoi()
    assert np.array_equal(rows_view(k), rows_view(ok))
    assert np.array_equal(k, k64)
    # same FP64 circumcentre path: identical coordinates, and within the oracle's tolerance
    assert np.array_equal(c, c64)
    order = np.lexsort(ok.T[::-1])
    assert float(np.max(np.abs(c - oc[order]))) < COORD_TOL
    # the filter passes some candidates to FP64, never more than it scanned
    assert 0 < st["fp64_rechecks"] < st["pairs_scanned"]
    assert st["fallbacks"] <= st["rays"]


@pytest.mark.parametrize("prec", MIXED)
def test_mixed_identical_vertex_set_at_5d_100k_mixture(ctx, prec):
    # the benchmark workload: 5D N=100k Gaussian mixture, cuboid(5, periodic=[])
    xs = synth.gaussian_mixture_5d()
    k, c, st = build(ctx, xs, prec)
    k64, c64, st64 = build(ctx, xs, "fp64")
    assert st["vertices"] == st64["vertices"] > 10_000_000
    assert key_digest(k) == key_digest(k64)
    assert np.array_equal(k, k64)
    assert np.array_equal(c, c64)


def edge_closure_defects_hashed(keys, n):
    """Walkable edges (d-subsets of a key whose smallest id is a generator) not in exactly two
    keys, over a large table: every edge hashed to 64 bits id by id (a collision among ~1e8
    edges has probability ~1e-3 and could only merge two counts)."""
    keys = np.asarray(keys, dtype=np.uint64)
    V, m = keys.shape
    M1, M2 = np.uint64(0xBF58476D1CE4E5B9), np.uint64(0x94D049BB133111EB)
    hs = []
    with np.errstate(over="ignore"):
        for p in range(m):
            e = np.delete(keys, p, axis=1)
            e = e[e[:, 0] <= n]
            h = np.full(len(e), 0x9E3779B97F4A7C15, dtype=np.uint64)
            for j in range(m - 1):
                h = (h ^ e[:, j]) * M1
                h ^= h >> np.uint64(31)
                h *= M2
                h ^= h >> np.uint64(29)
            hs.append(h)
    h = np.sort(np.concatenate(hs))
    starts = np.flatnonzero(np.r_[True, h[1:] != h[:-1]])
    counts = np.diff(np.r_[starts, len(h)])
    return int(np.sum(counts != 2))


def test_5d_100k_mixture_cells_equal_oracle(ctx):
    """The benchmark workload against the oracle: the vertices of cells 1..2000 (the oracle's
    walk over Iter = 1:2000) equal the FP64 build's vertices that hold one of them and the
    BF16 walk of those cells alone (hv_build_mesh_cells); the whole FP64 table closes every
    edge (each walkable edge in exactly two keys)."""
    xs = synth.gaussian_mixture_5d()
    n, b = len(xs), 2000
    lo, nrm = synth.cube_planes(5)
    ok, oc, _ = Oracle(xs, lo, nrm).voronoi(cells=(0, b))
    k64, c64, st = build(ctx, xs, "fp64")
    assert st["vertices"] > 10_000_000
    held = (k64 <= b).any(axis=1)
    assert np.array_equal(rows_view(k64[held]), rows_view(ok))
    order = np.lexsort(ok.T[::-1])
    assert float(np.max(np.abs(c64[held] - oc[order]))) < COORD_TOL
    assert edge_closure_defects_hashed(k64, n) == 0
    del k64, c64
    m = ctx.build_mesh_cells(np.arange(1, b + 1), precision="bf16")
    kc, cc = m.export()
    m.free()
    assert np.array_equal(rows_view(kc), rows_view(ok))
    assert float(np.max(np.abs(cc - oc[order]))) < COORD_TOL


@pytest.mark.parametrize("prec", MIXED)
def test_mixed_rejects_other_leaf_sizes(ctx, prec):
    xs = synth.uniform(3, 200, 1)
    b, nrm = synth.cube_planes(3)
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    with pytest.raises(HVError) as e:
        ctx.build_mesh(precision=prec, leaf_size=16)
    assert e.value.code == -7


@pytest.mark.parametrize("prec", MIXED)
def test_mixed_raycast_batch_matches_fp64(ctx, prec):
    d, n = 5, 800
    xs = synth.uniform(d, n, 31)
    b, nrm = synth.cube_planes(d)
    k64, c64, _ = build(ctx, xs, "fp64")
    rng = np.random.default_rng(3)
    rows = rng.permutation(len(k64))[:500]
    sig, drop, r = [], [], []
    for i in rows:
        for p in range(d + 1):
            e = np.delete(k64[i], p)
            if e[0] > n:
                continue
            sig.append(k64[i]); drop.append(p); r.append(c64[i])
    sig, drop, r = np.array(sig), np.array(drop), np.array(r)
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    a = ctx.raycast_batch(sig, drop, r, precision="fp64")
    m = ctx.raycast_batch(sig, drop, r, precision=prec)
    for x, y in zip(a, m):
        assert np.array_equal(x, y)
