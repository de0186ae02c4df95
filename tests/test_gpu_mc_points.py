"""GPU checks of the point-sample Monte-Carlo integrator (hv_integrate_mc_points).

The estimator is statistical, so it is pinned against the exact polytope integrator on the
same mesh (itself pinned against Qhull and the oracle): the volumes sum to the box volume
exactly (every sample lands in one cell), each cell volume is within 6 binomial standard
deviations, and the interface areas carry no first-order slab bias (total within 1.5%,
mean absolute error within 6%).  Same seed, same result.
"""
import numpy as np
import pytest

from hvgpu import synth
from hvgpu._lib import Context, HVError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def mesh_and_exact(ctx, d, n, seed, lo=0.0, hi=1.0):
    xs = lo + (hi - lo) * synth.uniform(d, n, seed)
    b, nrm = synth.cube_planes(d, lo, hi)
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    m = ctx.build_mesh(precision="bf16")
    return m, ctx.integrate(m, "polytope")


@pytest.mark.parametrize("d,n", [(2, 200), (3, 300), (4, 250)])
def test_points_mc_against_exact(ctx, d, n):
    m, ex = mesh_and_exact(ctx, d, n, 60 + d)
    ns = 20_000_000
    mc = ctx.integrate(m, "points", nsamples=ns, seed=7, slab=0.02)
    m.free()
    assert np.array_equal(mc["neighbors"], ex["neighbors"])
    assert np.array_equal(mc["nbr_ptr"], ex["nbr_ptr"])
    assert abs(mc["volumes"].sum() - 1.0) < 1e-12
    p = ex["volumes"]
    sigma = np.sqrt(p * (1.0 - p) / ns)
    assert float(np.max(np.abs(mc["volumes"] - p) / sigma)) < 6.0
    A, Am = ex["areas"], mc["areas"]
    assert np.all(Am >= 0.0)
    assert abs(Am.sum() - A.sum()) / A.sum() < 0.015
    assert np.sum(np.abs(Am - A)) / A.sum() < 0.06


def test_points_mc_box_volume_and_determinism(ctx):
    m, ex = mesh_and_exact(ctx, 3, 150, 9, lo=-1.0, hi=2.0)
    a = ctx.integrate(m, "points", nsamples=3_000_000, seed=11)
    b = ctx.integrate(m, "points", nsamples=3_000_000, seed=11)
    c = ctx.integrate(m, "points", nsamples=3_000_000, seed=12)
    m.free()
    assert abs(a["volumes"].sum() - 27.0) < 27.0 * 1e-12
    assert np.array_equal(a["volumes"], b["volumes"]) and np.array_equal(a["areas"], b["areas"])
    assert not np.array_equal(a["volumes"], c["volumes"])


def test_points_mc_rejects_bad_sample_count(ctx):
    m, _ = mesh_and_exact(ctx, 2, 50, 3)
    with pytest.raises(HVError) as e:
        ctx.integrate(m, "points", nsamples=0)
    m.free()
    assert e.value.code == -1
