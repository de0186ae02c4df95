"""GPU parity of integrand quadrature (hv_integrate_polytope_f): bulk and interface integrals
of f by the reference's polytope rule (src/polyintegrator.jl:174-353), against the CPU
restatement oracle_py.polytope_integrand on the same mesh (relative 1e-10), with f = 1
giving the volumes and areas of hv_integrate_polytope."""
import numpy as np
import pytest

from hvgpu import synth
from hvgpu.geometry import cuboid
from hvgpu._lib import Context
from oracle_py import polytope_integrand

pytestmark = pytest.mark.gpu

TOL = 1e-10


def integrand(p):
    d = p.shape[1]
    return np.stack([np.ones(len(p)), p[:, 0], p[:, 1] ** 2, np.sin(3 * p[:, 0]) * np.cos(p[:, d - 1])], axis=1)


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def check(I, bulk, iface, n):
    assert np.max(np.abs(I["bulk"] - bulk)) <= TOL * max(1.0, np.max(np.abs(bulk)))
    ptr, nbr = I["nbr_ptr"], I["neighbors"]
    for i in range(n):
        for t in range(ptr[i], ptr[i + 1]):
            ref = iface[(i, int(nbr[t]))]
            assert np.max(np.abs(I["interface"][t] - ref)) <= TOL * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("d,n,seed", [(2, 80, 1), (3, 50, 2), (4, 25, 3), (5, 14, 4)])
def test_integrand_matches_restatement(ctx, d, n, seed):
    xs = synth.uniform(d, n, seed)
    base, normal = synth.cube_planes(d)
    ctx.set_points(xs)
    ctx.set_boundary(base, normal)
    m = ctx.build_mesh(precision="bf16")
    keys, coords = m.export()
    I = ctx.integrate(m, "polytope", integrand=integrand, nf=4)
    P = ctx.integrate(m, "polytope")
    m.free()
    bulk, iface = polytope_integrand(xs, base, normal, keys, coords, integrand, 4)
    check(I, bulk, iface, n)
    assert np.allclose(I["bulk"][:, 0], P["volumes"], rtol=1e-12, atol=0)
    assert np.allclose(I["interface"][:, 0], P["areas"], rtol=1e-12, atol=1e-300)
    assert np.array_equal(I["neighbors"], P["neighbors"]) and np.array_equal(I["volumes"], P["volumes"])


@pytest.mark.parametrize("d,n,seed,per", [(2, 60, 5, (1, 2)), (3, 40, 6, (1, 2, 3)), (3, 40, 7, (2,))])
def test_integrand_on_periodic_mesh(ctx, d, n, seed, per):
    xs = synth.uniform(d, n, seed)
    base, normal, bc = cuboid(d, periodic=list(per)).arrays()
    ctx.set_points(xs)
    ctx.set_boundary(base, normal, bc)
    m = ctx.build_mesh(precision="fp64")
    keys, coords = m.export()
    orig, shift = m.images()
    I = ctx.integrate(m, "polytope", integrand=integrand, nf=4)
    m.free()
    ext = xs[orig - 1] + shift   # unit box: one length per shift
    bulk, iface = polytope_integrand(ext, base, normal, keys, coords, integrand, 4, ncell=n)
    check(I, bulk, iface, n)
    assert abs(I["bulk"][:, 0].sum() - 1.0) < 1e-12


def test_integrand_error_reaches_the_caller(ctx):
    xs = synth.uniform(3, 100, 9)
    base, normal = synth.cube_planes(3)
    ctx.set_points(xs)
    ctx.set_boundary(base, normal)
    m = ctx.build_mesh(precision="fp64")

    def bad(p):
        raise ValueError("integrand failed")

    with pytest.raises(ValueError):
        ctx.integrate(m, "polytope", integrand=bad, nf=1)
    # the context stays usable
    I = ctx.integrate(m, "polytope", integrand=lambda p: np.ones((len(p), 1)), nf=1)
    assert abs(I["bulk"][:, 0].sum() - 1.0) < 1e-12
    m.free()


def test_voronoi_geometry_integrand(ctx):
    from hvgpu import VI_FAST_POLYGON, VI_POLYGON, VoronoiData, VoronoiGeometry
    xs = synth.uniform(3, 200, 12)
    f = lambda p: np.stack([p[:, 0], p[:, 2] ** 2], axis=1)
    a = VoronoiData(VoronoiGeometry(xs, cuboid(3, periodic=[]), integrator=VI_POLYGON, integrand=f))
    b = VoronoiData(VoronoiGeometry(xs, cuboid(3, periodic=[]), integrator=VI_FAST_POLYGON, integrand=f))
    assert np.array_equal(a.bulk_integral, b.bulk_integral)
    assert abs(a.bulk_integral[:, 0].sum() - 0.5) < 1e-12   # affine: exact in 3D
    for i in (0, 99, 199):
        assert a.interface_integral[i].shape == (len(a.neighbors[i]), 2)


def test_6d_50k_face_passes_cover_every_face(ctx):
    # 6D N=50k: millions of facets and ridges, one wavefront each, walked grid-stride (a
    # dispatch holds at most 2^32 work-items); every face must be measured: the volumes sum to
    # the box, f = 1 gives the volumes, x0 integrates to ~1/2
    xs = synth.uniform(6, 50_000, 4)
    base, normal = synth.cube_planes(6)
    ctx.set_points(xs)
    ctx.set_boundary(base, normal)
    m = ctx.build_mesh(precision="bf16")
    I = ctx.integrate(m, "polytope", integrand=lambda p: np.stack([np.ones(len(p)), p[:, 0]], axis=1), nf=2)
    m.free()
    assert abs(I["volumes"].sum() - 1.0) < 1e-10
    assert np.allclose(I["bulk"][:, 0], I["volumes"], rtol=1e-10, atol=0)
    assert abs(I["bulk"][:, 1].sum() - 0.5) < 1e-3
