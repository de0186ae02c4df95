"""Parity at large sizes: 3D N=5M and 6D N=50k uniform in the Dirichlet cube.

Bars: the BF16-filter build and the FP64 build give the same key table and coordinates bit
for bit over the whole mesh (33.7M and 39.1M vertices); the vertices of the first cells equal
the oracle's walk over those cells (keys bit-exact, coordinates within 1e-10); every walkable
edge of the FP64 table lies in exactly two keys.
"""
import numpy as np
import pytest

from helpers import rows_view
from hvgpu import synth
from hvgpu._lib import Context
from oracle_py import Oracle
from test_gpu_mixed import edge_closure_defects_hashed

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


@pytest.mark.parametrize("d,n,seed,b", [(3, 5_000_000, 41, 2000), (6, 50_000, 42, 25)])
def test_large_mesh_mixed_equals_fp64_and_oracle(ctx, d, n, seed, b):
    xs = synth.uniform(d, n, seed)
    lo, nrm = synth.cube_planes(d)
    ctx.set_points(xs)
    ctx.set_boundary(lo, nrm)
    m = ctx.build_mesh(precision="fp64")
    k64, c64 = m.export()
    m.free()
    assert len(k64) > 10_000_000
    assert edge_closure_defects_hashed(k64, n) == 0
    m = ctx.build_mesh(precision="bf16")
    k16, c16 = m.export()
    m.free()
    assert np.array_equal(k16, k64) and np.array_equal(c16, c64)
    del k16, c16
    ok, oc, _ = Oracle(xs, lo, nrm).voronoi(cells=(0, b))
    held = (k64 <= b).any(axis=1)
    assert np.array_equal(rows_view(k64[held]), rows_view(ok))
    order = np.lexsort(ok.T[::-1])
    assert float(np.max(np.abs(c64[held] - oc[order]))) < 1e-10
