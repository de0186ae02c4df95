"""hv_raycast_batch_noalloc (the allocation-free variant of hv_raycast_batch, for callers that
cannot allocate device memory) against hv_raycast_batch through the C ABI: the same rays give
the same keys, coordinates, ray parameters and statuses, bit for bit, on two calls in a row (the
second with fewer rays), in each precision.  Device buffers for the inputs and outputs are the
caller's (allocated here with hipMalloc, outside the entry point)."""
import ctypes as C

import numpy as np
import pytest

from hvgpu import synth
from hvgpu._lib import Context, HVError, PRECISIONS, lib
from oracle_py import Oracle

pytestmark = pytest.mark.gpu


def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, h, nbytes):
        self.h = h
        self.p = C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), max(8, nbytes)) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert self.h.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape, dtype):
        out = np.zeros(shape, dtype=dtype)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def rays_of(xs, base, normal, count, seed):
    """(sig, drop, r) of `count` walkable edges of the oracle's mesh."""
    ok, oc, _ = Oracle(xs, base, normal).voronoi()
    n, d = xs.shape
    rng = np.random.default_rng(seed)
    sig, drop, r = [], [], []
    for i in rng.permutation(len(ok)):
        for p in range(d + 1):
            if np.delete(ok[i], p)[0] > n:
                continue
            sig.append(ok[i]); drop.append(p); r.append(oc[i])
        if len(sig) >= count:
            break
    return (np.array(sig[:count], dtype=np.int64), np.array(drop[:count], dtype=np.int32),
            np.array(r[:count], dtype=np.float64))


def noalloc(ctx, h, precision, sig, drop, r):
    n, d1 = sig.shape
    d = d1 - 1
    bufs = [DevBuf(h, sig.nbytes), DevBuf(h, drop.nbytes), DevBuf(h, r.nbytes),
            DevBuf(h, sig.nbytes), DevBuf(h, r.nbytes), DevBuf(h, 8 * n), DevBuf(h, 4 * n)]
    try:
        bufs[0].put(sig); bufs[1].put(drop); bufs[2].put(r)
        rc = lib().hv_raycast_batch_noalloc(ctx._h, PRECISIONS[precision], n, *[b.p for b in bufs])
        ctx._check(rc)
        return (bufs[3].get((n, d + 1), np.int64), bufs[4].get((n, d), np.float64),
                bufs[5].get(n, np.float64), bufs[6].get(n, np.int32))
    finally:
        for b in bufs:
            b.free()


@pytest.mark.parametrize("precision", ["fp64", "bf16", "fp32"])
def test_noalloc_equals_raycast_batch_on_two_calls(precision):
    h = hip()
    d, n = 4, 3000
    xs = synth.uniform(d, n, 123)
    base, normal = synth.cube_planes(d)
    sig, drop, r = rays_of(xs, base, normal, 5000, 4)
    ctx = Context(0)
    ctx.set_points(xs)
    ctx.set_boundary(base, normal)
    for count in (5000, 1700):          # the second call with fewer rays
        s, dp, rr = sig[:count], drop[:count], r[:count]
        want = ctx.raycast_batch(s, dp, rr, precision=precision)   # builds the tree of this precision
        got = noalloc(ctx, h, precision, s, dp, rr)
        for w, g in zip(want, got):
            assert np.array_equal(w, g)
        assert np.all(got[3] == 0)


def test_noalloc_refusals():
    h = hip()
    xs = synth.uniform(3, 500, 7)
    base, normal = synth.cube_planes(3)
    ctx = Context(0)
    ctx.set_points(xs)
    ctx.set_boundary(base, normal)
    sig, drop, r = rays_of(xs, base, normal, 10, 1)
    # no tree yet on this context's points
    with pytest.raises(HVError) as e:
        noalloc(ctx, h, "fp64", sig, drop, r)
    assert e.value.code == -8
    ctx.raycast_batch(sig, drop, r, precision="fp64")
    with pytest.raises(HVError) as e:        # FP64 tables only: no BF16 tables yet
        noalloc(ctx, h, "bf16", sig, drop, r)
    assert e.value.code == -8
    bad = sig.copy()
    bad[0, [0, 1]] = bad[0, [1, 0]]          # unsorted key: refused before any ray is cast
    with pytest.raises(HVError) as e:
        noalloc(ctx, h, "fp64", bad, drop, r)
    assert e.value.code == -1
    # more rays than the scratch holds: refused before anything touches the buffers
    buf = DevBuf(h, 64)
    try:
        rc = lib().hv_raycast_batch_noalloc(ctx._h, 0, 40000, *([buf.p] * 7))
        assert rc == -6
    finally:
        buf.free()
    got = noalloc(ctx, h, "fp64", sig, drop, r)    # still usable after the refusals
    want = ctx.raycast_batch(sig, drop, r, precision="fp64")
    assert all(np.array_equal(a, b) for a, b in zip(want, got))
