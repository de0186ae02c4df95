"""The reference's method parity check (test/rcmethods.jl:3-13): 4D N=1000 uniform in
cuboid(4, periodic=[]), VI_FAST_POLYGON with integrand x -> [1.0]; for every raycast method the
summed bulk integral is 1 within 1e-3.  Here the methods share one device walk, so beyond the
reference's bar every method gives the same vertex table and the same integrals bit for bit.
"""
import numpy as np
import pytest

from hvgpu import synth
from hvgpu.geometry import (METHODS, RCCombined, RCNonGeneral, RCNonGeneralFast, RCOriginal, VI_FAST_POLYGON,
                            VoronoiData, VoronoiGeometry, cuboid)

pytestmark = pytest.mark.gpu


def run(xs, method):
    VG = VoronoiGeometry(xs, cuboid(4, periodic=[]), integrator=VI_FAST_POLYGON, integrand=lambda x: np.ones((len(x), 1)),
                         search_settings={"method": method})
    return VG, VoronoiData(VG)


def test_raycast_methods_give_the_same_tessellation():
    xs = synth.uniform(4, 1000, 7)
    ref_keys = ref_bulk = None
    for method in (RCCombined, RCOriginal, RCNonGeneralFast, RCNonGeneral):
        VG, vd = run(xs, method)
        v = float(np.sum(VG.integral["bulk"][:, 0]))
        assert abs(v - 1.0) < 0.001
        keys, _ = VG.vertex_table()
        if ref_keys is None:
            ref_keys, ref_bulk = keys, VG.integral["bulk"]
        assert np.array_equal(keys, ref_keys)
        assert np.array_equal(VG.integral["bulk"], ref_bulk)


def test_unknown_method_is_refused():
    with pytest.raises(ValueError):
        run(synth.uniform(4, 100, 8), "RCSomethingElse")
    assert len(METHODS) == 8
