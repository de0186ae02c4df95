"""CPU tests of the C-ABI library: it is built, it loads, and it exports every entry
point include/hv.h declares (no compute call: there is no GPU in this container)."""
import ctypes

import numpy as np
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "highvoronoi.jl_amd", "libhv.so")


def declared_symbols():
    with open(os.path.join(ROOT, "include", "hv.h")) as f:
        src = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(hv_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_the_path():
    syms = declared_symbols()
    for s in ["hv_create", "hv_destroy", "hv_last_error", "hv_set_points", "hv_set_points_device",
              "hv_set_boundary", "hv_build_mesh", "hv_mesh_counts", "hv_mesh_export", "hv_mesh_stats",
              "hv_mesh_free", "hv_raycast_batch", "hv_nn_halfspace_batch", "hv_inrange_batch"]:
        assert s in syms


def test_library_loads_and_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libhv.so not built (run __graft_entry__.build())"
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_resolves_all_symbols():
    import hvgpu._lib as L
    lib = L.lib()
    for s in declared_symbols():
        assert getattr(lib, s) is not None


def test_library_is_built_for_gfx950():
    with open(LIB, "rb") as f:
        blob = f.read()
    assert b"gfx950" in blob


def test_cuboid_mirrors_reference_plane_order():
    from hvgpu.geometry import cuboid
    from hvgpu import synth
    b = cuboid(4, periodic=[])
    base, normal, bc = b.arrays()
    want_b, want_n = synth.cube_planes(4)
    assert np.array_equal(base, want_b) and np.array_equal(normal, want_n)
    assert np.all(bc == 0)
    # the reference's default is periodic in every direction, partners 2i <-> 2i-1
    bp = cuboid(3).arrays()[2]
    assert list(bp) == [2, 1, 4, 3, 6, 5]
    bn = cuboid(2, periodic=[], neumann=[1, -2]).arrays()[2]
    assert list(bn) == [-1, 0, 0, -1]
