"""hvgpu -- host side of the MI355X raycast Voronoi path.

The compute lives in libhv.so (HIP kernels for gfx950 behind the C ABI declared in
include/hv.h); this package binds it with ctypes and mirrors the reference's
public API for the path: ``VoronoiGeometry``, ``VoronoiData``, ``cuboid``,
``RaycastParameter`` and the integrator singletons.
"""
from . import synth  # noqa: F401
