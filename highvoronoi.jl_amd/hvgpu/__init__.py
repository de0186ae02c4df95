"""hvgpu -- host side of the MI355X raycast Voronoi path.

The compute lives in libhv.so (HIP kernels for gfx950 behind the C ABI declared in
include/hv.h); this package binds it with ctypes and mirrors the reference's
public API for the path: ``VoronoiGeometry``, ``VoronoiData``, ``cuboid``, ``boundary`` and the
integrator singletons (hvgpu.geometry), imported lazily so that the CPU-only parts
(synth) load without the HIP library.
"""
from . import synth  # noqa: F401

__all__ = ["synth", "VoronoiGeometry", "VoronoiData", "cuboid", "Boundary", "Plane", "boundary",
           "BC_Dirichlet", "BC_Neumann", "BC_Periodic",
           "VI_GEOMETRY", "VI_POLYGON", "VI_MONTECARLO", "VI_FAST_POLYGON"]


def __getattr__(name):
    if name in __all__ and name != "synth":
        from . import geometry
        return getattr(geometry, name)
    raise AttributeError(name)
