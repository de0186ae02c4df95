"""User-facing mirror of the reference's API for this path, on top of libhv.so.

    VG = VoronoiGeometry(xs, cuboid(5, periodic=[]), integrator=VI_POLYGON,
                         search_settings={"precision": "bf16"})
    data = VoronoiData(VG)
    data.volume[i], data.neighbors[i], data.area[i], data.vertices[i]

follows VoronoiGeometry(xs, boundary; integrator=..., search_settings=(...))
(src/geometry.jl:136-260) and VoronoiData(VG) (src/voronoidata.jl:545-621) with the same
field meanings: neighbours sorted, 1-based, boundary plane p as n + p; area[i][k] belongs to
neighbors[i][k]; vertices[i] maps each vertex key of cell i (sorted 1-based tuple) to its
coordinates.  The mesh and the integrals are computed by the device path only; there is
no CPU fallback.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ._lib import Context

# integrator singletons (src/HighVoronoi.jl:102-117)
VI_GEOMETRY = "geometry"
VI_POLYGON = "polygon"
VI_MONTECARLO = "montecarlo"
VI_FAST_POLYGON = "fast_polygon"

DIRICHLET = 0
NEUMANN = -1

# raycast methods (src/raycast-types.jl:245-310).  Every method computes the same tessellation
# (test/rcmethods.jl checks exactly that through the volume integral); on the device they are
# one walk, the FP64 decision rule of Raycast_Non_General_HP behind the filter the precision
# names, so the method is accepted and recorded, and changes nothing.
RCOriginal = "RCOriginal"
RCOriginalSafety = "RCOriginalSafety"
RCOriginalHP = "RCOriginalHP"
RCNonGeneralFast = "RCNonGeneralFast"
RCNonGeneralSkip = "RCNonGeneralSkip"
RCCombined = "RCCombined"
RCNonGeneralHP = "RCNonGeneralHP"
RCNonGeneralCutoff = "RCNonGeneralCutoff"
RCNonGeneral = RCNonGeneralHP
RCStandard = RCNonGeneral
METHODS = (RCOriginal, RCOriginalSafety, RCOriginalHP, RCNonGeneralFast, RCNonGeneralSkip, RCCombined,
           RCNonGeneralHP, RCNonGeneralCutoff)


@dataclass
class Plane:
    """A boundary plane (src/boundary.jl:22-29): base point, outer normal, and BC
    (0 Dirichlet, -1 Neumann, > 0 the 1-based index of the periodic partner plane)."""
    base: np.ndarray
    normal: np.ndarray
    BC: int = DIRICHLET


@dataclass
class Boundary:
    planes: list = field(default_factory=list)

    @property
    def dimension(self):
        return len(self.planes[0].base) if self.planes else 0

    def arrays(self):
        base = np.array([p.base for p in self.planes], dtype=np.float64)
        normal = np.array([p.normal for p in self.planes], dtype=np.float64)
        bc = np.array([p.BC for p in self.planes], dtype=np.int16)
        return base, normal, bc


def BC_Dirichlet(base, normal) -> Plane:
    """A Dirichlet plane with base point `base` and outer unit normal `normal`
    (src/boundary.jl:37-39)."""
    return Plane(np.asarray(base, dtype=np.float64), np.asarray(normal, dtype=np.float64), DIRICHLET)


def BC_Neumann(base, normal) -> Plane:
    """A Neumann plane (src/boundary.jl:41-43)."""
    return Plane(np.asarray(base, dtype=np.float64), np.asarray(normal, dtype=np.float64), NEUMANN)


@dataclass
class BC_Periodic:
    """Two periodic planes: `base1` with outer normal `normal1`, `base2` with -normal1
    (src/boundary.jl:45-54)."""
    base1: np.ndarray
    base2: np.ndarray
    normal1: np.ndarray


def boundary(*planes) -> Boundary:
    """Boundary(planes...) (src/boundary.jl:109-131): planes from BC_Dirichlet / BC_Neumann, and
    BC_Periodic pairs whose two planes name each other (1-based) as partners.  Normals must have
    unit length (the mirror image x + 2<n, b - x> n assumes it).  The device walk takes the planes
    of cuboid(d; ...) or any planes of Dirichlet / Neumann type that bound a convex polytope."""
    out = []
    for p in planes:
        if isinstance(p, BC_Periodic):
            k = len(out) + 1
            n1 = np.asarray(p.normal1, dtype=np.float64)
            out.append(Plane(np.asarray(p.base1, dtype=np.float64), n1, k + 1))
            out.append(Plane(np.asarray(p.base2, dtype=np.float64), -n1, k))
        else:
            out.append(p)
    return Boundary(out)


def cuboid(dim: int, dimensions=None, periodic=None, neumann=(), offset=None) -> Boundary:
    """cuboid(dim; dimensions=ones, periodic=1:dim, neumann=[], offset=zeros)
    (src/boundary.jl:510-534).  As in the reference the default is periodic in every
    direction; pass periodic=[] for the Dirichlet box."""
    dims = np.ones(dim) if dimensions is None else np.asarray(dimensions, dtype=np.float64)
    off = np.zeros(dim) if offset is None else np.asarray(offset, dtype=np.float64)
    per = set(range(1, dim + 1)) if periodic is None else set(periodic)
    planes = [None] * (2 * dim)
    for i in range(1, dim + 1):
        unit = np.zeros(dim)
        unit[i - 1] = 1.0
        if i in per:
            planes[2 * i - 2] = Plane(off + dims[i - 1] * unit, unit.copy(), 2 * i)
            planes[2 * i - 1] = Plane(off.copy(), -unit, 2 * i - 1)
            continue
        planes[2 * i - 2] = Plane(off + dims[i - 1] * unit, unit.copy(), NEUMANN if i in neumann else DIRICHLET)
        planes[2 * i - 1] = Plane(off.copy(), -unit, NEUMANN if -i in neumann else DIRICHLET)
    return Boundary(planes)


def _cell_vertex_csr(keys: np.ndarray, n: int):
    nv, w = keys.shape
    rows = np.repeat(np.arange(nv, dtype=np.int64), w)
    ids = keys.reshape(-1)
    real = (ids >= 1) & (ids <= n)
    cell = ids[real] - 1
    order = np.argsort(cell, kind="stable")
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(ptr, cell + 1, 1)
    return np.cumsum(ptr), rows[real][order]


class VoronoiGeometry:
    """Mesh (and, with an integrator, volumes and interface areas) of the generators xs
    in `boundary`, computed on one MI355X.

    search_settings: {"precision": "bf16" | "fp32" | "fp64", "leaf_size": int, "device": int,
    "method": one of METHODS (RaycastParameter.method; all give the same mesh)}.
    mc_accurate: (directions per cell, unused, unused) as the reference's tuple
    (src/geometry.jl:139); seed drives the Monte-Carlo directions.
    integrand (VI_POLYGON / VI_FAST_POLYGON): f(points[npts, d]) -> values[npts, k], the
    reference's integrand=f (src/geometry.jl:136-160): bulk and interface integrals by the
    polytope rule.  VI_FAST_POLYGON computes the same quantities as VI_POLYGON (the reference's
    two exact integrators share the centre rule, src/fastpolyintegrator.jl:272-476) and runs the
    same device integrator.
    """

    def __init__(self, xs, boundary: Boundary | None = None, integrator=VI_GEOMETRY, search_settings=None,
                 mc_accurate=(1000, 100, 20), seed: int = 0, integrand=None):
        self.nodes = np.ascontiguousarray(np.asarray(xs, dtype=np.float64))
        n, d = self.nodes.shape
        self.boundary = boundary if boundary is not None else Boundary([])
        s = dict(search_settings or {})
        self.precision = s.get("precision", "bf16")
        self.method = s.get("method", RCStandard)
        if self.method not in METHODS:
            raise ValueError(f"unknown raycast method {self.method!r} (one of {', '.join(METHODS)})")
        self.integrator = integrator
        if integrand is not None and integrator not in (VI_POLYGON, VI_FAST_POLYGON):
            raise NotImplementedError("integrands are integrated by VI_POLYGON / VI_FAST_POLYGON on the device path")
        self._ctx = Context(int(s.get("device", 0)))
        self._ctx.set_points(self.nodes)
        base, normal, bc = self.boundary.arrays()
        self._ctx.set_boundary(base.reshape(-1, d), normal.reshape(-1, d), bc if len(bc) else None)
        self.leaf_size = int(s.get("leaf_size", 0))
        self.integrand, self.mc_accurate, self.seed = integrand, mc_accurate, seed
        self._mesh = self._ctx.build_mesh(precision=self.precision, leaf_size=self.leaf_size)
        self.stats = self._mesh.stats()
        self._integrate()

    def _integrate(self):
        self.integral = None
        if self.integrator in (VI_POLYGON, VI_FAST_POLYGON):
            if self.integrand is None:
                self.integral = self._ctx.integrate(self._mesh, "polytope")
            else:
                k = np.asarray(self.integrand(self.nodes[:1]), dtype=np.float64).reshape(1, -1).shape[1]
                self.integral = self._ctx.integrate(self._mesh, "polytope", integrand=self.integrand, nf=k)
        elif self.integrator == VI_MONTECARLO:
            self.integral = self._ctx.integrate(self._mesh, "montecarlo", ndir=int(self.mc_accurate[0]),
                                                seed=self.seed)
        self._keys = None
        self._coords = None

    def refine(self, xs):
        """refine!(VG, xs) (src/geometry.jl:404-438): appends the points xs[m, d] to the nodes
        (ids n+1..n+m), refines the mesh on the device (hv_refine_mesh: the old vertices whose
        circumsphere holds no new point, plus a walk of the new cells) and integrates again
        with the same integrator; in a periodic box the mesh of the union is built afresh (the new
        points bring their own images).  Returns self."""
        xs = np.ascontiguousarray(np.asarray(xs, dtype=np.float64)).reshape(-1, self.nodes.shape[1])
        mesh, self.removed = self._ctx.refine(self._mesh, xs, precision=self.precision, leaf_size=self.leaf_size)
        self._mesh.free()
        self._mesh = mesh
        self.nodes = np.vstack([self.nodes, xs])
        self.stats = self._mesh.stats()
        self._integrate()
        return self

    def vertex_table(self):
        """(keys[V, d+1] sorted 1-based, coords[V, d]) in lexicographic key order."""
        if self._keys is None:
            self._keys, self._coords = self._mesh.export()
        return self._keys, self._coords

    def __del__(self):
        m = getattr(self, "_mesh", None)
        if m is not None:
            m.free()


class _CellVertices:
    def __init__(self, keys, coords, ptr, rows):
        self._k, self._c, self._p, self._r = keys, coords, ptr, rows

    def __len__(self):
        return len(self._p) - 1

    def __getitem__(self, i):
        rows = self._r[self._p[i]:self._p[i + 1]]
        return {tuple(int(v) for v in self._k[r]): self._c[r] for r in rows}


class VoronoiData:
    """Read-only views of a VoronoiGeometry (src/voronoidata.jl:545-621): nodes, vertices,
    neighbors, volume, area.  volume / neighbors / area need an integrator.

    Periodic boxes: ids are public as in the reference (VoronoiDataShift,
    src/voronoidata.jl:540-542): a periodic image is named by its original's index and
    plane p is n + p; vertices[i] holds cell i's vertices in its own frame.  `images` keeps the
    device labels: (orig, shift) per device generator id (identity without periodic planes)."""

    def __init__(self, VG: VoronoiGeometry):
        self.geometry = VG
        self.nodes = VG.nodes
        self.boundary = VG.boundary
        n = len(VG.nodes)
        keys, coords = VG.vertex_table()
        orig, shift = VG._mesh.images()
        self.images = (orig, shift)
        n_ext = len(orig)

        def public(ids):
            ids = np.asarray(ids, dtype=np.int64)
            out = ids - n_ext + n
            inner = ids <= n_ext
            out[inner] = orig[ids[inner] - 1]
            return out

        ptr, rows = _cell_vertex_csr(keys, n)
        pkeys = np.sort(public(keys), axis=1) if n_ext > n else keys
        self.vertices = _CellVertices(pkeys, coords, ptr, rows)
        I = VG.integral
        self.bulk_integral = self.interface_integral = None
        if I is None:
            self.volume = self.neighbors = self.area = None
            return
        self.volume = I["volumes"]
        p = I["nbr_ptr"]
        nb = public(I["neighbors"]) if n_ext > n else I["neighbors"]
        self.neighbors = [nb[p[i]:p[i + 1]] for i in range(n)]
        self.area = [I["areas"][p[i]:p[i + 1]] for i in range(n)]
        if "bulk" in I:   # integrand=f: per cell, and per interface aligned with neighbors[i]
            self.bulk_integral = I["bulk"]
            self.interface_integral = [I["interface"][p[i]:p[i + 1]] for i in range(n)]
