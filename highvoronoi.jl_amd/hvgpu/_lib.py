"""ctypes binding of libhv.so (C ABI in include/hv.h).

The shared library is the product: there is no CPU fallback.  Importing this module
without a built libhv.so raises, and so does any call when no GPU is visible.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_PKG), "libhv.so")

HV_OK = 0
ERRORS = {
    -1: "HV_ERR_ARG", -2: "HV_ERR_DOMAIN", -3: "HV_ERR_FEW_POINTS", -4: "HV_ERR_DEGENERATE",
    -5: "HV_ERR_DEVICE", -6: "HV_ERR_CAPACITY", -7: "HV_ERR_UNSUPPORTED", -8: "HV_ERR_STATE",
}
PRECISIONS = {"fp64": 0, "fp32": 1, "bf16": 2}


class HVError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class hv_opts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("leaf_size", C.c_int32), ("rays_per_block", C.c_int32),
                ("reserved0", C.c_int32), ("plane_tolerance", C.c_double), ("ray_tol", C.c_double),
                ("vertex_capacity", C.c_int64), ("seed", C.c_uint64), ("packet_floor", C.c_int32),
                ("packet_min", C.c_int32)]


class hv_stats_t(C.Structure):
    _fields_ = [("rays", C.c_int64), ("levels", C.c_int64), ("vertices", C.c_int64), ("duplicates", C.c_int64),
                ("pairs_scanned", C.c_int64), ("leaves_visited", C.c_int64), ("near_ties", C.c_int64),
                ("degenerate", C.c_int64), ("fallbacks", C.c_int64), ("fp64_rechecks", C.c_int64),
                ("retries", C.c_int64), ("unbounded", C.c_int64), ("max_t_relerr", C.c_double),
                ("ms_total", C.c_double), ("ms_raycast", C.c_double), ("rays_skipped", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


# void f(void* user, int64_t npts, const double* pts /* d x npts */, double* values /* nf x npts */)
INTEGRAND = C.CFUNCTYPE(None, C.c_void_p, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc for gfx950)")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    L.hv_create.argtypes = [P(vp), C.c_int]
    L.hv_destroy.argtypes = [vp]
    L.hv_last_error.argtypes = [vp]
    L.hv_last_error.restype = C.c_char_p
    L.hv_set_points.argtypes = [vp, C.c_int, C.c_int64, P(C.c_double)]
    L.hv_set_points_device.argtypes = [vp, C.c_int, C.c_int64, vp]
    L.hv_set_boundary.argtypes = [vp, C.c_int, P(C.c_double), P(C.c_double), P(C.c_int16)]
    L.hv_build_mesh.argtypes = [vp, P(hv_opts), P(vp)]
    L.hv_build_mesh_cells.argtypes = [vp, P(hv_opts), C.c_int64, P(C.c_int64), P(vp)]
    L.hv_tune.argtypes = [vp, C.c_int, C.c_char_p, P(C.c_int32), P(C.c_int32), P(C.c_double)]
    L.hv_batch_release.argtypes = [C.c_int]
    L.hv_batch_release.restype = C.c_int
    L.hv_build_mesh_batch.argtypes = [C.c_int, C.c_int32, P(C.c_int32), P(C.c_int64), P(P(C.c_double)),
                                      P(C.c_int32), P(P(C.c_double)), P(P(C.c_double)), P(P(C.c_int16)),
                                      P(hv_opts), C.c_int32, P(vp), P(C.c_int32)]
    L.hv_mesh_counts.argtypes = [vp, P(C.c_int64), P(C.c_int32), P(C.c_int64)]
    L.hv_mesh_export.argtypes = [vp, P(C.c_int64), P(C.c_double)]
    L.hv_mesh_stats.argtypes = [vp, P(hv_stats_t)]
    L.hv_mesh_free.argtypes = [vp]
    L.hv_mesh_images.argtypes = [vp, P(C.c_int64), P(C.c_int64), P(C.c_int8)]
    L.hv_refine_mesh.argtypes = [vp, P(hv_opts), vp, C.c_int64, P(C.c_double), P(vp), P(C.c_int64)]
    L.hv_mesh_export_db.argtypes = [vp, P(C.c_int64), P(C.c_int64), P(C.c_int64)]
    L.hv_raycast_batch_keys.argtypes = [vp, C.c_int, C.c_int64, P(C.c_int64), P(C.c_int32), P(C.c_double),
                                        C.c_int32, P(C.c_int64), P(C.c_int32), P(C.c_double), P(C.c_double),
                                        P(C.c_int32)]
    L.hv_raycast_batch.argtypes = [vp, C.c_int, C.c_int64, P(C.c_int64), P(C.c_int32), P(C.c_double),
                                   P(C.c_int64), P(C.c_double), P(C.c_double), P(C.c_int32)]
    L.hv_raycast_batch_noalloc.argtypes = [vp, C.c_int, C.c_int64, vp, vp, vp, vp, vp, vp, vp]
    L.hv_nn_halfspace_batch.argtypes = [vp, C.c_int64, P(C.c_double), P(C.c_double), P(C.c_double),
                                        P(C.c_int64), P(C.c_double)]
    L.hv_inrange_batch.argtypes = [vp, C.c_int64, P(C.c_double), P(C.c_double), P(C.c_int64), P(C.c_int64)]
    L.hv_integrate_polytope.argtypes = [vp, vp, P(vp)]
    L.hv_integrate_mc.argtypes = [vp, vp, C.c_int64, C.c_uint64, P(vp)]
    L.hv_integrate_mc_points.argtypes = [vp, vp, C.c_int64, C.c_uint64, C.c_double, P(vp)]
    L.hv_integral_counts.argtypes = [vp, P(C.c_int64), P(C.c_int64)]
    L.hv_integral_work.argtypes = [vp, P(C.c_int64), P(C.c_double)]
    L.hv_integral_export.argtypes = [vp, P(C.c_double), P(C.c_int64), P(C.c_int64), P(C.c_double), P(C.c_double)]
    L.hv_integral_free.argtypes = [vp]
    L.hv_integrate_polytope_f.argtypes = [vp, vp, C.c_int32, INTEGRAND, vp, P(vp)]
    L.hv_integral_export_f.argtypes = [vp, P(C.c_int32), P(C.c_double), P(C.c_double)]
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i64(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


class Mesh:
    """A device-resident vertex set (hv_mesh)."""

    def __init__(self, ctx, handle, d: int = 0):
        self._ctx = ctx
        self._h = handle
        self.d = d if d else ctx.d

    def __del__(self):
        self.free()

    def free(self):
        if getattr(self, "_h", None):
            lib().hv_mesh_free(self._h)
            self._h = None

    def counts(self):
        nv = C.c_int64(0); ms = C.c_int32(0); nu = C.c_int64(0)
        lib().hv_mesh_counts(self._h, C.byref(nv), C.byref(ms), C.byref(nu))
        return nv.value, ms.value, nu.value

    def export(self):
        """(keys[V, d+1] 1-based sorted int64, coords[V, d]) in lexicographic key order."""
        nv, ms, _ = self.counts()
        keys = np.zeros((nv, ms), dtype=np.int64)
        coords = np.zeros((nv, self.d), dtype=np.float64)
        rc = lib().hv_mesh_export(self._h, _i64(keys), _dp(coords))
        if rc != HV_OK:
            raise HVError(rc, "export failed")
        return keys, coords

    def export_db(self):
        """DatabaseVertexStorage layout (hv_mesh_export_db): heap[V * (2d + 2)] int64 words
        [d + 1, sig..., r... as FP64 bits] per vertex in export order, cell_ptr[n_ext + 1] and
        refs[cell_ptr[-1]]: per generator the 1-based heap positions of its vertices, negative
        where it is the key's first entry."""
        nv, _, _ = self.counts()
        n = C.c_int64(0)
        lib().hv_mesh_images(self._h, C.byref(n), None, None)
        ptr = np.zeros(n.value + 1, dtype=np.int64)
        rc = lib().hv_mesh_export_db(self._h, None, _i64(ptr), None)
        if rc != HV_OK:
            raise HVError(rc, "export failed")
        heap = np.zeros(nv * (2 * self.d + 2), dtype=np.int64)
        refs = np.zeros(max(int(ptr[-1]), 1), dtype=np.int64)
        rc = lib().hv_mesh_export_db(self._h, _i64(heap), _i64(ptr), _i64(refs))
        if rc != HV_OK:
            raise HVError(rc, "export failed")
        return heap, ptr, refs[:int(ptr[-1])]

    def stats(self) -> dict:
        s = hv_stats_t()
        lib().hv_mesh_stats(self._h, C.byref(s))
        return s.as_dict()

    def images(self):
        """(orig[n_ext] 1-based, shift[n_ext, d] int8): generator id i+1 of this mesh is the
        original orig[i] moved by shift[i] box lengths (the identity for a non-periodic mesh)."""
        n = C.c_int64(0)
        rc = lib().hv_mesh_images(self._h, C.byref(n), None, None)
        if rc != HV_OK:
            raise HVError(rc, "images failed")
        orig = np.zeros(n.value, dtype=np.int64)
        shift = np.zeros((n.value, self.d), dtype=np.int8)
        lib().hv_mesh_images(self._h, C.byref(n), _i64(orig), shift.ctypes.data_as(C.POINTER(C.c_int8)))
        return orig, shift


class Context:
    """One hv_ctx: points, boundary and a HIP stream on one device."""

    def __init__(self, device: int = 0):
        L = lib()
        h = C.c_void_p()
        rc = L.hv_create(C.byref(h), device)
        if rc != HV_OK:
            raise HVError(rc, "hv_create failed (no GPU visible?)")
        self._h = h
        self.d = 0
        self.n = 0

    def __del__(self):
        if getattr(self, "_h", None):
            lib().hv_destroy(self._h)
            self._h = None

    def _check(self, rc):
        if rc != HV_OK:
            raise HVError(rc, lib().hv_last_error(self._h).decode())

    def set_points(self, xs: np.ndarray):
        xs = np.ascontiguousarray(xs, dtype=np.float64)
        n, d = xs.shape
        self._check(lib().hv_set_points(self._h, d, n, _dp(xs)))
        self.d, self.n = d, n

    def set_points_device(self, ptr: int, n: int, d: int):
        self._check(lib().hv_set_points_device(self._h, d, n, C.c_void_p(ptr)))
        self.d, self.n = d, n

    def set_boundary(self, base: np.ndarray, normal: np.ndarray, bc=None):
        base = np.ascontiguousarray(base, dtype=np.float64)
        normal = np.ascontiguousarray(normal, dtype=np.float64)
        npl = base.shape[0]
        bcp = None
        if bc is not None:
            bcp = np.ascontiguousarray(bc, dtype=np.int16).ctypes.data_as(C.POINTER(C.c_int16))
        self._check(lib().hv_set_boundary(self._h, npl, _dp(base), _dp(normal), bcp))

    def build_mesh(self, precision: str = "fp64", leaf_size: int = 0, vertex_capacity: int = 0,
                   chunks: int = 0, rays_per_block: int = 0, packet_floor: int = 0, seed: int = 0,
                   packet_min: int = 0) -> Mesh:
        o = hv_opts()
        o.precision = PRECISIONS[precision]
        o.leaf_size = leaf_size
        o.reserved0 = chunks
        o.vertex_capacity = vertex_capacity
        o.rays_per_block = rays_per_block
        o.packet_floor = packet_floor
        o.packet_min = packet_min
        o.seed = seed
        m = C.c_void_p()
        self._check(lib().hv_build_mesh(self._h, C.byref(o), C.byref(m)))
        return Mesh(self, m)

    def tune(self, precision: str = "bf16", cache_path: str | None = None):
        """Autotune the scan tiles on this context's points (hv_tune): (rays_per_block,
        packet_floor, ms) chosen for this dimension, precision and size class."""
        a = C.c_int32(0); b = C.c_int32(0); t = C.c_double(0.0)
        path = cache_path.encode() if cache_path else None
        self._check(lib().hv_tune(self._h, PRECISIONS[precision], path, C.byref(a), C.byref(b), C.byref(t)))
        return a.value, b.value, t.value

    def refine(self, mesh: "Mesh", new_points, precision: str = "fp64", leaf_size: int = 0):
        """refine!(VG, xs) (hv_refine_mesh): the mesh of mesh's generators (of a periodic mesh:
        its originals) followed by new_points[m, d] (ids n+1..n+m); the context's points become
        the union.  Returns (mesh, number of old vertices dropped)."""
        x = np.ascontiguousarray(new_points, dtype=np.float64)
        if x.ndim != 2 or x.shape[1] != mesh.d:
            raise ValueError("new_points must be [m, d]")
        o = _opts(precision, leaf_size)
        m = C.c_void_p()
        rm = C.c_int64(0)
        self._check(lib().hv_refine_mesh(self._h, C.byref(o), mesh._h, x.shape[0], _dp(x), C.byref(m), C.byref(rm)))
        out = Mesh(self, m, mesh.d)
        orig, _ = out.images()   # periodic meshes: originals first, then images
        self.d, self.n = mesh.d, int(orig.max()) if len(orig) else 0
        return out, rm.value

    def build_mesh_cells(self, cells, precision: str = "fp64", leaf_size: int = 0) -> Mesh:
        """The vertices of the cells `cells` (1-based generator ids) only: every vertex whose
        key holds one of them (voronoi(mesh; Iter=cells) in the reference)."""
        cells = np.ascontiguousarray(cells, dtype=np.int64).ravel()
        o = _opts(precision, leaf_size)
        m = C.c_void_p()
        self._check(lib().hv_build_mesh_cells(self._h, C.byref(o), len(cells),
                                              _i64(cells) if len(cells) else None, C.byref(m)))
        return Mesh(self, m)

    def raycast_batch_keys(self, sig, drop, r, max_sig: int = 32, precision: str = "fp64"):
        """hv_raycast_batch_keys: walkray with the reference's degenerate inclusion.  Returns
        (keys[n, max_sig] 0-padded, lengths[n], coords[n, d], t[n], status[n]); status 3 marks a
        key of more than d + 1 ids."""
        sig = np.ascontiguousarray(sig, dtype=np.int64)
        drop = np.ascontiguousarray(drop, dtype=np.int32)
        r = np.ascontiguousarray(r, dtype=np.float64)
        n, d = sig.shape[0], self.d
        keys = np.zeros((n, max_sig), dtype=np.int64)
        lens = np.zeros(n, dtype=np.int32)
        out_r = np.zeros((n, d))
        out_t = np.zeros(n)
        st = np.zeros(n, dtype=np.int32)
        self._check(lib().hv_raycast_batch_keys(self._h, PRECISIONS[precision], n, _i64(sig),
                                                drop.ctypes.data_as(C.POINTER(C.c_int32)), _dp(r), max_sig,
                                                _i64(keys), lens.ctypes.data_as(C.POINTER(C.c_int32)), _dp(out_r),
                                                _dp(out_t), st.ctypes.data_as(C.POINTER(C.c_int32))))
        return keys, lens, out_r, out_t, st

    def raycast_batch(self, sig, drop, r, precision: str = "fp64"):
        sig = np.ascontiguousarray(sig, dtype=np.int64)
        drop = np.ascontiguousarray(drop, dtype=np.int32)
        r = np.ascontiguousarray(r, dtype=np.float64)
        n = sig.shape[0]
        d = self.d
        out_sig = np.zeros((n, d + 1), dtype=np.int64)
        out_r = np.zeros((n, d))
        out_t = np.zeros(n)
        out_st = np.zeros(n, dtype=np.int32)
        self._check(lib().hv_raycast_batch(self._h, PRECISIONS[precision], n, _i64(sig),
                                           drop.ctypes.data_as(C.POINTER(C.c_int32)), _dp(r), _i64(out_sig),
                                           _dp(out_r), _dp(out_t), out_st.ctypes.data_as(C.POINTER(C.c_int32))))
        return out_sig, out_r, out_t, out_st

    def nn_halfspace(self, q, u, c):
        q = np.ascontiguousarray(q, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        c = np.ascontiguousarray(c, dtype=np.float64)
        n = q.shape[0]
        idx = np.zeros(n, dtype=np.int64)
        dist = np.zeros(n)
        self._check(lib().hv_nn_halfspace_batch(self._h, n, _dp(q), _dp(u), _dp(c), _i64(idx), _dp(dist)))
        return idx, dist

    def integrate(self, mesh: Mesh, method: str = "polytope", ndir: int = 1000, seed: int = 0,
                  nsamples: int = 10_000_000, slab: float = 0.0,
                  integrand=None, nf: int = 1) -> dict:
        """Cell volumes and interface areas of `mesh` (built from this context's points):
        method "polytope" (exact), "montecarlo" (ndir directions per cell) or "points"
        (nsamples stratified point samples, slab width factor `slab`).  Returns
        {"volumes", "nbr_ptr", "neighbors" (1-based, planes n+p), "areas", "ms"}.
        integrand (polytope only): f(points[npts, d]) -> values[npts, nf], called on the host;
        adds "bulk" [ncells, nf] and "interface" [nrows, nf] (the reference's quadrature)."""
        L = lib()
        h = C.c_void_p()
        if method == "polytope" and integrand is not None:
            d = self.d
            err = []

            def tramp(_user, npts, pts, vals):
                try:
                    x = np.ctypeslib.as_array(pts, shape=(npts, d))
                    out = np.ctypeslib.as_array(vals, shape=(npts, nf))
                    out[:] = np.asarray(integrand(x.copy()), dtype=np.float64).reshape(npts, nf)
                except Exception as e:   # never raise through C; report after the call
                    err.append(e)

            cb = INTEGRAND(tramp)
            rc = L.hv_integrate_polytope_f(self._h, mesh._h, int(nf), cb, None, C.byref(h))
            if err:
                if rc == HV_OK:
                    L.hv_integral_free(h)
                # popped, not indexed: the list would tie the exception to this frame (and the
                # context) in a cycle that only the collector frees
                raise err.pop()
            self._check(rc)
        elif method == "polytope":
            self._check(L.hv_integrate_polytope(self._h, mesh._h, C.byref(h)))
        elif method == "montecarlo":
            self._check(L.hv_integrate_mc(self._h, mesh._h, int(ndir), int(seed), C.byref(h)))
        elif method == "points":
            self._check(L.hv_integrate_mc_points(self._h, mesh._h, int(nsamples), int(seed), float(slab),
                                                 C.byref(h)))
        else:
            raise ValueError(f"unknown integration method {method!r}")
        try:
            nc = C.c_int64(0); nn = C.c_int64(0)
            L.hv_integral_counts(h, C.byref(nc), C.byref(nn))
            vol = np.zeros(nc.value)
            ptr = np.zeros(nc.value + 1, dtype=np.int64)
            nbr = np.zeros(max(1, nn.value), dtype=np.int64)
            area = np.zeros(max(1, nn.value))
            ms = C.c_double(0.0)
            rc = L.hv_integral_export(h, _dp(vol), _i64(ptr), _i64(nbr), _dp(area), C.byref(ms))
            if rc != HV_OK:
                raise HVError(rc, "integral export failed")
            k = C.c_int32(0)
            L.hv_integral_export_f(h, C.byref(k), None, None)
            extra = {}
            if k.value:
                bulk = np.zeros((nc.value, k.value))
                iface = np.zeros((max(1, nn.value), k.value))
                L.hv_integral_export_f(h, C.byref(k), _dp(bulk), _dp(iface))
                extra = {"bulk": bulk, "interface": iface[: nn.value]}
            work, kms = C.c_int64(0), C.c_double(0.0)
            L.hv_integral_work(h, C.byref(work), C.byref(kms))
            if work.value:
                extra.update(pairs=work.value, ms_kernel=kms.value)
        finally:
            L.hv_integral_free(h)
        return {"volumes": vol, "nbr_ptr": ptr, "neighbors": nbr[: nn.value], "areas": area[: nn.value],
                "ms": ms.value, **extra}

    def inrange(self, q, radius):
        q = np.ascontiguousarray(q, dtype=np.float64)
        radius = np.ascontiguousarray(radius, dtype=np.float64)
        n = q.shape[0]
        ptr = np.zeros(n + 1, dtype=np.int64)
        self._check(lib().hv_inrange_batch(self._h, n, _dp(q), _dp(radius), _i64(ptr), None))
        idx = np.zeros(max(int(ptr[-1]), 1), dtype=np.int64)
        self._check(lib().hv_inrange_batch(self._h, n, _dp(q), _dp(radius), _i64(ptr), _i64(idx)))
        return ptr, idx[: ptr[-1]]


def _opts(precision: str = "fp64", leaf_size: int = 0, vertex_capacity: int = 0, chunks: int = 0):
    o = hv_opts()
    o.precision = PRECISIONS[precision]
    o.leaf_size = leaf_size
    o.reserved0 = chunks
    o.vertex_capacity = vertex_capacity
    return o


def build_mesh_batch(point_sets, planes, precision: str = "fp64", nworkers: int = 0, device: int = 0):
    """Independent meshes in one hv_build_mesh_batch call.  point_sets[k]: (n_k, d_k) array;
    planes[k]: (base, normal), each (nplanes_k, d_k), or (base, normal, bc) with the planes'
    conditions (cuboid(...).arrays(): periodic sides are built with images).  Returns (meshes,
    status): a Mesh or None per input and the per-mesh HV_* codes."""
    L = lib()
    m = len(point_sets)
    xs = [np.ascontiguousarray(x, dtype=np.float64) for x in point_sets]
    bases = [np.ascontiguousarray(pl[0], dtype=np.float64) for pl in planes]
    normals = [np.ascontiguousarray(pl[1], dtype=np.float64) for pl in planes]
    bcs = [np.ascontiguousarray(pl[2], dtype=np.int16) if len(pl) > 2 and pl[2] is not None else None
           for pl in planes]
    d = np.array([x.shape[1] for x in xs], dtype=np.int32)
    n = np.array([x.shape[0] for x in xs], dtype=np.int64)
    npl = np.array([b.shape[0] for b in bases], dtype=np.int32)
    PD = C.POINTER(C.c_double)
    xp = (PD * max(1, m))(*[_dp(x) for x in xs])
    bp = (PD * max(1, m))(*[_dp(b) for b in bases])
    np_ = (PD * max(1, m))(*[_dp(nr) for nr in normals])
    PS = C.POINTER(C.c_int16)
    bcp = None
    if any(b is not None for b in bcs):
        bcp = (PS * max(1, m))(*[b.ctypes.data_as(PS) if b is not None else PS() for b in bcs])
    out = (C.c_void_p * max(1, m))()
    status = np.zeros(max(1, m), dtype=np.int32)
    o = _opts(precision)
    L.hv_build_mesh_batch(device, m, d.ctypes.data_as(C.POINTER(C.c_int32)), _i64(n), xp,
                          npl.ctypes.data_as(C.POINTER(C.c_int32)), bp, np_, bcp, C.byref(o), nworkers, out,
                          status.ctypes.data_as(C.POINTER(C.c_int32)))
    meshes = [Mesh(None, C.c_void_p(out[k]), int(d[k])) if out[k] else None for k in range(m)]
    return meshes, status[:m].copy()


def release_batch(device: int = 0):
    """Frees the contexts and device tables that build_mesh_batch's fused walks keep between
    calls on `device` (all devices when negative); meshes already built stay valid."""
    lib().hv_batch_release(int(device))
