"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this,
and only as the checker / CPU baseline.  Keys follow the reference's 1-based
convention: generators 1..n, boundary plane p is n+p (src/extended.jl:27-34).

polytope_integrand (the reference's integrand rule) is pinned by a hand-computed 2D cell,
by f = 1 and by affine exactness in 2D and 3D; for non-affine f in d >= 4 its values are
parity unpinned beyond that dimension-independent recursion.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build(force: bool = False) -> str:
    srcs = [os.path.join(_HERE, f) for f in ("hv_oracle.c", "hv_integrate.c")]
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < max(map(os.path.getmtime, srcs)):
        subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.orc_ctx_create.argtypes = [C.c_int, C.c_int64, P(C.c_double), C.c_int, P(C.c_double), P(C.c_double), P(C.c_void_p)]
        L.orc_ctx_free.argtypes = [C.c_void_p]
        L.orc_voronoi.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_uint64, P(C.c_void_p)]
        L.orc_mesh_nvert.argtypes = [C.c_void_p]
        L.orc_mesh_nvert.restype = C.c_int64
        L.orc_mesh_maxsig.argtypes = [C.c_void_p]
        L.orc_mesh_export.argtypes = [C.c_void_p, P(C.c_int64), P(C.c_double)]
        L.orc_mesh_stats.argtypes = [C.c_void_p, P(C.c_int64)]
        L.orc_ctx_stats.argtypes = [C.c_void_p, P(C.c_int64)]
        L.orc_mesh_free.argtypes = [C.c_void_p]
        L.orc_walkray.argtypes = [C.c_void_p, P(C.c_int64), C.c_int, P(C.c_double), C.c_int64,
                                  P(C.c_int64), P(C.c_int), P(C.c_double), P(C.c_double)]
        L.orc_last_error.restype = C.c_char_p
        L.orc_polytope.argtypes = [C.c_int, C.c_int64, P(C.c_double), C.c_int, P(C.c_double), P(C.c_double),
                                   C.c_int64, P(C.c_int64), P(C.c_double), P(C.c_int64), P(C.c_int64),
                                   P(C.c_int64), P(C.c_int64), P(C.c_double), P(C.c_double)]
        L.orc_mc_dirs.argtypes = [C.c_int, C.c_int64, P(C.c_double), C.c_int, P(C.c_double), P(C.c_double),
                                  P(C.c_int64), P(C.c_int64), C.c_int64, C.c_uint64, P(C.c_double), P(C.c_double)]
        L.orc_mc_direction.argtypes = [C.c_int, C.c_uint64, C.c_int64, C.c_int64, C.c_int64, P(C.c_double)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


STAT_NAMES = ["walkrays", "nn_calls", "inrange_calls", "descents", "tolerance_breach",
              "irreparable", "duplicates", "fraud", "unbounded", "hp_full_mode", "near_ties", "unused"]


class Oracle:
    """FP64 CPU restatement of VoronoiGeometry's mesh step (RCNonGeneralHP, VI_KD)."""

    def __init__(self, xs: np.ndarray, base: np.ndarray | None = None, normal: np.ndarray | None = None):
        L = lib()
        self.xs = np.ascontiguousarray(xs, dtype=np.float64)
        self.n, self.d = self.xs.shape
        if base is None:
            base = np.zeros((0, self.d)); normal = np.zeros((0, self.d))
        self.base = np.ascontiguousarray(base, dtype=np.float64)
        self.normal = np.ascontiguousarray(normal, dtype=np.float64)
        self.np = self.base.shape[0]
        h = C.c_void_p()
        rc = L.orc_ctx_create(self.d, self.n, _dp(self.xs), self.np, _dp(self.base), _dp(self.normal), C.byref(h))
        if rc != 0:
            raise ValueError(L.orc_last_error().decode())
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_ctx_free(self._h)
            self._h = None

    def voronoi(self, cells: tuple[int, int] | None = None, seed: int = 0):
        """Returns (keys[V, d+1] 1-based sorted, coords[V, d], stats dict)."""
        L = lib()
        b, e = (0, self.n) if cells is None else cells
        m = C.c_void_p()
        rc = L.orc_voronoi(self._h, b, e, seed, C.byref(m))
        if rc != 0:
            raise RuntimeError(f"oracle walk failed ({rc}): {L.orc_last_error().decode()}")
        try:
            nv = L.orc_mesh_nvert(m)
            ms = L.orc_mesh_maxsig(m)
            keys = np.zeros((nv, ms), dtype=np.int64)
            coords = np.zeros((nv, self.d), dtype=np.float64)
            L.orc_mesh_export(m, _ip(keys), _dp(coords))
            st = np.zeros(len(STAT_NAMES), dtype=np.int64)
            L.orc_mesh_stats(m, _ip(st))
        finally:
            L.orc_mesh_free(m)
        return keys, coords, dict(zip(STAT_NAMES, st.tolist()))

    def walkray_stats(self):
        """Counters (STAT_NAMES) of the last walkray call."""
        st = np.zeros(len(STAT_NAMES), dtype=np.int64)
        lib().orc_ctx_stats(self._h, _ip(st))
        return dict(zip(STAT_NAMES, st.tolist()))

    def walkray(self, sig, drop: int, r, cell: int):
        """One reference walkray; returns (new_key, coords, u) or None if unbounded."""
        L = lib()
        sig = np.ascontiguousarray(sig, dtype=np.int64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        out_sig = np.zeros(64, dtype=np.int64)
        out_len = C.c_int(0)
        out_r = np.zeros(self.d)
        out_u = np.zeros(self.d)
        rc = L.orc_walkray(self._h, _ip(sig), drop, _dp(r), cell, _ip(out_sig), C.byref(out_len), _dp(out_r), _dp(out_u))
        if rc == 1:
            return None
        if rc != 0:
            raise RuntimeError(L.orc_last_error().decode())
        return out_sig[: out_len.value].copy(), out_r, out_u


def sort_rows(keys: np.ndarray) -> np.ndarray:
    """Canonical order of a key table: lexicographic over the (already sorted) rows."""
    keys = np.asarray(keys)
    if len(keys) == 0:
        return keys
    order = np.lexsort(keys.T[::-1])
    return keys[order]


# ---- integration (src/polyintegrator.jl, src/mcintegrator.jl) ---------------------------

def adjacency(keys: np.ndarray, n: int):
    """Cell -> vertex rows and cell -> neighbour ids (sorted, 1-based, planes n+p) from a
    key table (src/neighbors.jl:219-251: every id co-occurring in a vertex key)."""
    keys = np.asarray(keys, dtype=np.int64)
    nv, w = keys.shape
    rows = np.repeat(np.arange(nv, dtype=np.int64), w)
    ids = keys.reshape(-1)
    real = (ids >= 1) & (ids <= n)
    cell = ids[real] - 1
    order = np.lexsort((rows[real], cell))
    cell_vert = rows[real][order]
    cell_ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(cell_ptr, cell[order] + 1, 1)
    cell_ptr = np.cumsum(cell_ptr)
    a = np.repeat(keys, w, axis=1).reshape(-1)
    b = np.tile(keys, (1, w)).reshape(-1)
    keep = (a != b) & (a >= 1) & (a <= n) & (b >= 1)
    pairs = np.unique(np.stack([a[keep] - 1, b[keep]], axis=1), axis=0)
    nbr_ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(nbr_ptr, pairs[:, 0] + 1, 1)
    nbr_ptr = np.cumsum(nbr_ptr)
    return cell_ptr, cell_vert, nbr_ptr, np.ascontiguousarray(pairs[:, 1])


def polytope(xs, base, normal, keys, coords):
    """Exact polytope volumes and interface areas: (vol[n], nbr_ptr, nbr, area)."""
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    n, d = xs.shape
    base = np.ascontiguousarray(base, dtype=np.float64)
    normal = np.ascontiguousarray(normal, dtype=np.float64)
    keys = np.ascontiguousarray(keys, dtype=np.int64)
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    cell_ptr, cell_vert, nbr_ptr, nbr = adjacency(keys, n)
    vol = np.zeros(n)
    area = np.zeros(len(nbr))
    rc = lib().orc_polytope(d, n, _dp(xs), base.shape[0], _dp(base), _dp(normal), len(keys), _ip(keys), _dp(coords),
                            _ip(cell_ptr), _ip(np.ascontiguousarray(cell_vert)), _ip(nbr_ptr), _ip(nbr),
                            _dp(vol), _dp(area))
    if rc != 0:
        raise RuntimeError(f"oracle polytope integration failed ({rc})")
    return vol, nbr_ptr, nbr, area


def mc_dirs(xs, base, normal, nbr_ptr, nbr, ndir: int, seed: int):
    """Directional Monte-Carlo volumes and areas (directions drawn per cell): (vol, area)."""
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    n, d = xs.shape
    base = np.ascontiguousarray(base, dtype=np.float64)
    normal = np.ascontiguousarray(normal, dtype=np.float64)
    nbr_ptr = np.ascontiguousarray(nbr_ptr, dtype=np.int64)
    nbr = np.ascontiguousarray(nbr, dtype=np.int64)
    vol = np.zeros(n)
    area = np.zeros(len(nbr))
    rc = lib().orc_mc_dirs(d, n, _dp(xs), base.shape[0], _dp(base), _dp(normal), _ip(nbr_ptr), _ip(nbr),
                           ndir, seed, _dp(vol), _dp(area))
    if rc != 0:
        raise RuntimeError(f"oracle Monte-Carlo integration failed ({rc})")
    return vol, area


def mc_direction(d: int, seed: int, cell: int, ndir: int, q: int):
    u = np.zeros(d)
    lib().orc_mc_direction(d, seed, cell, ndir, q, _dp(u))
    return u


def polytope_integrand(xs, base, normal, keys, coords, f, nf: int, ncell: int | None = None):
    """Integrals of f (points[npts, d] -> values[npts, nf]) over the cells and the interfaces
    by the reference's polytope rule, pure Python over the face lattice of the vertex keys
    (small cases only).  Returns (bulk[n, nf], {(i, j): interface[nf]}) with 0-based cell i and
    1-based neighbour id j (planes n + p).

    The rule (src/polyintegrator.jl:174-353; the same in src/fastpolyintegrator.jl:272-476):
    a face G (the vertices holding the generator set S) is cut into the simplices of the chains
    of face centres (mean of a face's vertices, src/polyintegrator.jl:470-507) down to an edge,
    f entering only at the edge ends (src/polyintegrator.jl:262-269, the edge value
    (f(r1) + f(r2)) / 2 times the chain's measure, :176-189):
        J(edge) = |e| (f(v1) + f(v2)) / 2,   J(G) = sum_H dist(c_G, aff H) J(H) / dim G;
    a facet F of cell i: I(F) = (d-1)/d J(F) + f(c_F) |F| / d (:317-322); and the cell, with
    apex at its generator: bulk = f(x_i) vol_i / (d+1) + sum_F dist(x_i, F) I(F) / (d+1)
    (:323-328), where |F| and vol_i follow the same recursion with f = 1.
    ncell: integrate only the cells 0..ncell-1 (the originals of a periodic mesh).
    """
    xs = np.asarray(xs, dtype=np.float64)
    n, d = xs.shape
    keys = np.asarray(keys, dtype=np.int64)
    coords = np.asarray(coords, dtype=np.float64)
    base = np.asarray(base, dtype=np.float64)
    normal = np.asarray(normal, dtype=np.float64)
    fv = np.asarray(f(coords), dtype=np.float64).reshape(len(coords), nf)
    ksets = [frozenset(int(v) for v in row) for row in keys]
    by_gen: dict[int, list[int]] = {}
    for r, ks in enumerate(ksets):
        for g in ks:
            by_gen.setdefault(g, []).append(r)
    memo: dict[frozenset, tuple] = {}

    def verts(S):
        rows = None
        for g in S:
            rg = set(by_gen.get(g, ()))
            rows = rg if rows is None else rows & rg
        return sorted(rows)

    def dist_to_aff(c, pts, k):
        # distance from c to the affine hull of pts, of dimension k
        v0 = pts[0]
        if k == 0:
            return float(np.linalg.norm(c - v0))
        A = (pts[1:] - v0).T
        u, sv, _ = np.linalg.svd(A, full_matrices=False)
        B = u[:, :k]
        r = (c - v0) - B @ (B.T @ (c - v0))
        return float(np.linalg.norm(r))

    def face(S):
        """(measure, J[nf], centre, vertex rows) of the face with generator set S."""
        if S in memo:
            return memo[S]
        rows = verts(S)
        dimG = d + 1 - len(S)
        pts = coords[rows]
        if dimG == 1:
            assert len(rows) == 2
            L = float(np.linalg.norm(pts[0] - pts[1]))
            out = (L, L * 0.5 * (fv[rows[0]] + fv[rows[1]]), pts.mean(axis=0), rows)
        else:
            c = pts.mean(axis=0)
            meas = 0.0
            J = np.zeros(nf)
            subs = set()
            for r in rows:
                for g in ksets[r]:
                    if g not in S:
                        subs.add(g)
            for g in sorted(subs):
                H = S | {g}
                mH, JH, _, rH = face(H)
                h = dist_to_aff(c, coords[rH], dimG - 1)
                meas += h * mH / dimG
                J = J + h * JH / dimG
            out = (meas, J, c, rows)
        memo[S] = out
        return out

    nc = n if ncell is None else ncell
    bulk = np.zeros((nc, nf))
    iface = {}
    fx = np.asarray(f(xs[:nc]), dtype=np.float64).reshape(nc, nf)
    for i in range(nc):
        gi = i + 1
        nbrs = set()
        for r in by_gen.get(gi, ()):
            nbrs |= ksets[r]
        nbrs.discard(gi)
        vol = 0.0
        y = np.zeros(nf)
        for j in sorted(nbrs):
            mF, JF, cF, _ = face(frozenset((gi, j)))
            fc = np.asarray(f(cF[None, :]), dtype=np.float64).reshape(nf)
            I = (d - 1) / d * JF + fc * mF / d
            if j > n:
                p = j - n - 1
                dist = float(np.dot(normal[p], base[p] - xs[i]))
            else:
                dist = 0.5 * float(np.linalg.norm(xs[j - 1] - xs[i]))
            iface[(i, j)] = I
            vol += dist * mF / d
            y = y + dist * I / (d + 1)
        bulk[i] = fx[i] * vol / (d + 1) + y
    return bulk, iface
