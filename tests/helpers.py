"""Shared helpers for the parity tests."""
from __future__ import annotations

import glob
import json
import os
from fractions import Fraction

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_cases():
    return sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(GOLDEN, "*.json")))


def load_golden(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        g = json.load(f)
    xs = np.array([[float.fromhex(v) for v in row] for row in g["points_hex"]])
    keys = np.array(g["keys"], dtype=np.int64)
    exact = [[Fraction(v) for v in row] for row in g["coords_exact"]]
    return g, xs, keys, exact


def key_set(keys):
    return set(map(tuple, np.asarray(keys).tolist()))


def rows_view(keys):
    """Sorted unique-row view of an int64 key table (for set algebra on large tables)."""
    keys = np.ascontiguousarray(keys, dtype=np.int64)
    v = keys.view(np.dtype((np.void, keys.dtype.itemsize * keys.shape[1]))).ravel()
    return np.sort(v)


def key_digest(keys):
    """Order-independent digest of a key table (sum and xor of a mixed row hash)."""
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    h = np.full(k.shape[0], 0x9E3779B97F4A7C15, dtype=np.uint64)
    with np.errstate(over="ignore"):
        for j in range(k.shape[1]):
            z = h ^ k[:, j]
            z = (z ^ (z >> np.uint64(33))) * np.uint64(0xFF51AFD7ED558CCD)
            z = (z ^ (z >> np.uint64(33))) * np.uint64(0xC4CEB9FE1A85EC53)
            h = z ^ (z >> np.uint64(33))
        s = int(np.sum(h, dtype=np.uint64))
    x = int(np.bitwise_xor.reduce(h)) if len(h) else 0
    return (len(h), s, x)


def edge_closure_defects(keys, n):
    """Number of walkable edges (d-subsets of a key holding a generator) that do not
    appear in exactly two vertex keys -- zero for a complete bounded tessellation."""
    keys = np.asarray(keys, dtype=np.int64)
    V, m = keys.shape
    edges = []
    for p in range(m):
        e = np.delete(keys, p, axis=1)
        e = e[e[:, 0] <= n]
        edges.append(e)
    e = np.concatenate(edges)
    v = rows_view(e)
    _, counts = np.unique(v, return_counts=True)
    return int(np.sum(counts != 2))


def coords_by_key(keys, coords):
    return {tuple(k): c for k, c in zip(np.asarray(keys).tolist(), np.asarray(coords))}


# ---- periodic boxes ---------------------------------------------------------------------------
# Labels make a periodic mesh comparable across constructions: generator (original o, shift s)
# is (0, o, *s), boundary plane p (1-based) is (1, p, 0, ..., 0); a vertex is the sorted tuple
# of its generators' labels.

def periodic_tiling(xs, per, copies=1, lo=0.0, hi=1.0):
    """The points and their images shifted by -copies..copies box lengths in the directions
    `per` (0-based), originals first, in the box widened by `copies` lengths there (Dirichlet
    planes in the cuboid order).  Returns (tiled, orig, shift, base, normal)."""
    from itertools import product
    n, d = xs.shape
    lo = np.broadcast_to(np.asarray(lo, dtype=np.float64), (d,))
    hi = np.broadcast_to(np.asarray(hi, dtype=np.float64), (d,))
    L = hi - lo
    shifts = [np.zeros(d, dtype=np.int64)]
    for s in product(range(-copies, copies + 1), repeat=len(per)):
        if any(s):
            v = np.zeros(d, dtype=np.int64)
            v[list(per)] = s
            shifts.append(v)
    tiled = np.vstack([xs + L * s for s in shifts])
    orig = np.tile(np.arange(n), len(shifts))
    shift = np.repeat(np.array(shifts), n, axis=0)
    base, normal = [], []
    for k in range(d):
        w = copies * L[k] if k in per else 0.0
        e = np.zeros(d)
        e[k] = 1.0
        b = np.zeros(d); b[k] = hi[k] + w
        base.append(b); normal.append(e.copy())
        b = np.zeros(d); b[k] = lo[k] - w
        base.append(b); normal.append(-e)
    return tiled, orig, shift, np.array(base), np.array(normal)


def vertex_labels(keys, n_ext, orig0, shift):
    """Label tuples of the rows of a 1-based key table whose ids 1..n_ext are generators
    (orig0: 0-based originals, shift: int rows) and whose larger ids are planes."""
    d = shift.shape[1]
    out = []
    for row in np.asarray(keys).tolist():
        lab = []
        for i in row:
            if i <= n_ext:
                lab.append((0, int(orig0[i - 1])) + tuple(int(v) for v in shift[i - 1]))
            else:
                lab.append((1, int(i - n_ext)) + (0,) * d)
        out.append(tuple(sorted(lab)))
    return out


def periodic_oracle(xs, per, copies=1, lo=0.0, hi=1.0):
    """Vertices of the originals' cells in the periodic box [lo, hi] (directions `per`), from
    the oracle's walk over the cells 1..n of the tiled set: {label tuple: coordinates}."""
    from oracle_py import Oracle
    n = xs.shape[0]
    tiled, orig, shift, base, normal = periodic_tiling(xs, per, copies, lo, hi)
    keys, coords, _ = Oracle(tiled, base, normal).voronoi(cells=(0, n))
    labs = vertex_labels(keys, len(tiled), orig, shift)
    return dict(zip(labs, coords))


def periodic_mesh_labels(mesh):
    """{label tuple: coordinates} of a device mesh built on a periodic box."""
    keys, coords = mesh.export()
    orig, shift = mesh.images()
    labs = vertex_labels(keys, len(orig), orig - 1, shift)
    return dict(zip(labs, coords))


def near_tie_points(delta: float, swap: bool = False, theta: float = np.pi / 6):
    """A 2D point set whose ray from the vertex of generators 1, 2, 3 (centre (0.5, 0.425)),
    walking away from generator 3 along +y, meets two candidates whose end points differ by
    `delta` in t: generator 4 (ids 1-based) at t = 0.175 and generator 5 at 0.175 + delta.
    Without swap, 4 lies at the top of its circle (the larger <u, x - x0> as well as the
    smaller t); with swap, 5 does.  Eight more points keep the rest of the box away."""
    a, b, c = np.array([0.4, 0.5]), np.array([0.6, 0.5]), np.array([0.5, 0.3])
    r4, r5 = np.hypot(0.1, 0.1), np.hypot(0.1, 0.1 + delta)
    top = lambda cy, rr: np.array([0.5, cy + rr])
    side = lambda cy, rr: np.array([0.5 + rr * np.cos(theta), cy + rr * np.sin(theta)])
    p, q = (side(0.6, r4), top(0.6 + delta, r5)) if swap else (top(0.6, r4), side(0.6 + delta, r5))
    fill = np.array([[0.1, 0.1], [0.9, 0.1], [0.1, 0.9], [0.9, 0.9], [0.2, 0.5], [0.8, 0.5], [0.5, 0.05], [0.5, 0.97]])
    return np.vstack([a, b, c, p, q, fill]), np.array([0.5, 0.425])


def lattice(d: int, k: int, eps: float = 0.0, seed: int = 1):
    """k^d cell centres of the unit cube, each moved by U(-eps, eps) per coordinate."""
    g = (np.stack(np.meshgrid(*[np.arange(k)] * d, indexing="ij"), -1).reshape(-1, d) + 0.5) / k
    if eps:
        g = g + eps * np.random.default_rng(seed).uniform(-1, 1, g.shape)
    return g


def lattice_boundary_rays(d: int, k: int):
    """Rays of the exact lattice(d, k) (k a power of two: coordinates exact in binary) that
    start at a general boundary vertex and end at a vertex out of general position, with the
    cell the reference walks them from.  2D: the vertex between cells (0, j) and (0, j+1) on
    the face x_1 = 0, walking into the box (its next vertex is a square's centre, 4 generators
    on its circle).  3D: the vertex between cells (0, 0, l) and (0, 0, l+1) on the box edge
    x_1 = x_2 = 0, walking along the face x_1 = 0 (its next vertex has 4 generators and the
    face on its sphere).  Returns (sig 1-based sorted, drop, r, cell) lists."""
    n = k ** d
    idx = lambda *c: int(np.ravel_multi_index(c, (k,) * d)) + 1
    sig, drop, r, cell = [], [], [], []
    if d == 2:
        for j in range(k - 1):
            key = sorted([idx(0, j), idx(0, j + 1), n + 2])   # plane 2: x_1 = 0
            sig.append(key); drop.append(key.index(n + 2)); r.append([0.0, (j + 1) / k]); cell.append(idx(0, j))
    else:
        for l in range(k - 1):
            key = sorted([idx(0, 0, l), idx(0, 0, l + 1), n + 2, n + 4])   # planes x_1 = 0, x_2 = 0
            sig.append(key); drop.append(key.index(n + 4)); r.append([0.0, 0.0, (l + 1) / k]); cell.append(idx(0, 0, l))
    return np.array(sig), np.array(drop), np.array(r), cell
