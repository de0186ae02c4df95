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
