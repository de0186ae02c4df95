"""Exact-rational brute force of the bounded Voronoi vertex set (TEST INFRASTRUCTURE).

Pins the oracle and the GPU path on tiny inputs without the reference.  Every
double is an exact rational, so with ``fractions.Fraction`` the circumcentre of a
candidate key and all in-sphere / in-box tests are exact.

A candidate key is a set of d+1 elements drawn from the generators and the planes
of the box [0,1]^d, at most one plane per axis, at least one generator
(src/extended.jl:27-34, src/boundary.jl:510-534: plane 2k-1 is x_k = 1, plane 2k is
x_k = 0, key index n+plane).  Its point p solves
    2 <x_g - x_g0, p> = |x_g|^2 - |x_g0|^2   (generators g of the key)
    p_k = side                                (planes of the key)
and the key is a vertex iff p lies in the closed box and no generator is strictly
closer to p than the key's generators.  A generator at exactly the same distance,
or p on a plane outside the key, means the input is not in general position and is
reported as DegenerateInput.
"""
from __future__ import annotations

import itertools
from fractions import Fraction

import numpy as np


class DegenerateInput(Exception):
    pass


def _solve(A, b):
    n = len(A)
    M = [row[:] + [bb] for row, bb in zip(A, b)]
    for c in range(n):
        piv = next((r for r in range(c, n) if M[r][c] != 0), None)
        if piv is None:
            return None
        M[c], M[piv] = M[piv], M[c]
        inv = 1 / M[c][c]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c] * inv
                M[r] = [a - f * bq for a, bq in zip(M[r], M[c])]
    return [M[i][n] / M[i][i] for i in range(n)]


def exact_vertices(xs: np.ndarray):
    """All vertices of the Voronoi tessellation of xs restricted to [0,1]^d.

    Returns (keys, coords): keys is a sorted list of 1-based sorted tuples, coords
    the matching list of exact Fraction tuples.
    """
    xs = np.asarray(xs, dtype=np.float64)
    n, d = xs.shape
    X = [[Fraction(float(v)) for v in row] for row in xs]
    N2 = [sum(v * v for v in row) for row in X]
    out = {}
    axes = list(range(d))
    for m in range(0, d + 1):
        ng = d + 1 - m
        if ng < 1 or ng > n:
            continue
        for ax in itertools.combinations(axes, m):
            for sides in itertools.product((1, 0), repeat=m):
                fixed = dict(zip(ax, (Fraction(s) for s in sides)))
                free = [k for k in axes if k not in fixed]
                for gens in itertools.combinations(range(n), ng):
                    g0 = gens[0]
                    A, b = [], []
                    for g in gens[1:]:
                        row = [2 * (X[g][k] - X[g0][k]) for k in axes]
                        rhs = N2[g] - N2[g0]
                        rhs -= sum(row[k] * fixed[k] for k in fixed)
                        A.append([row[k] for k in free])
                        b.append(rhs)
                    if free:
                        sol = _solve(A, b)
                        if sol is None:
                            continue
                    else:
                        sol = []
                    p = [None] * d
                    for k, v in fixed.items():
                        p[k] = v
                    for k, v in zip(free, sol):
                        p[k] = v
                    if any(v < 0 or v > 1 for v in p):
                        continue
                    rho = sum((p[k] - X[g0][k]) ** 2 for k in axes)
                    ok = True
                    tie = False
                    gs = set(gens)
                    for h in range(n):
                        if h in gs:
                            continue
                        dh = sum((p[k] - X[h][k]) ** 2 for k in axes)
                        if dh < rho:
                            ok = False
                            break
                        if dh == rho:
                            tie = True
                    if not ok:
                        continue
                    if tie:
                        raise DegenerateInput(f"co-spherical generators at key {gens}+planes{ax}")
                    for k in free:
                        if p[k] == 0 or p[k] == 1:
                            raise DegenerateInput("vertex on a plane outside its key")
                    key = [g + 1 for g in gens] + [n + (2 * k + 1 if s == 0 else 2 * k) + 1 for k, s in zip(ax, sides)]
                    key = tuple(sorted(key))
                    out[key] = tuple(p)
    keys = sorted(out)
    return keys, [out[k] for k in keys]
