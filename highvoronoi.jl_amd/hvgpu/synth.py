"""Synthetic generator sets for the benchmark and parity workloads.

A counter-based splitmix64 stream gives U[0,1) doubles (53-bit mantissa), so the
same (seed, index) yields the same point in Python and in C++.  Normals use the
Box-Muller transform on consecutive pairs of uniforms.

Workloads (all in the unit cube [0,1]^d):
  * uniform(d, n, seed)                      i.i.d. uniform points
  * gaussian_mixture_5d(n=100_000, seed=2)   8 isotropic Gaussians (centres uniform in
    [0.2,0.8]^5, sigma=0.05) holding 90% of the points plus 10% uniform background.
    Gaussian samples that fall outside the cube are redrawn (a truncated normal):
    clipping would put points exactly on a boundary plane, where a generator
    coincides with its own mirror image and the vertex set stops being unique.
  * batched_meshes(count=512, seed=3)        independent meshes, d in {3,4,5,6},
    n log-uniform in [200, 5000].
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, start: int, count: int) -> np.ndarray:
    """Outputs start..start+count-1 of the splitmix64 stream seeded with `seed`."""
    with np.errstate(over="ignore"):
        idx = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniforms(seed: int, count: int, start: int = 0) -> np.ndarray:
    z = splitmix64(seed, start, count)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform(d: int, n: int, seed: int) -> np.ndarray:
    """(n, d) array of i.i.d. U[0,1) points."""
    return uniforms(seed, n * d).reshape(n, d)


def normals(seed: int, count: int, start: int = 0) -> np.ndarray:
    m = (count + 1) // 2
    u = uniforms(seed, 2 * m, 2 * start)
    u1 = np.maximum(u[0::2], 1e-300)
    u2 = u[1::2]
    rad = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = rad * np.cos(2 * np.pi * u2)
    out[1::2] = rad * np.sin(2 * np.pi * u2)
    return out[:count]


def gaussian_mixture_5d(n: int = 100_000, seed: int = 2, clusters: int = 8, sigma: float = 0.05,
                        background: float = 0.10, d: int = 5, clip: bool = False) -> np.ndarray:
    """n points in [0, 1]^d: `clusters` Gaussian clusters (centres U[0.2, 0.8]^d, standard
    deviation sigma) and a uniform background fraction.  Cluster samples outside the open cube
    are drawn again, or with clip=True clipped onto the cube's faces (generators on a boundary
    plane)."""
    nb = int(round(n * background))
    nc = n - nb
    per = [nc // clusters + (1 if i < nc % clusters else 0) for i in range(clusters)]
    centres = 0.2 + 0.6 * uniforms(seed ^ 0xC3A5C85C97CB3127, clusters * d).reshape(clusters, d)
    pts = []
    stream = 0
    for ci in range(clusters):
        need = per[ci]
        got = []
        while need > 0:
            batch = max(2 * need, 64)
            z = normals(seed, batch * d, start=stream).reshape(batch, d)
            stream += batch * d
            cand = centres[ci] + sigma * z
            if clip:
                cand = np.clip(cand[:need], 0.0, 1.0)
            else:
                ok = np.all((cand > 0.0) & (cand < 1.0), axis=1)
                cand = cand[ok][:need]
            got.append(cand)
            need -= len(cand)
        pts.append(np.concatenate(got))
    pts.append(uniforms(seed ^ 0x9AE16A3B2F90404F, nb * d).reshape(nb, d))
    return np.ascontiguousarray(np.concatenate(pts))


def batched_meshes(count: int = 512, seed: int = 3, nmin: int = 200, nmax: int = 5000):
    """List of (d, points) for `count` independent meshes."""
    u = uniforms(seed, 2 * count)
    out = []
    for k in range(count):
        d = 3 + int(u[2 * k] * 4) % 4
        n = int(np.floor(np.exp(np.log(nmin) + u[2 * k + 1] * (np.log(nmax) - np.log(nmin)))))
        n = max(n, nmin)
        out.append((d, uniform(d, n, seed * (1 << 32) + k + 1)))
    return out


def cube_planes(d: int, lo: float = 0.0, hi: float = 1.0):
    """Planes of cuboid(d, periodic=[]) in the reference's order (src/boundary.jl:510-534):
    plane 2k-1 is x_k = hi with outer normal +e_k, plane 2k is x_k = lo with normal -e_k."""
    base = np.zeros((2 * d, d))
    normal = np.zeros((2 * d, d))
    for k in range(d):
        base[2 * k, k] = hi
        normal[2 * k, k] = 1.0
        base[2 * k + 1, k] = lo
        normal[2 * k + 1, k] = -1.0
    return base, normal
