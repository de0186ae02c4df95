"""CPU pinning of the integrand restatement (oracle_py.polytope_integrand), the reference's
polytope quadrature: f = 1 gives the oracle's exact volumes and areas; in d <= 3 affine f is
integrated exactly (sum over the cells and the divergence theorem per cell); a hand-computed
square cell fixes the weights of the non-affine rule.  From d = 4 on the rule is not exact for
affine f (below the facets it evaluates f only at edge ends, not at the intermediate face
centres of its simplices), so there only the f = 1 identities apply."""
import numpy as np
import pytest

from hvgpu import synth
from oracle_py import Oracle, polytope, polytope_integrand


def test_square_cell_by_hand():
    # one generator at the centre of the unit square: the cell is the square, vertices are
    # the corners (planes n+1: x=1, n+2: x=0, n+3: y=1, n+4: y=0)
    xs = np.array([[0.5, 0.5]])
    base, normal = synth.cube_planes(2)
    keys = np.array([[1, 3, 5], [1, 2, 5], [1, 3, 4], [1, 2, 4]])
    coords = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [1.0, 1.0]])
    f = lambda p: np.stack([np.ones(len(p)), p[:, 0], p[:, 0] ** 2], axis=1)
    bulk, iface = polytope_integrand(xs, base, normal, keys, coords, f, 3)
    # facet rule in 2D: I = |e| ((f(a) + f(b)) / 4 + f(mid) / 2); x^2 on the four sides:
    # x = 1: 1, x = 0: 0, y = 0 and y = 1: 1/4 + 1/8 = 3/8; bulk = (f(x) vol + sum h I) / 3
    assert np.allclose(bulk[0], [1.0, 0.5, (0.25 + 0.5 * (1 + 0 + 0.375 + 0.375)) / 3], atol=1e-15)
    assert np.allclose(iface[(0, 2)], [1.0, 1.0, 1.0]) and np.allclose(iface[(0, 3)], [1.0, 0.0, 0.0])
    assert np.allclose(iface[(0, 5)], [1.0, 0.5, 0.375])


def normals_of(xs, base, normal, i, j):
    n = len(xs)
    if j > n:
        return normal[j - n - 1]
    v = xs[j - 1] - xs[i]
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("d,n,seed", [(2, 40, 1), (3, 25, 2), (4, 12, 3)])
def test_restatement_exact_for_affine_f(d, n, seed):
    xs = synth.uniform(d, n, seed)
    base, normal = synth.cube_planes(d)
    keys, coords, _ = Oracle(xs, base, normal).voronoi()
    vol, ptr, nbr, area = polytope(xs, base, normal, keys, coords)
    f = lambda p: np.concatenate([np.ones((len(p), 1)), p], axis=1)
    bulk, iface = polytope_integrand(xs, base, normal, keys, coords, f, d + 1)
    assert np.allclose(bulk[:, 0], vol, rtol=1e-11, atol=0)
    for i in range(n):
        for t in range(ptr[i], ptr[i + 1]):
            assert abs(iface[(i, nbr[t])][0] - area[t]) <= 1e-11 * max(area[t], 1e-300) + 1e-15
    if d > 3:
        return
    # the cells tile the box: sum of the integrals of x_k is the box's, 1/2
    assert np.allclose(bulk[:, 1:].sum(axis=0), 0.5, atol=1e-12)
    # divergence theorem per cell for f = x_k: sum_F n_F,l int_F x_k = delta_kl vol
    for i in range(n):
        M = np.zeros((d, d))
        for t in range(ptr[i], ptr[i + 1]):
            nv = normals_of(xs, base, normal, i, nbr[t])
            M += np.outer(nv, iface[(i, nbr[t])][1:])
        assert np.allclose(M, vol[i] * np.eye(d), atol=1e-11)
