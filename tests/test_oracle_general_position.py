"""The oracle at the edge of general position (the reference's winner rule,
src/raycast.jl:877-950): a constructed near tie resolved by the larger <u, x - x0>, perturbed
lattices that stay in general position, and an exact lattice, which the oracle reports as a
non-general vertex (its FastEdgeIterator path, src/edgeiterate.jl:82-295, is not restated).

The perturbation scan pins where the rule's two bands sit: the near-tie window
(10 + d) * full_error is narrower than the band within which the reference puts a further
generator into the key, so a perturbed lattice goes from no near ties straight to keys longer
than d + 1 (2D: general down to a perturbation of 2e-11, 3D: at 1e-5, not at 1e-6).
"""
import numpy as np
import pytest

from helpers import lattice, near_tie_points
from hvgpu import synth
from oracle_py import Oracle


def walk(delta, swap):
    xs, r = near_tie_points(delta, swap)
    b, nrm = synth.cube_planes(2)
    o = Oracle(xs, b, nrm)
    res = o.walkray(np.array([1, 2, 3]), 2, r, 1)
    return res, o.walkray_stats()


@pytest.mark.parametrize("delta,ties", [(1e-12, 0), (1e-13, 0), (3e-15, 1), (1e-15, 1)])
def test_constructed_near_tie(delta, ties):
    (key, r2, _), st = walk(delta, False)
    assert key.tolist() == [1, 2, 4]
    assert st["near_ties"] == ties
    # the new vertex is the centre of generator 4's circle, (0.5, 0.6)
    assert np.max(np.abs(r2 - [0.5, 0.6])) < 1e-13


@pytest.mark.parametrize("delta", [1e-13, 3e-15, 1e-15])
def test_swapped_candidates_are_no_tie(delta):
    # generator 5 at the top of its circle: it is not within the window of 4 (the first found)
    (key, _, _), st = walk(delta, True)
    assert key.tolist() == [1, 2, 4] and st["near_ties"] == 0


@pytest.mark.parametrize("d,k,eps", [(2, 30, 1e-9), (2, 30, 2e-11), (3, 8, 1e-5)])
def test_perturbed_lattice_stays_general(d, k, eps):
    xs = lattice(d, k, eps)
    b, nrm = synth.cube_planes(d)
    keys, _, st = Oracle(xs, b, nrm).voronoi()
    assert keys.shape[1] == d + 1 and st["near_ties"] == 0 and len(keys) > k ** d


@pytest.mark.parametrize("d,k,eps", [(2, 30, 0.0), (2, 30, 1e-11), (3, 8, 0.0), (3, 8, 1e-6)])
def test_lattice_is_not_general(d, k, eps):
    xs = lattice(d, k, eps)
    b, nrm = synth.cube_planes(d)
    with pytest.raises(RuntimeError, match="non-general"):
        Oracle(xs, b, nrm).voronoi()


@pytest.mark.parametrize("eps", [1e-10, 5e-11])
def test_rays_along_edges_shorter_than_plane_tolerance_end_without_a_vertex(eps):
    """A near-co-circular square has a Voronoi edge of length ~eps: a ray along it whose
    nearest candidate comes within t < plane_tolerance (1e-12) finds no vertex (its window is
    empty, src/raycast.jl:883-894); both end points are reached through other edges, and the
    mesh is the 2 (k - 1)^2 + 4 k vertices of a triangulated lattice."""
    xs = lattice(2, 30, eps)
    b, nrm = synth.cube_planes(2)
    keys, _, st = Oracle(xs, b, nrm).voronoi()
    assert st["unbounded"] > 0
    assert len(keys) == 2 * 29 ** 2 + 4 * 30


def test_clipped_mixture_is_not_general():
    """The 5D N=100k mixture with out-of-cube samples clipped onto the faces (instead of drawn
    again) holds generators on a boundary plane; such a generator coincides with its own
    mirror image and the walk of its cell reaches a vertex with more than d + 1 generators."""
    xs = synth.gaussian_mixture_5d(clip=True)
    on = np.nonzero(((xs == 0.0) | (xs == 1.0)).any(axis=1))[0]
    assert len(on) > 0
    b, nrm = synth.cube_planes(5)
    with pytest.raises(RuntimeError, match="non-general"):
        Oracle(xs, b, nrm).voronoi(cells=(int(on[0]), int(on[0]) + 1))


@pytest.mark.parametrize("d,k", [(2, 8), (3, 4)])
def test_degenerate_inclusion_on_exact_lattices(d, k):
    """The oracle's walkray keeps every generator and plane on the new vertex's sphere
    (src/raycast.jl:926-950): from a boundary vertex of an exact lattice the ray ends at the
    centre of a lattice square (2D: its 4 generators) or on a face at the centre of 4 cells
    (3D: 4 generators and the face), the closed-form keys."""
    from helpers import lattice_boundary_rays
    xs = lattice(d, k)
    b, nrm = synth.cube_planes(d)
    o = Oracle(xs, b, nrm)
    sig, drop, r, cell = lattice_boundary_rays(d, k)
    n = k ** d
    for q in range(len(sig)):
        key, r2, _ = o.walkray(sig[q], int(drop[q]), r[q], cell[q])
        if d == 2:
            j = q
            want = sorted([j + 1, j + 2, k + j + 1, k + j + 2])
            centre = [1.0 / k, (j + 1) / k]
        else:
            l = q
            want = sorted([l + 1, l + 2, k + l + 1, k + l + 2, n + 2])
            centre = [0.0, 1.0 / k, (l + 1) / k]
        assert key.tolist() == want
        assert np.max(np.abs(r2 - centre)) < 1e-12
