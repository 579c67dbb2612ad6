"""Pins the oracle's mesh-collider functions: Collisions::ClossestPointOnTriangle
/ DistanceToTriangle (src/3d/collisions.cpp:21-218), TriangleMesh::IsInside
(src/3d/triangle_mesh.cpp:120-163), Mesh2SDF::Build (src/3d/mesh_2_sdf.cpp:14-95)
and ColliderCollection::BuildSdf over a TriangleMeshCollider
(src/3d/collider_collection.cpp:33-50).

  * the reference's known answers (test/collisions_test.cpp:13-49);
  * closest points against an independent candidate method (plane projection
    or clamped edge projections) on random triangles and points, which visit
    all seven parameter regions;
  * IsInside on a closed box with the answers of test/collisions_test.cpp:51-70
    (its test_model.obj is not in the source tree; a box [-1, 1]^3 gives the
    same answers) and on an icosphere away from its surface;
  * Mesh2SDF on a box: the band holds the exact box distances, nodes outside
    the band hold max(h) * (nx + ny + nz), nodes inside the box are negative;
    the band's index box divides by h.x without subtracting the grid origin;
  * the scene's collider SDF is the running minimum of the sampled collider
    SDFs.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from meshes import box, box_distance, closest_point_reference_free, icosphere


def test_known_answers():
    tri = ((1, 0, 0), (-1, 0, 0), (0, 0, 1))
    assert np.array_equal(orc.closest_on_triangle((0, 1, 0.5), *tri), [0, 0, 0.5])
    c = orc.closest_on_triangle((0, 0, 2), *tri)
    assert np.linalg.norm(c - np.array([0, 0, 2])) == 1.0


def test_closest_point_all_regions():
    rng = np.random.default_rng(3)
    for _ in range(3000):
        p = rng.normal(size=(3, 3))
        x = rng.normal(size=3) * 2
        got = orc.closest_on_triangle(x, *p)
        want, dist = closest_point_reference_free(x, *p)
        assert abs(np.linalg.norm(got - x) - dist) <= 1e-12 * max(1.0, dist)
        assert np.abs(got - want).max() <= 1e-9


def test_closest_point_degenerate_and_on_triangle():
    # a point on the triangle is its own closest point; sliver triangles
    p = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    for x in ([0.25, 0.25, 0.0], [0.0, 0.0, 0.0], [0.5, 0.5, 0.0]):
        assert np.allclose(orc.closest_on_triangle(x, *p), x, atol=1e-15)
    sliver = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [2.0, 1e-9, 0.0]])
    got = orc.closest_on_triangle([0.5, 1.0, 0.0], *sliver)
    want, dist = closest_point_reference_free([0.5, 1.0, 0.0], *sliver)
    assert abs(np.linalg.norm(got - [0.5, 1.0, 0.0]) - dist) < 1e-9


def test_is_inside_box_reference_answers():
    v, t = box((-1, -1, -1), (1, 1, 1))
    pts = [(0, 0, 0), (0.5, 0.5, 0.5), (0.99, 0.99, 0.99), (1, 1, 1), (1, 0, 1), (2, 1, 1)]
    want = [True, True, True, False, False, False]
    assert [orc.mesh_inside(v, t, p) for p in pts] == want
    # orientation does not matter (the magnitude of the angle sum is tested)
    assert [orc.mesh_inside(v, t[:, ::-1], p) for p in pts] == want


def test_is_inside_icosphere():
    v, t = icosphere((0.1, -0.2, 0.3), 0.7, levels=2)
    rng = np.random.default_rng(8)
    x = rng.uniform(-1.2, 1.2, (400, 3)) + np.array([0.1, -0.2, 0.3])
    r = np.linalg.norm(x - np.array([0.1, -0.2, 0.3]), axis=1)
    keep = np.abs(r - 0.7) > 0.05  # the facets lie within 0.7*(1 - cos) of the sphere
    for p, rr in zip(x[keep], r[keep]):
        assert orc.mesh_inside(v, t, p) == (rr < 0.7)


def test_mesh_to_sdf_box_band_fill_and_sign():
    n = (12, 10, 11)
    h = (0.1, 0.1, 0.1)
    lo, hi = (0.32, 0.27, 0.35), (0.71, 0.66, 0.78)
    v, t = box(lo, hi)
    sdf = orc.mesh_to_sdf(n, (0.0, 0.0, 0.0), h, v, t)
    Z, Y, X = np.meshgrid(*[np.arange(n[a]) * h[a] for a in (2, 1, 0)], indexing="ij")
    P = np.stack([X, Y, Z], axis=-1)
    inside = np.all((P > np.array(lo)) & (P < np.array(hi)), axis=-1)
    fill = 0.1 * sum(n)
    # band: the union of the triangles' index boxes (vertex / h.x, -1 / +2, clamped)
    band = np.zeros(sdf.shape, bool)
    for tri in t:
        f = v[tri] / h[0]
        b0 = [min(max(int(np.floor(f[:, a].min())) - 1, 0), n[a] - 1) for a in range(3)]
        b1 = [min(max(int(np.floor(f[:, a].max())) + 2, 0), n[a] - 1) for a in range(3)]
        band[b0[2]:b1[2] + 1, b0[1]:b1[1] + 1, b0[0]:b1[0] + 1] = True
    d = box_distance(P, lo, hi)
    assert np.allclose(np.abs(sdf[band]), d[band], rtol=0, atol=1e-12)
    assert np.all(np.abs(sdf[~band]) == fill)
    assert np.array_equal(sdf < 0, inside)


def test_mesh_to_sdf_band_ignores_origin():
    """The band's index box ignores the grid origin: with the origin moved by
    three cells the band sits three cells away from the mesh."""
    n = (14, 14, 14)
    h = (0.1, 0.1, 0.1)
    v, t = box((0.42, 0.41, 0.43), (0.61, 0.62, 0.63))
    sdf = orc.mesh_to_sdf(n, (-0.3, -0.3, -0.3), h, v, t)
    fill = 0.1 * 42
    near = np.abs(sdf) < fill
    idx = np.argwhere(near)  # (k, j, i)
    assert idx.min() >= 3 and idx.max() <= 8  # boxes [4-1, 6+2] from 0.42/0.1 .. 0.63/0.1
    # the mesh itself is at indices 7..9 of this grid: its nearest nodes are outside the band
    assert np.all(np.abs(sdf[7:10, 7:10, 7:10][~near[7:10, 7:10, 7:10]]) == fill)


def test_add_mesh_collider_running_minimum():
    n = (10, 9, 8)
    s = orc.Scene(n)
    h = tuple(1.0 / q for q in n)
    v1, t1 = icosphere((0.3, 0.4, 0.5), 0.2, levels=1)
    v2, t2 = box((0.55, 0.1, 0.2), (0.9, 0.5, 0.7))
    s.add_mesh_collider(n, (0.0, 0.0, 0.0), h, v1, t1)  # the scene's own node grid
    c1 = orc.mesh_to_sdf(n, (0.0, 0.0, 0.0), h, v1, t1)

    def sampled(c, o, hh):
        out = np.empty((n[2], n[1], n[0]))
        for k in range(n[2]):
            for j in range(n[1]):
                for i in range(n[0]):
                    out[k, j, i] = orc.sample_array(c, o, hh, (i * h[0], j * h[1], k * h[2]))
        return out

    # sampled at the nodes of its own grid: the node values up to the rounding
    # of (i h) / h in the trilinear weights
    want = sampled(c1, (0.0, 0.0, 0.0), h)
    assert np.array_equal(s.get(orc.COLLIDER_SDF), want)
    assert np.allclose(want, c1, rtol=1e-14, atol=1e-15)
    # a second collider on a finer grid of its own: sampled at the scene's nodes
    m = (17, 15, 13)
    o2 = (-0.05, -0.02, 0.0)
    h2 = (0.07, 0.07, 0.07)
    s.add_mesh_collider(m, o2, h2, v2, t2)
    want = np.minimum(want, sampled(orc.mesh_to_sdf(m, o2, h2, v2, t2), o2, h2))
    assert np.array_equal(s.get(orc.COLLIDER_SDF), want)
    with pytest.raises(ValueError):
        s.add_mesh_collider(n, (0.0, 0.0, 0.0), h, v1, t1 + 1000)
