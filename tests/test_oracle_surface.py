"""Pins the oracle's MarchingCubesSolver::BuildSurface restatement
(src/3d/marching_cubes_solver.cpp) by geometry, not by retyping it:

  * a sphere SDF away from the walls: a closed, consistently oriented
    surface (every directed edge once, its reverse once), Euler
    characteristic 2, outward normals, vertices on the sphere to O(h^2) and
    the enclosed volume within the inscribed-polygon error;
  * a planar slab touching the walls: the cube part and the boundary caps
    close the slab exactly (planar SDF: linear interpolation is exact), so the
    enclosed volume is the slab between the outer node planes;
  * a random SDF: the surface with its caps is closed after merging the caps'
    own vertex copies (one vertex set per boundary plane, as the reference's
    separate vertex maps), with consistent orientation.
"""
from collections import Counter

import numpy as np

from oracle import oracle as orc


def _nodes(n, size=(1.0, 1.0, 1.0)):
    h = [size[a] / n[a] for a in range(3)]
    ax = [(np.arange(n[a]) + 0.5) * h[a] for a in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return X, Y, Z, h


def _merged_edges(v, t, tol=1e-9):
    """Directed edge counts after merging vertices closer than tol."""
    key = np.round(v / tol).astype(np.int64)
    _, ids = np.unique(key, axis=0, return_inverse=True)
    ids = ids.reshape(-1)
    E = Counter()
    for a, b, c in t:
        for p, q in ((a, b), (b, c), (c, a)):
            E[(ids[p], ids[q])] += 1
    return E, ids


def _closed(E):
    return all(k == 1 and E.get((q, p), 0) == 1 for (p, q), k in E.items())


def _volume(v, t):
    return float(np.einsum("ij,ij->i", v[t[:, 0]], np.cross(v[t[:, 1]], v[t[:, 2]])).sum() / 6.0)


def test_sphere_closed_oriented_accurate():
    n = (24, 22, 26)
    X, Y, Z, h = _nodes(n)
    c, r = np.array([0.5, 0.47, 0.52]), 0.3
    s = orc.Scene(n)
    s.set(orc.SDF, np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) - r)
    v, nrm, t = s.surface()
    E = Counter()
    for a, b, cc in t:
        for p, q in ((a, b), (b, cc), (cc, a)):
            E[(p, q)] += 1
    assert _closed(E)
    assert len(v) - len(E) // 2 + len(t) == 2
    d = np.sqrt(((v - c) ** 2).sum(1)) - r
    assert np.abs(d).max() < max(h) ** 2 / r
    radial = (v - c) / np.linalg.norm(v - c, axis=1)[:, None]
    assert (radial * nrm).sum(1).min() > 0.99
    vol = _volume(v, t)
    assert 0 < 4 / 3 * np.pi * r ** 3 - vol < 0.03 * 4 / 3 * np.pi * r ** 3


def test_planar_slab_with_caps_is_exact():
    n = (10, 12, 8)
    X, Y, Z, h = _nodes(n)
    level = 0.31
    s = orc.Scene(n)
    s.set(orc.SDF, Y - level)
    v, nrm, t = s.surface()
    E, _ = _merged_edges(v, t)
    assert _closed(E)
    lo = [0.5 * h[a] for a in range(3)]
    hi = [1.0 - 0.5 * h[a] for a in range(3)]
    want = (hi[0] - lo[0]) * (level - lo[1]) * (hi[2] - lo[2])
    assert abs(_volume(v, t) - want) < 1e-12
    top = np.abs(v[:, 1] - level) < 1e-12
    assert np.allclose(nrm[top & (np.abs(v[:, 0] - 0.5) < 0.3) & (np.abs(v[:, 2] - 0.5) < 0.3)], [0, 1, 0])


def test_random_field_closed_with_caps():
    n = (9, 7, 8)
    rng = np.random.default_rng(4)
    s = orc.Scene(n)
    s.set(orc.SDF, rng.uniform(-1, 1, (n[2], n[1], n[0])))
    v, nrm, t = s.surface()
    assert len(t) > 50
    E, _ = _merged_edges(v, t)
    assert _closed(E)
    assert _volume(v, t) > 0
