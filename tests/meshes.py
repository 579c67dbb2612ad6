"""Closed triangle meshes for the mesh-collider tests (outward, counter-
clockwise triangles): an axis-aligned box and a subdivided icosahedron."""
import numpy as np


def box(lo, hi):
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    v = np.array([[lo[0] if (c & 1) == 0 else hi[0], lo[1] if (c & 2) == 0 else hi[1],
                   lo[2] if (c & 4) == 0 else hi[2]] for c in range(8)])
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    t = []
    for a, b, c, d in quads:
        t += [(a, b, c), (a, c, d)]
    return v, np.array(t, dtype=np.int32)


def icosphere(center, radius, levels=2):
    g = (1 + 5 ** 0.5) / 2
    v = [(-1, g, 0), (1, g, 0), (-1, -g, 0), (1, -g, 0), (0, -1, g), (0, 1, g), (0, -1, -g), (0, 1, -g),
         (g, 0, -1), (g, 0, 1), (-g, 0, -1), (-g, 0, 1)]
    v = [np.array(p, float) / np.linalg.norm(p) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(levels):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[key] = len(v) - 1
            return cache[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    verts = np.asarray(center, float) + radius * np.array(v)
    return verts, np.array(f, dtype=np.int32)


def closest_point_reference_free(x, p0, p1, p2):
    """Closest point of a triangle by candidates: the plane projection when it
    falls inside, else the clamped projections on the three edges."""
    x, p0, p1, p2 = (np.asarray(a, float) for a in (x, p0, p1, p2))
    n = np.cross(p1 - p0, p2 - p0)
    cands = []
    if np.dot(n, n) > 0:
        q = x - np.dot(x - p0, n) / np.dot(n, n) * n
        # barycentric signs
        inside = all(np.dot(np.cross(b - a, q - a), n) >= 0 for a, b in ((p0, p1), (p1, p2), (p2, p0)))
        if inside:
            cands.append(q)
    for a, b in ((p0, p1), (p1, p2), (p2, p0)):
        d = b - a
        dd = np.dot(d, d)
        t = 0.0 if dd == 0 else min(max(np.dot(x - a, d) / dd, 0.0), 1.0)
        cands.append(a + t * d)
    dist = [np.linalg.norm(c - x) for c in cands]
    return cands[int(np.argmin(dist))], min(dist)


def box_distance(p, lo, hi):
    """Unsigned distance from p (..., 3) to the surface of the box [lo, hi]."""
    p = np.asarray(p, float)
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    c, e = (lo + hi) / 2, (hi - lo) / 2
    q = np.abs(p - c) - e
    outside = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
    inside = np.minimum(np.max(q, axis=-1), 0.0)
    return np.abs(outside + inside)
