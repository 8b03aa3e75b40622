"""Closed meshes and rays through their edges and vertices, for the
watertightness tests (tests/test_watertight.py on the CPU oracle,
tests/test_gpu_watertight.py on the GPU).

Every vertex has small integer coordinates and every ray direction has
power-of-two ratios to its largest component, so the ray-space shear of the
watertight test (Woop, Benthin, Wald 2013) and its 2-D edge functions are
computed without rounding: a ray aimed at an edge or a vertex meets it
exactly, and the edge functions there are exactly zero. A watertight test with
a consistent tie rule then accepts exactly one triangle at each crossing of
the surface: none would be a hole, two a double hit.
"""
from __future__ import annotations

import numpy as np


def closed_heightfield(n: int = 12, seed: int = 5):
    """A solid block [0, n]^2 x [0, h(x, y)]: a heightfield top with integer
    heights in [3, 9] (cell diagonals alternating direction), a flat bottom
    gridded the same way and four vertical side walls, all wound outward.
    Returns (tris (m, 3, 3) float32, heights (n+1, n+1))."""
    rng = np.random.default_rng(seed)
    h = rng.integers(3, 10, size=(n + 1, n + 1))
    tris = []

    def top(i, j):
        return (i, j, int(h[i, j]))

    def bot(i, j):
        return (i, j, 0)

    for i in range(n):
        for j in range(n):
            for f, flip in ((top, False), (bot, True)):
                p00, p10, p11, p01 = f(i, j), f(i + 1, j), f(i + 1, j + 1), f(i, j + 1)
                if (i + j) % 2 == 0:
                    cell = [(p00, p10, p11), (p00, p11, p01)]
                else:
                    cell = [(p00, p10, p01), (p10, p11, p01)]
                for a, b, c in cell:
                    tris.append((a, c, b) if flip else (a, b, c))
    # side walls: boundary of the grid walked counter-clockwise seen from +z
    ring = [(i, 0) for i in range(n)] + [(n, j) for j in range(n)] + \
           [(i, n) for i in range(n, 0, -1)] + [(0, j) for j in range(n, 0, -1)]
    for k in range(len(ring)):
        a, b = ring[k], ring[(k + 1) % len(ring)]
        ta, tb, ba, bb = top(*a), top(*b), bot(*a), bot(*b)
        tris.append((ba, bb, tb))
        tris.append((ba, tb, ta))
    return np.asarray(tris, np.float32), h


def gridded_cube(m: int = 8):
    """The cube [0, m]^3 with every face split into m x m cells (diagonals
    alternating), wound outward: 12 m^2 triangles."""
    tris = []
    for axis in range(3):
        for side in (0, m):
            u, v = (axis + 1) % 3, (axis + 2) % 3
            for i in range(m):
                for j in range(m):
                    def P(a, b):
                        p = [0, 0, 0]
                        p[axis], p[u], p[v] = side, a, b
                        return tuple(p)
                    p00, p10, p11, p01 = P(i, j), P(i + 1, j), P(i + 1, j + 1), P(i, j + 1)
                    cell = [(p00, p10, p11), (p00, p11, p01)] if (i + j) % 2 == 0 else \
                           [(p00, p10, p01), (p10, p11, p01)]
                    for a, b, c in cell:
                        # (u, v, axis) is right-handed: counter-clockwise in
                        # (u, v) faces +axis, outward on the side = m face
                        tris.append((a, b, c) if side == m else (a, c, b))
    return np.asarray(tris, np.float32)


def check_closed(tris: np.ndarray) -> float:
    """Asserts every directed edge appears once and its reverse once (a closed,
    consistently wound 2-manifold); returns the enclosed volume (> 0 when the
    winding faces outward)."""
    from collections import Counter
    e = Counter()
    for t in tris.astype(np.int64):
        for k in range(3):
            e[(tuple(t[k]), tuple(t[(k + 1) % 3]))] += 1
    for (a, b), c in e.items():
        assert c == 1 and e.get((b, a), 0) == 1, (a, b)
    t = tris.astype(np.float64)
    return float(np.einsum("ij,ij->i", t[:, 0], np.cross(t[:, 1], t[:, 2])).sum() / 6.0)


def heightfield_rays(h: np.ndarray):
    """Rays through vertices, edge midpoints, edge quarter points and cell
    diagonals of the heightfield top. Vertical rays (direction (0, 0, -1),
    origin at z = 20) cross the top and the bottom once each; slanted rays
    (directions with ratios 1/4 and 1/2 to their z component) pass through the
    target at t = 8 exactly. Returns (origins, dirs, target t, kind)."""
    n = h.shape[0] - 1
    targets = []
    for i in range(1, n):
        for j in range(1, n):
            targets.append((i, j, float(h[i, j]), "vertex"))
    for i in range(n):
        for j in range(1, n):
            for f in (0.5, 0.25):
                targets.append((i + f, j, float(h[i, j]) + f * float(h[i + 1, j] - h[i, j]), "edge"))
                targets.append((j, i + f, float(h[j, i]) + f * float(h[j, i + 1] - h[j, i]), "edge"))
    for i in range(n):
        for j in range(n):
            if (i + j) % 2 == 0:  # diagonal (i, j) - (i+1, j+1)
                targets.append((i + 0.5, j + 0.5, 0.5 * float(h[i, j] + h[i + 1, j + 1]), "diagonal"))
            else:                 # diagonal (i+1, j) - (i, j+1)
                targets.append((i + 0.5, j + 0.5, 0.5 * float(h[i + 1, j] + h[i, j + 1]), "diagonal"))
    o, d, tt, kind = [], [], [], []
    for x, y, z, k in targets:
        o.append((x, y, 20.0))
        d.append((0.0, 0.0, -1.0))
        tt.append(20.0 - z)
        kind.append("vertical-" + k)
    dirs = [(1.0, 2.0, -4.0), (-2.0, 1.0, -4.0), (1.0, -1.0, -2.0), (-1.0, -2.0, -4.0), (2.0, 2.0, -4.0)]
    for q, (x, y, z, k) in enumerate(targets):
        dd = np.array(dirs[q % len(dirs)])
        o.append(tuple(np.array((x, y, z)) - 8.0 * dd))
        d.append(tuple(dd))
        tt.append(8.0)
        kind.append("slanted-" + k)
    return (np.asarray(o, np.float32), np.asarray(d, np.float32), np.asarray(tt, np.float64), kind)


def cube_rays(m: int = 8):
    """Rays from the cube's centre through every grid vertex, grid-edge
    midpoint and face-diagonal midpoint of its surface (cube edges and corners
    included); each crosses the surface once, at t = 1."""
    c = np.full(3, m / 2.0)
    pts = set()
    for axis in range(3):
        for side in (0, m):
            u, v = (axis + 1) % 3, (axis + 2) % 3
            for a2 in range(0, 2 * m + 1):
                for b2 in range(0, 2 * m + 1):
                    a, b = a2 / 2.0, b2 / 2.0  # both halves odd: a cell centre, on its diagonal
                    p = [0.0, 0.0, 0.0]
                    p[axis], p[u], p[v] = float(side), a, b
                    pts.add(tuple(p))
    pts = sorted(pts)
    o = np.repeat(c[None].astype(np.float32), len(pts), 0)
    d = (np.asarray(pts) - c).astype(np.float32)
    return o, d, np.ones(len(pts)), ["cube"] * len(pts)


def crossings(prims, ts, sep=1e-3):
    """Groups hits of one ray by distance: [(t, [prims])] sorted by t."""
    order = np.argsort(ts, kind="stable")
    groups = []
    for k in order:
        if groups and abs(float(ts[k]) - groups[-1][0]) <= sep * max(1.0, abs(groups[-1][0])):
            groups[-1][1].append(int(prims[k]))
        else:
            groups.append((float(ts[k]), [int(prims[k])]))
    return groups
