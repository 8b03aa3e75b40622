"""Watertightness of the FP32 ray-triangle test (Woop, Benthin, Wald 2013,
with a tie rule for edge functions that are exactly zero), on the CPU oracle:
rays through the edges and vertices of closed meshes, with exact arithmetic
(tests/_watertight.py), cross the surface with exactly one triangle accepted
at each crossing -- no holes, no double hits -- by brute force over every
triangle. The reference's own test (src/Geometry.cpp:11-25: matrix inverse,
|det| < 1e-8 reject, u, v >= 0 and u + v <= 1 on rounded values) leaves holes
or double hits on the same rays, which shows that the check discriminates."""
import numpy as np
import pytest

from _watertight import check_closed, closed_heightfield, crossings, cube_rays, gridded_cube, heightfield_rays


def test_meshes_are_closed():
    tris, h = closed_heightfield()
    assert check_closed(tris) == pytest.approx(float(sum(
        (h[i, j] + h[i + 1, j] + h[i + 1, j + 1] + h[i, j + 1]) / 4.0 for i in range(h.shape[0] - 1)
        for j in range(h.shape[1] - 1))), rel=0.05)  # bilinear-cell volume, close to the triangulated one
    assert check_closed(gridded_cube(8)) == pytest.approx(512.0)


def _cases():
    hf, h = closed_heightfield()
    cube = gridded_cube(8)
    return [("heightfield", hf) + heightfield_rays(h), ("cube", cube) + cube_rays(8)]


def _check(name, tris, o, d, tt, kinds, hits):
    """Rays whose hits break watertightness. Each hit enters (+1: the ray
    against the outward normal) or leaves (-1) the solid; hits at one point
    of the ray form a group. Walking the groups in order, the count of solids
    the ray is inside must stay 0 or 1 and end at 0: a hole (a missed
    triangle) or a double hit (two triangles of one sheet at one point)
    breaks that. A ray that only touches the surface at a ridge meets it with
    one hit entering and one leaving there, or none; transversal crossings
    (every vertical ray through the heightfield, every ray from the cube's
    centre) have exactly one hit each."""
    t64 = tris.astype(np.float64)
    nrm = np.cross(t64[:, 1] - t64[:, 0], t64[:, 2] - t64[:, 0])  # exact: integer vertices
    bad = []
    for i, (prims, ts) in enumerate(hits):
        g = crossings(prims, ts)
        inside = 1 if name == "cube" else 0  # the cube's rays start at its centre
        ok = True
        for _, ps in g:
            s = [(-1 if float(np.dot(d[i].astype(np.float64), nrm[p])) > 0 else 1) for p in ps]
            inside += sum(s)
            ok = ok and inside in (0, 1)
        ok = ok and inside == 0
        if kinds[i].startswith("vertical"):
            ok = ok and len(g) == 2 and all(len(ps) == 1 for _, ps in g)   # top, then bottom
            ok = ok and abs(g[0][0] - tt[i]) <= 1e-5 * tt[i]
        if name == "cube":
            ok = ok and len(g) == 1 and len(g[0][1]) == 1 and abs(g[0][0] - 1.0) <= 1e-6
        if not ok:
            bad.append((kinds[i], tuple(o[i]), tuple(d[i]), [(round(t, 5), ps) for t, ps in g]))
    return bad


def test_watertight_one_triangle_per_crossing(oracle):
    total = 0
    for name, tris, o, d, tt, kinds in _cases():
        hits = oracle.all_hits(tris, o, d, "watertight")
        bad = _check(name, tris, o, d, tt, kinds, hits)
        assert not bad, (name, len(bad), bad[:5])
        # closest hit (BVH and brute force of the oracle) at the aimed crossing
        ps = oracle.PhysScene(tris, np.full((len(tris), 3), 0.5, np.float32), None, "f32")
        import ctypes as C
        L = oracle.lib()
        for i in range(0, len(o), 7):
            for brute in (0, 1):
                t, u, v = C.c_float(), C.c_float(), C.c_float()
                p = L.or_phys_closest_f32(ps.h, np.ascontiguousarray(o[i]), np.ascontiguousarray(d[i]),
                                          float("inf"), brute, C.byref(t), C.byref(u), C.byref(v))
                if len(hits[i][0]) == 0:  # touches a ridge (no hit) and leaves the block
                    assert p == -1, (name, i, brute, p)
                    continue
                # any triangle of the nearest crossing (a ridge touched there
                # holds one entering and one leaving at the same point)
                t0, near = crossings(*hits[i])[0]
                assert p in near, (name, i, brute, p, near)
                assert t.value == hits[i][1][list(hits[i][0]).index(p)], (name, i, brute, t.value)
        ps.close()
        total += len(o)
        kinds_seen = {k.split("-")[-1] for k in kinds}
        assert kinds_seen >= ({"vertex", "edge", "diagonal"} if name == "heightfield" else {"cube"})
    assert total > 1500


def test_reference_triangle_test_is_not_watertight(oracle):
    """The same rays through the reference's test: holes or double hits."""
    counts = {}
    for name, tris, o, d, tt, kinds in _cases():
        hits = oracle.all_hits(tris, o, d, "compat")
        counts[name] = len(_check(name, tris, o, d, tt, kinds, hits))
    print("rays failing with the reference's triangle test:", counts)
    assert sum(counts.values()) > 0, counts
