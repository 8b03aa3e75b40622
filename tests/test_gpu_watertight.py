"""Watertightness on the GPU: the BVH8 traversal with the watertight triangle
test, through rn_trace_rays, on rays aimed exactly at the edges and vertices
of closed meshes (tests/_watertight.py). Every closest hit is a triangle of
the nearest crossing that the oracle's brute force over all triangles finds
(never a miss where that crossing exists), at the same distance bit for bit;
any-hit queries are occluded just past that crossing and not just before it.
The meshes sit on grid lines, so child boxes have faces in the planes the
axis-aligned rays run in."""
import numpy as np
import pytest

from _watertight import closed_heightfield, crossings, cube_rays, gridded_cube, heightfield_rays

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mesh", ["heightfield", "cube", "cube_shifted"])
def test_gpu_closest_and_any_at_edges_and_vertices(rn, oracle, mesh):
    if mesh == "heightfield":
        tris, h = closed_heightfield()
        o, d, tt, kinds = heightfield_rays(h)
    else:
        tris = gridded_cube(8)
        o, d, tt, kinds = cube_rays(8)
        if mesh == "cube_shifted":  # the same rays and cube moved by an exact offset
            off = np.array([1024.0, -512.0, 256.0], np.float32)
            tris = tris + off
            o = o + off
    hits = oracle.all_hits(tris, o, d, "watertight")
    sc = rn.Scene.from_arrays(tris, np.full((len(tris), 3), 0.5, np.float32), None)
    info = sc.build(8)
    assert info["leaf_tris"] == len(tris)
    prim, t, _, _ = sc.trace_rays(o, d)
    first_t = np.full(len(o), np.inf)
    misses = wrong = 0
    for i in range(len(o)):
        if len(hits[i][0]) == 0:  # touches a ridge only, and leaves the block
            assert prim[i] == -1, (i, kinds[i])
            continue
        t0, near = crossings(*hits[i])[0]
        first_t[i] = t0
        if prim[i] < 0:
            misses += 1
        elif prim[i] not in near or t[i] != hits[i][1][list(hits[i][0]).index(prim[i])]:
            wrong += 1
    print(f"{mesh}: {len(o)} rays, misses {misses}, other hits {wrong}")
    assert misses == 0 and wrong == 0
    has = np.isfinite(first_t)
    occ_after = sc.trace_rays(o[has], d[has], tmax=(first_t[has] * (1 + 1e-4)).astype(np.float32),
                              query=rn.QUERY_ANY)
    occ_before = sc.trace_rays(o[has], d[has], tmax=(first_t[has] * (1 - 1e-4)).astype(np.float32),
                               query=rn.QUERY_ANY)
    assert occ_after.all() and not occ_before.any()
    sc.close()
