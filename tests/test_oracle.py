"""CPU tests that pin the oracle: the reference's own known-answer tests, golden
vectors from the compiled reference, and closed forms of the physical
estimator. None of these needs a GPU."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from conftest import ASSETS, GOLDEN

# ----------------------------------------------------------------- Philox --


def test_philox_known_answers(oracle):
    """Random123 known-answer vectors for philox4x32-10."""
    ctr = np.array([[0, 0, 0, 0], [0xFFFFFFFF] * 4, [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344]], np.uint32)
    key = np.array([[0, 0], [0xFFFFFFFF] * 2, [0xA4093822, 0x299F31D0]], np.uint32)
    exp = np.array([[0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8],
                    [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD],
                    [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]], np.uint32)
    assert np.array_equal(oracle.philox(ctr, key), exp)


def test_philox_uniformity(oracle):
    rng = np.random.default_rng(0)
    ctr = np.zeros((20000, 4), np.uint32)
    ctr[:, 0] = np.arange(20000)
    key = np.tile(rng.integers(0, 2 ** 32, 2, dtype=np.uint64).astype(np.uint32), (20000, 1))
    u = (oracle.philox(ctr, key) >> 8).astype(np.float64) / 2 ** 24
    assert abs(u.mean() - 0.5) < 0.005
    assert abs(u.var() - 1 / 12) < 0.002
    h, _ = np.histogram(u, bins=16, range=(0, 1))
    assert (abs(h - u.size / 16) < 6 * math.sqrt(u.size / 16)).all()


# ------------------------------------------------ reference triangle test --

def tri(o, d, v0, v1, v2, oracle):
    t, u, v = C.c_float(), C.c_float(), C.c_float()
    a = [np.asarray(x, np.float32) for x in (o, d, v0, v1, v2)]
    h = oracle.lib().or_compat_intersect_tri(*a, C.byref(t), C.byref(u), C.byref(v))
    return h, t.value, u.value, v.value


def test_triangle_closed_forms(oracle):
    """tests/tests.cpp:150-165: distance-2 hit, miss, nearer of two wins."""
    v0, v1, v2 = (-1, -1, -2), (1, -1, -2), (0, 1, -2)
    h, t, u, v = tri((0, 0, 0), (0, 0, -1), v0, v1, v2, oracle)
    assert h and abs(t - 2.0) < 1e-4
    assert not tri((0, 0, 0), (0, 0, 1), v0, v1, v2, oracle)[0]
    verts = np.array([[v0, v1, v2], [(-1, -1, -1), (1, -1, -1), (0, 1, -1)]], np.float32)
    tt, uu, vv = C.c_float(), C.c_float(), C.c_float()
    p = oracle.lib().or_compat_closest_brute(verts.reshape(-1), 2, np.zeros(3, np.float32),
                                             np.array([0, 0, -1], np.float32), -1, C.byref(tt), C.byref(uu),
                                             C.byref(vv))
    assert p == 1 and abs(tt.value - 1.0) < 1e-4


def test_triangle_cramer_agreement(oracle):
    """tests/tests.cpp:324-329: the inverse solve equals Cramer's rule."""
    o, d = np.array([0.1, 0, 0]), np.array([0, 0, -1.0])
    v0, v1, v2 = np.array([-1, -1, -2.0]), np.array([1, -1, -2.0]), np.array([0, 1, -2.0])
    h, t, u, v = tri(o, d, v0, v1, v2, oracle)
    e0, e1, b = v1 - v0, v2 - v0, o - v0
    det = np.linalg.det(np.column_stack([-d, e0, e1]))
    tc = np.linalg.det(np.column_stack([b, e0, e1])) / det
    uc = np.linalg.det(np.column_stack([-d, b, e1])) / det
    vc = np.linalg.det(np.column_stack([-d, e0, b])) / det
    assert h and abs(t - tc) < 1e-4 and abs(u - uc) < 1e-4 and abs(v - vc) < 1e-4


def test_triangle_parallel_reject(oracle):
    assert not tri((0, 0, 0), (1, 0, 0), (-1, -1, -2), (1, -1, -2), (0, 1, -2), oracle)[0]


def test_compat_camera_bit_exact_vs_reference(oracle):
    """Oracle rn_render camera rays equal the reference Camera::primaryRay bits."""
    for asset in ("cornell", "quad"):
        g = np.load(os.path.join(GOLDEN, f"primary_{asset}.npz"))
        W, H = int(g["width"]), int(g["height"])
        o = np.empty(3, np.float32)
        d = np.empty(3, np.float32)
        for k, i in enumerate(g["sub"]):
            oracle.lib().or_compat_primary_ray(int(i % W), int(i // W), W, H, float(g["cam_z"]), 0.0, 0.0, o, d)
            assert np.array_equal(o, g["o_sub"][k]) and np.array_equal(d, g["d_sub"][k])


def test_camera_projection_closed_form(oracle):
    """tests/tests.cpp:113-130: the image centre looks straight at the origin."""
    o = np.empty(3, np.float32)
    d = np.empty(3, np.float32)
    # image-plane offset x + 0.5 + jx - W/2 = 0 for x = 159, jx = 0.5
    oracle.lib().or_compat_primary_ray(159, 119, 320, 240, 4.0, 0.5, 0.5, o, d)
    assert np.array_equal(o, [0, 0, 4]) and np.array_equal(d, [0, 0, -1])
    oracle.lib().or_compat_primary_ray(319, 119, 320, 240, 4.0, 0.0, 0.5, o, d)
    assert d[0] > 0 and d[1] == 0


@pytest.mark.parametrize("asset", ["cornell", "quad", "sphere"])
def test_compat_bvh_vs_reference_goldens(oracle, asset):
    """Oracle median-split BVH + glm-order test against BVH::intersect and
    getClosestIntersection of the compiled reference: ids equal outside the
    edge band, (t,u,v) bit-identical whenever ids agree."""
    ob = np.load(os.path.join(GOLDEN, f"obj_{asset}.npz"))
    g = np.load(os.path.join(GOLDEN, f"rays_{asset}.npz"))
    prim, t, u, v = oracle.compat_bvh_intersect(ob["verts"], g["o"], g["d"])
    same = prim == g["prim"]
    band = np.minimum(np.minimum(g["u"], g["v"]), 1 - g["u"] - g["v"]) < 1e-6
    assert (~same & ~band).sum() == 0
    h = same & (prim >= 0)
    assert np.array_equal(t[h], g["t"][h]) and np.array_equal(u[h], g["u"][h]) and np.array_equal(v[h], g["v"][h])
    bp, bt, bu, bv = oracle.compat_brute(ob["verts"], g["o"], g["d"])
    assert np.array_equal(bp, g["brute_prim"])
    hb = bp >= 0
    assert np.array_equal(bt[hb], g["brute_t"][hb])
    # BVH::occluded: the restated traversal and brute force over every
    # triangle both give the reference's answer on every ray
    import ctypes as C
    L = oracle.lib()
    v = np.ascontiguousarray(ob["verts"], np.float32)
    h = L.or_compat_bvh_build(v, v.size // 9)
    o3, d3 = g["o"].reshape(-1, 3), g["d"].reshape(-1, 3)
    occ = np.array([L.or_compat_bvh_occluded(h, np.ascontiguousarray(o3[i]), np.ascontiguousarray(d3[i]),
                                             float(g["maxd"][i]), -1) for i in range(len(o3))])
    L.or_compat_bvh_free(h)
    assert np.array_equal(occ, g["occluded"])
    brute = np.array([int((ts < g["maxd"][i]).any()) for i, (_, ts) in
                      enumerate(oracle.all_hits(v, o3, d3, "compat", cap=1 << 12))])
    assert np.array_equal(brute, g["occluded"])


def test_compat_primary_vs_reference_goldens(oracle):
    for asset in ("cornell", "quad", "sphere"):
        ob = np.load(os.path.join(GOLDEN, f"obj_{asset}.npz"))
        g = np.load(os.path.join(GOLDEN, f"primary_{asset}.npz"))
        W, H = int(g["width"]), int(g["height"])
        o, d = oracle.compat_primary_rays(W, H, float(g["cam_z"]))
        prim, t, u, v = oracle.compat_bvh_intersect(ob["verts"], o, d)
        assert np.array_equal(prim, g["prim"])
        h = prim >= 0
        assert np.array_equal(t[h], g["t"][h]) and np.array_equal(u[h], g["u"][h])


def test_bvh_fan_matches_brute_force(oracle):
    """tests/tests.cpp:174-197: a 400-ray fan over the Cornell asset, BVH ==
    brute force (same index, |dt| <= 1e-3)."""
    ob = np.load(os.path.join(GOLDEN, "obj_cornell.npz"))
    a = (np.arange(400) % 20) / 20.0 - 0.5
    b = (np.arange(400) // 20) / 20.0 - 0.5
    d = np.stack([a, b, -np.ones(400)], 1)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    o = np.tile(np.array([0, 0, 4], np.float32), (400, 1))
    p1, t1, _, _ = oracle.compat_bvh_intersect(ob["verts"], o, d)
    p2, t2, _, _ = oracle.compat_brute(ob["verts"], o, d)
    assert (p1 >= 0).sum() > 0
    h = p1 >= 0
    assert np.array_equal(p1, p2) and np.all(np.abs(t1[h] - t2[h]) <= 1e-3)


def test_fractal_noise_vs_reference(oracle):
    g = np.load(os.path.join(GOLDEN, "noise.npz"))
    out = np.array([oracle.lib().or_fractal_noise(np.ascontiguousarray(p), 4) for p in g["p"]], np.float32)
    assert np.array_equal(out, g["fbm4"])


def test_texel_closed_form(oracle):
    """Nearest-texel selection (src/Renderer.cpp:584-591): corners and clamp."""
    tp = np.array([0, 0, 1, 0, 1, 1], np.float32)  # vt of textured-quad.obj face 1
    L = oracle.lib()
    assert L.or_compat_texel(0.0, 0.0, tp, 480, 395) == 394 * 480 + 0      # (u,v)=(0,0): texcoord (0,0)
    assert L.or_compat_texel(1.0, 0.0, tp, 480, 395) == 394 * 480 + 479    # vertex 1: texcoord (1,0)
    assert L.or_compat_texel(0.0, 1.0, tp, 480, 395) == 0 * 480 + 479      # vertex 2: texcoord (1,1)


# ------------------------------------------------------------------- LBVH --

def _compact(x):
    r = 0
    for i in range(10):
        r |= ((x >> (3 * i)) & 1) << i
    return r


def test_morton_known_values(oracle):
    L = oracle.lib()
    assert L.or_morton3(1, 0, 0) == 1 and L.or_morton3(0, 1, 0) == 2 and L.or_morton3(0, 0, 1) == 4
    assert L.or_morton3(1023, 1023, 1023) == 0x3FFFFFFF
    c = L.or_morton3(123, 456, 789)  # tests/tests.cpp:430-431 round trip
    assert (_compact(c), _compact(c >> 1), _compact(c >> 2)) == (123, 456, 789)


def test_morton_codes_quantisation(oracle):
    v = np.zeros((3, 3, 3), np.float32)
    v[0] = [[0, 0, 0]] * 3
    v[1] = [[1, 2, 4]] * 3
    v[2] = [[0.5, 1, 2]] * 3
    codes = np.empty(3, np.uint32)
    oracle.lib().or_morton_codes(v.reshape(-1), 3, codes)
    assert codes[0] == 0 and codes[1] == 0x3FFFFFFF
    assert codes[2] == oracle.lib().or_morton3(512, 512, 512)


def _check_radix_tree(codes, left, right):
    n = len(codes)
    keys = [(int(codes[i]) << 32) | i for i in range(n)]
    seen_leaves = set()
    stack = [(0, 0, n - 1)]
    while stack:
        node, lo, hi = stack.pop()
        lc, rc = int(left[node]), int(right[node])
        common = 64 - (keys[lo] ^ keys[hi]).bit_length()
        split = max(i for i in range(lo, hi) if 64 - (keys[lo] ^ keys[i]).bit_length() > common)
        for c, a, b in ((lc, lo, split), (rc, split + 1, hi)):
            if c < 0:
                assert a == b == ~c
                seen_leaves.add(~c)
            else:
                assert a < b
                stack.append((c, a, b))
    assert seen_leaves == set(range(n))


@pytest.mark.parametrize("n", [2, 3, 17, 500])
def test_radix_tree_structure(oracle, n):
    rng = np.random.default_rng(n)
    codes = np.sort(rng.integers(0, 64, n).astype(np.uint32))  # heavy duplication
    left = np.empty(n - 1, np.int32)
    right = np.empty(n - 1, np.int32)
    oracle.lib().or_radix_tree(codes, n, left, right)
    _check_radix_tree(codes, left, right)


def _boxes_row(xs):
    """Triangles in sequence order, the k-th with box [x, x+1] x [0, 1] x [0, 1]:
    a merged box spanning dx along x has half-area 2 dx + 1."""
    return np.array([[[x, 0, 0], [x + 1, 1, 0], [x, 0, 1]] for x in xs], np.float32)


def _ploc(oracle, v, order=None):
    n = len(v)
    order = np.arange(n, dtype=np.int32) if order is None else order
    left = np.empty(max(n - 1, 1), np.int32)
    right = np.empty(max(n - 1, 1), np.int32)
    rounds = oracle.lib().or_ploc_tree(np.ascontiguousarray(v.reshape(-1)), order, n, left, right)
    return rounds, left[: n - 1].tolist(), right[: n - 1].tolist()


# Hand-worked trees: (x positions, rounds, left, right). Ids count down from
# n - 2 in merge order; the smaller position of a pair is the left child.
PLOC_CASES = [
    ([0, 5], 1, [~0], [~1]),
    # two separate pairs, one round each level
    ([0, 1.5, 10, 11.5], 2, [2, ~2, ~0], [1, ~3, ~1]),
    # gaps growing to the right: one merge a round, a caterpillar from the left
    ([0, 2, 5, 11, 24], 4, [1, 2, 3, ~0], [~4, ~3, ~2, ~1]),
    # its mirror: the caterpillar grows from the right, owners on the left
    ([0, 13, 19, 22, 24], 4, [~0, ~1, ~2, ~3], [1, 2, 3, ~4]),
    # equal gaps: on equal areas the pairs at even positions win, so every
    # other pair merges (ordering by position alone would merge one a round)
    ([0, 2, 4, 6, 8, 10], 3, [1, 4, ~4, ~2, ~0], [2, 3, ~5, ~3, ~1]),
    # duplicates merge first
    ([0, 0, 10], 2, [1, ~0], [~2, ~1]),
]


@pytest.mark.parametrize("xs,rounds,left,right", PLOC_CASES)
def test_ploc_tree_known(oracle, xs, rounds, left, right):
    assert _ploc(oracle, _boxes_row(xs)) == (rounds, left, right)


def _check_ploc_tree(n, left, right):
    """Every internal id 0..n-2 and every leaf is a child exactly once (the
    root 0 excepted), and children carry larger ids than their parent."""
    internal = [c for c in left + right if c >= 0]
    leaves = [~c for c in left + right if c < 0]
    assert sorted(internal) == list(range(1, n - 1))
    assert sorted(leaves) == list(range(n))
    for p in range(n - 1):
        for c in (left[p], right[p]):
            assert c < 0 or c > p


@pytest.mark.parametrize("n", [2, 3, 200, 3000])
def test_ploc_tree_structure_and_first_round(oracle, n):
    rng = np.random.default_rng(n)
    v = (rng.uniform(-5, 5, (n, 1, 3)) + rng.normal(scale=0.3, size=(n, 3, 3))).astype(np.float32)
    codes = np.empty(n, np.uint32)
    oracle.lib().or_morton_codes(np.ascontiguousarray(v.reshape(-1)), n, codes)
    order = np.argsort(codes, kind="stable").astype(np.int32)
    rounds, left, right = _ploc(oracle, v, order)
    assert rounds >= 1
    _check_ploc_tree(n, left, right)
    # brute force of the first round: mutual nearest neighbours within 8
    # positions by merged half-area (float64; random data has no ties)
    lo = v[order].min(axis=1).astype(np.float64)
    hi = v[order].max(axis=1).astype(np.float64)
    nn = []
    for i in range(n):
        js = [j for j in range(max(0, i - 8), min(n, i + 9)) if j != i]
        d = [np.maximum(hi[i], hi[j]) - np.minimum(lo[i], lo[j]) for j in js]
        nn.append(js[int(np.argmin([e[0] * e[1] + e[1] * e[2] + e[2] * e[0] for e in d]))])
    pairs = [(i, nn[i]) for i in range(n) if nn[nn[i]] == i and i < nn[i]]
    first = [(left[n - 2 - k], right[n - 2 - k]) for k in range(len(pairs))]
    assert first == [(~i, ~j) for i, j in pairs]


def test_ploc_rounds_on_a_strip(oracle):
    """A strip of 4000 identical quads (8000 triangles) in Morton order: the
    equal areas along it must not serialise the rounds."""
    q = np.arange(4000, dtype=np.float32)
    a = np.stack([q, 0 * q, 0 * q], 1)
    b, c, d = a + [1, 0, 0], a + [1, 1, 0], a + [0, 1, 0]
    v = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)]).astype(np.float32)
    n = len(v)
    codes = np.empty(n, np.uint32)
    oracle.lib().or_morton_codes(np.ascontiguousarray(v.reshape(-1)), n, codes)
    rounds, left, right = _ploc(oracle, v, np.argsort(codes, kind="stable").astype(np.int32))
    _check_ploc_tree(n, left, right)
    assert 0 < rounds <= 40, rounds


# ---------------------------------------------------- physical estimator --

def sphere_inside(levels=3):
    from rednoise_amd import scenes
    v, f = scenes.geodesic_sphere(levels)
    return v[f].astype(np.float32)


def test_white_furnace_closed_form(oracle):
    """Closed sphere, every triangle emitting Le with albedo a, camera inside:
    the estimator's mean is Le (1 - a^(D+1)) / (1 - a) for depth D, with and
    without Russian roulette."""
    tv = sphere_inside(3)
    n = len(tv)
    a, le = 0.5, 1.0
    alb = np.full((n, 3), a, np.float32)
    em = np.full((n, 3), le, np.float32)
    ps = oracle.PhysScene(tv, alb, em, "f64")
    W = H = 24
    cam = oracle.camera_basis((0.1, 0.05, -0.02), (1, 0.3, 0.2), (0, 1, 0), 60.0, W, H)
    for D, rr in ((4, 100), (12, 2)):
        p = oracle.params(W, H, 256, D, rr, 0.0, 5, 0, 1e-5)
        img, cnt = ps.render(cam, p)
        exp = le * (1 - a ** (D + 1)) / (1 - a)
        vals = img.reshape(-1, 3)[:, 0]
        se = vals.std() / math.sqrt(len(vals))
        assert abs(vals.mean() - exp) < max(4 * se, 1e-3 * exp), (D, vals.mean(), exp, se)


def _polygon_irradiance(p, n, poly):
    """Irradiance factor of a unit-radiance polygon at point p with normal n
    (Lambert's formula): 1/2 sum theta_i n.(g_i x g_{i+1})/|g_i x g_{i+1}|."""
    g = [(q - p) / np.linalg.norm(q - p) for q in poly]
    s = 0.0
    for i in range(len(g)):
        a, b = g[i], g[(i + 1) % len(g)]
        c = np.cross(a, b)
        th = math.acos(max(-1.0, min(1.0, float(a @ b))))
        s += th * float(n @ c) / np.linalg.norm(c)
    return 0.5 * abs(s)


def test_one_bounce_polygon_irradiance(oracle):
    """Floor lit by a square light, depth 1 (primary + NEE): pixel value =
    albedo / pi * Le * E_polygon at the visible floor point."""
    s, h, a, le = 1.0, 2.0, 0.6, 3.0
    floor = np.array([[[-50, 0, -50], [50, 0, -50], [50, 0, 50]], [[-50, 0, -50], [50, 0, 50], [-50, 0, 50]]])
    light = np.array([[[-s, h, -s], [s, h, -s], [s, h, s]], [[-s, h, -s], [s, h, s], [-s, h, s]]])
    tv = np.concatenate([floor, light]).astype(np.float32)
    alb = np.array([[a] * 3] * 2 + [[0.0] * 3] * 2, np.float32)
    em = np.array([[0.0] * 3] * 2 + [[le] * 3] * 2, np.float32)
    ps = oracle.PhysScene(tv, alb, em, "f64")
    W = H = 5
    eye = np.array([3.0, 1.0, 0.3])
    tgt = np.array([0.4, 0.0, 0.2])
    cam = oracle.camera_basis(eye, tgt, (0, 1, 0), 0.5, W, H)
    p = oracle.params(W, H, 20000, 1, 100, 0.0, 1, 0, 1e-6)
    img, _ = ps.render(cam, p)
    centre = img[H // 2, W // 2, 0]
    d = tgt - eye
    d /= np.linalg.norm(d)
    hit = eye + d * (-eye[1] / d[1])
    poly = [np.array(q, float) for q in light[0][[0, 1, 2]].tolist()] + [np.array(light[1][2], float)]
    exp = a / math.pi * le * _polygon_irradiance(hit, np.array([0, 1.0, 0]), poly)
    assert abs(centre - exp) < 0.01 * exp, (centre, exp)


def test_fp32_oracle_tracks_fp64(oracle):
    from rednoise_amd import scenes
    sd = scenes.cornell_box()
    W, H = 48, 36
    cam = oracle.camera_basis(sd.camera["eye"], sd.camera["target"], sd.camera["up"], 40.0, W, H)
    p = oracle.params(W, H, 8, 5, 3, 0.0, 0, 0, 0.05)
    a32, _ = oracle.PhysScene(sd.verts, sd.albedo, sd.emission, "f32").render(cam, p)
    a64, _ = oracle.PhysScene(sd.verts, sd.albedo, sd.emission, "f64").render(cam, p)
    close = np.abs(a32 - a64) <= 1e-3 * np.maximum(1.0, np.abs(a64))
    assert close.mean() > 0.97


def test_oracle_bvh_matches_brute(oracle):
    rng = np.random.default_rng(4)
    c = rng.uniform(-5, 5, size=(800, 1, 3))
    tv = (c + rng.normal(scale=0.4, size=(800, 3, 3))).astype(np.float32)
    ps = oracle.PhysScene(tv, np.full((800, 3), 0.5, np.float32), None, "f64")
    L = oracle.lib()
    for i in range(500):
        o = rng.uniform(-7, 7, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        r = []
        for brute in (0, 1):
            t, u, v = C.c_double(), C.c_double(), C.c_double()
            r.append(L.or_phys_closest_f64(ps.h, o, d, math.inf, brute, C.byref(t), C.byref(u), C.byref(v)))
        assert r[0] == r[1]


def test_adjoint_matches_finite_differences(oracle):
    """FP64 path-replay adjoint against central differences (h = 1e-3) of the
    FP64 forward estimator with common random numbers (no roulette, so the
    estimator is a polynomial in albedo and emission). Parameters are float32,
    so the difference quotient divides by the step the float32 values took."""
    from rednoise_amd import scenes
    sd = scenes.cornell_box()
    W, H, spp, D = 12, 9, 4, 3
    cam = oracle.camera_basis(sd.camera["eye"], sd.camera["target"], sd.camera["up"], 40.0, W, H)
    p = oracle.params(W, H, spp, D, 100, 0.0, 3, 0, 0.05)
    rng = np.random.default_rng(0)
    wgt = rng.normal(size=(H, W, 3)).astype(np.float32)
    ps = oracle.PhysScene(sd.verts, sd.albedo, sd.emission, "f64")
    ga, ge = ps.adjoint(cam, p, wgt)

    def loss(alb, em):
        s = oracle.PhysScene(sd.verts, alb, em, "f64")
        img, _ = s.render(cam, p)
        return float((img * wgt).sum())

    h = 1e-3
    checked = 0
    for tri in (0, 4, 6, 12, 20, 30):
        for ch in range(3):
            ap = sd.albedo.copy()
            am = sd.albedo.copy()
            ap[tri, ch] += h
            am[tri, ch] -= h
            step = float(ap[tri, ch]) - float(am[tri, ch])  # the float32 step actually taken
            fd = (loss(ap, sd.emission) - loss(am, sd.emission)) / step
            assert abs(fd - ga[tri, ch]) <= 1e-4 * max(1.0, abs(fd)), (tri, ch, fd, ga[tri, ch])
            checked += 1
    for tri in (30, 31):
        for ch in range(3):
            ep = sd.emission.copy()
            em_ = sd.emission.copy()
            ep[tri, ch] += h
            em_[tri, ch] -= h
            step = float(ep[tri, ch]) - float(em_[tri, ch])
            fd = (loss(sd.albedo, ep) - loss(sd.albedo, em_)) / step
            assert abs(fd - ge[tri, ch]) <= 1e-4 * max(1.0, abs(fd)), (tri, ch, fd, ge[tri, ch])
    assert checked == 18


def test_rmse_fixtures_match_their_workloads(oracle):
    """The FP64 65,536-spp pixel subsets that bench.py scores every config
    against carry the parameters of tests/golden/make_oracle_rmse.py, hold
    finite non-negative radiance, and config 0's agree with a fresh oracle
    render of 6 of their pixels at 4,096 spp on another seed (within 5
    standard errors of that render)."""
    import sys
    sys.path.insert(0, GOLDEN)
    import make_oracle_rmse as M
    from rednoise_amd import scenes
    for name, (_, W, H, D, rr, eps, block) in M.WORKLOADS.items():
        path = os.path.join(GOLDEN, f"oracle64k_{name}.npz")
        if not os.path.isfile(path):
            pytest.skip(f"{path} not generated")
        f = np.load(path)
        assert int(f["spp"]) == M.SPP and int(f["block"]) == block
        assert (int(f["width"]), int(f["height"]), int(f["max_depth"]), int(f["rr_depth"])) == (W, H, D, rr)
        assert abs(float(f["ray_eps"]) - eps) < 1e-9
        assert np.array_equal(f["pixels"], M.pixel_subset(W, H, block, seed=17))
        assert np.isfinite(f["values"]).all() and (f["values"] >= 0).all()
    path2 = os.path.join(GOLDEN, "oracle64k_config2.npz")
    if not os.path.isfile(path2):
        pytest.skip(f"{path2} not generated")
    f2 = np.load(path2)
    specs = scenes.config2_batch(64, seed=2)
    for k in M.CONFIG2_SCENES:
        s = specs[k]
        sel = f2["scene"] == k
        assert np.array_equal(f2["pixels"][sel], M.pixel_subset(s.width, s.height, M.CONFIG2_BLOCK, seed=17))
    # config 0 against a fresh render of a few of its pixels
    f0 = np.load(os.path.join(GOLDEN, "oracle64k_config0.npz"))
    sd = scenes.cornell_box()
    cam = oracle.camera_basis(sd.camera["eye"], sd.camera["target"], sd.camera["up"], sd.camera["vfov_deg"], 320, 240)
    ps = oracle.PhysScene(sd.verts, sd.albedo, sd.emission, "f64")
    pick = np.arange(0, len(f0["pixels"]), len(f0["pixels"]) // 6)[:6]
    spp = 4096
    runs = []
    for seed in (101, 202, 303, 404):
        p = oracle.params(320, 240, spp // 4, 5, 3, 0.0, seed, 0, 0.05)
        out, _ = ps.render(cam, p, pixels=f0["pixels"][pick])
        runs.append(out)
    runs = np.array(runs)
    mean = runs.mean(0)
    se = runs.std(0, ddof=1) / 2.0 + 1e-3 * np.abs(mean) + 1e-6
    z = np.abs(mean - f0["values"][pick]) / se
    assert (z < 5).all(), z


def test_adjoint_is_linear_in_the_pixel_weights(oracle):
    """The adjoint is linear in dL/dI: the gradient of a weight image equals
    the sum of the gradients of two weight images with disjoint pixel support
    (pixels of zero weight contribute nothing and are not traced, so this
    also checks that no pixel with a non-zero weight in any channel is
    skipped: some weights below are zero in one or two channels only)."""
    from rednoise_amd import scenes
    sd = scenes.cornell_box()
    W, H, spp, D = 16, 12, 4, 4
    cam = oracle.camera_basis(sd.camera["eye"], sd.camera["target"], sd.camera["up"], 40.0, W, H)
    p = oracle.params(W, H, spp, D, 100, 0.0, 5, 0, 0.05)
    rng = np.random.default_rng(9)
    wgt = rng.normal(size=(H, W, 3)).astype(np.float32)
    wgt[rng.uniform(size=(H, W, 3)) < 0.3] = 0.0  # single zero channels
    part = rng.uniform(size=(H, W)) < 0.5
    w1 = np.where(part[..., None], wgt, 0).astype(np.float32)
    w2 = np.where(part[..., None], 0, wgt).astype(np.float32)
    ps = oracle.PhysScene(sd.verts, sd.albedo, sd.emission, "f64")
    ga, ge = ps.adjoint(cam, p, wgt)
    a1, e1 = ps.adjoint(cam, p, w1)
    a2, e2 = ps.adjoint(cam, p, w2)
    assert np.abs(ga).max() > 0 and np.abs(ge).max() > 0
    assert np.allclose(ga, a1 + a2, rtol=1e-12, atol=1e-15)
    assert np.allclose(ge, e1 + e2, rtol=1e-12, atol=1e-15)
    z, zz = ps.adjoint(cam, p, np.zeros_like(wgt))
    assert not z.any() and not zz.any()


def test_mt19937_seeds_past_the_int_overflow(oracle):
    """renderPathTraced seeds each pixel's std::mt19937 with the int
    (y*W+x)*9781+1 (src/Renderer.cpp:1139), which overflows past pixel
    219,553. The compiled reference's image at 1024x432, 1 spp holds box
    pixels on both sides of that point (tests/golden/modeR_overflow.npz);
    the restatement, which seeds with the same product taken modulo 2^32,
    gives it pixel for pixel on both sides."""
    import ctypes as C
    g = np.load(os.path.join(GOLDEN, "modeR_overflow.npz"))
    W, H = int(g["width"]), int(g["height"])
    (y0, y1), (x0, x1) = g["rows"], g["cols"]
    ob = np.load(os.path.join(GOLDEN, "obj_cornell.npz"))
    from rednoise_amd import Scene
    sc = Scene.load_obj(os.path.join(ASSETS, "cornell-box.obj"), float(g["scale"]))
    assert np.array_equal(sc.triangles()[0].reshape(-1), ob["verts"])
    cs = oracle.CompatArrays.from_scene(sc)
    out = np.zeros(W * H * 4, np.uint8)
    oracle.lib().or_compat_render_mt(C.byref(cs.s), W, H, float(g["cam_z"]), int(g["spp"]), int(y0), int(y1), out)
    band = out.reshape(H, W, 4)[y0:y1, x0:x1]
    ys, xs = np.mgrid[y0:y1, x0:x1]
    past = ys * W + xs > 219_553
    same = (band == g["rgba"]).all(axis=2)
    print("identical before / past the overflow:", same[~past].mean(), same[past].mean())
    assert past.sum() > 10_000 and (~past).sum() > 10_000
    assert same.all()
