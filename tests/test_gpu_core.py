"""GPU parity: Philox, radix sort, LBVH stages, BVH8 structure, ray queries,
the reference camera/triangle test (bit-exact against BVH::intersect goldens)
and the physical estimator against the CPU oracle."""
import os

import numpy as np
import pytest

from conftest import ASSETS, GOLDEN

pytestmark = pytest.mark.gpu


def soup(n, seed=0, dup=0):
    """Random triangle soup in a 10-unit cube, with `dup` duplicated triangles."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-5, 5, size=(n, 1, 3))
    v = (c + rng.normal(scale=0.3, size=(n, 3, 3))).astype(np.float32)
    if dup:
        v[-dup:] = v[:dup]
    return v


def test_device_is_gfx950(rn):
    assert rn.device_name().startswith("gfx950"), rn.device_name()


def test_philox_bitwise(rn, oracle):
    rng = np.random.default_rng(1)
    ctr = rng.integers(0, 2 ** 32, size=(8192, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 2 ** 32, size=(8192, 2), dtype=np.uint64).astype(np.uint32)
    ctr[0] = 0
    key[0] = 0
    g = rn.philox(ctr, key)
    assert np.array_equal(g, oracle.philox(ctr, key))
    assert g[0].tolist() == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]


@pytest.mark.parametrize("n", [1, 2, 7, 2047, 2048, 2049, 100_003])
def test_radix_sort_stable(rn, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << 12, size=n, dtype=np.uint64).astype(np.uint32)  # many duplicates
    keys[::5] = rng.integers(0, 1 << 30, size=len(keys[::5]), dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    k, v = rn.radix_sort_pairs(keys, vals, bits=30)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order])
    assert np.array_equal(v, vals[order])


@pytest.mark.parametrize("n,dup", [(2, 0), (3, 1), (1000, 0), (5000, 300)])
def test_lbvh_stages_match_oracle(rn, oracle, n, dup):
    v = soup(n, seed=n, dup=dup)
    codes, order, left, right = rn.lbvh_stages(v)
    oc = np.empty(n, np.uint32)
    oracle.lib().or_morton_codes(np.ascontiguousarray(v.reshape(-1)), n, oc)
    o2 = np.argsort(oc, kind="stable")
    assert np.array_equal(order, o2.astype(np.int32))
    assert np.array_equal(codes, oc[o2])
    ol = np.empty(max(n - 1, 1), np.int32)
    orr = np.empty(max(n - 1, 1), np.int32)
    oracle.lib().or_radix_tree(np.ascontiguousarray(oc[o2]), n, ol, orr)
    assert np.array_equal(left, ol[: n - 1])
    assert np.array_equal(right, orr[: n - 1])


def strip(q):
    """q identical unit quads in a row, two triangles each: runs of equal
    merged areas."""
    x = np.arange(q, dtype=np.float32)
    a = np.stack([x, 0 * x, 0 * x], 1)
    b, c, d = a + [1, 0, 0], a + [1, 1, 0], a + [0, 1, 0]
    return np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)]).astype(np.float32)


def grid(g):
    y, x = np.mgrid[0:g, 0:g].astype(np.float32)
    a = np.stack([x.ravel(), y.ravel(), 0 * x.ravel()], 1)
    b, c, d = a + [1, 0, 0], a + [1, 1, 0], a + [0, 1, 0]
    return np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)]).astype(np.float32)


@pytest.mark.parametrize("case", ["soup1", "soup2", "soup3", "soup1000", "soup5000dup", "soup70000", "strip", "grid"])
def test_ploc_tree_matches_oracle(rn, oracle, case):
    """The default builder's binary tree, node for node, against the oracle's
    sequential PLOC; 70000 triangles runs the multi-kernel rounds before the
    one-work-group tail, the strip and grid have many equal areas."""
    v = {"soup1": lambda: soup(1, seed=1), "soup2": lambda: soup(2, seed=2), "soup3": lambda: soup(3, seed=3), "soup1000": lambda: soup(1000, 1),
         "soup5000dup": lambda: soup(5000, 5, dup=700), "soup70000": lambda: soup(70000, 7),
         "strip": lambda: strip(20000), "grid": lambda: grid(120)}[case]()
    n = len(v)
    order, left, right = rn.ploc_stages(v)
    oc = np.empty(n, np.uint32)
    oracle.lib().or_morton_codes(np.ascontiguousarray(v.reshape(-1)), n, oc)
    o2 = np.argsort(oc, kind="stable").astype(np.int32)
    assert np.array_equal(order, o2)
    ol = np.empty(max(n - 1, 1), np.int32)
    orr = np.empty(max(n - 1, 1), np.int32)
    rounds = oracle.lib().or_ploc_tree(np.ascontiguousarray(v.reshape(-1)), o2, n, ol, orr)
    assert rounds > 0 if n >= 2 else rounds == 0
    assert np.array_equal(left, ol[: n - 1])
    assert np.array_equal(right, orr[: n - 1])


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_non_finite_vertices_are_refused(rn, bad):
    v = soup(100, seed=9)
    v[37, 1, 2] = bad
    with pytest.raises(RuntimeError, match="triangle 37 has a non-finite vertex"):
        rn.ploc_stages(v)
    with pytest.raises(RuntimeError, match="triangle 37 has a non-finite vertex"):
        rn.lbvh_stages(v)
    sc = rn.Scene.from_arrays(v, np.full((100, 3), 0.5, np.float32), None)
    with pytest.raises(RuntimeError, match="triangle 37 has a non-finite vertex"):
        sc.build(8)


NODE = np.dtype([("p", "<f4", 3), ("e", "u1", 3), ("imask", "u1"), ("child_base", "<u4"), ("tri_base", "<u4"),
                 ("tmask", "<u4"), ("pad", "<u4"), ("qlo", "u1", (3, 8)), ("qhi", "u1", (3, 8))])


def leaf_of(nd, s):
    """(first leaf-ordered record, triangle count) of slot s: three tmask
    bits per slot, leaves consecutive in slot order from tri_base."""
    t = int(nd["tmask"])
    bits = (t >> (3 * s)) & 7
    assert bits in (0, 1, 3, 7), "a slot's triangle bits are a unary count"
    return int(nd["tri_base"]) + bin(t & ((1 << (3 * s)) - 1)).count("1"), bin(bits).count("1")


def check_bvh(verts, nodes_raw, prims, width):
    nodes = np.frombuffer(nodes_raw.tobytes(), dtype=NODE)
    n = len(verts)
    assert sorted(prims.tolist()) == list(range(n)), "each triangle must sit in exactly one leaf"
    vmin = verts.min(axis=1).astype(np.float64)
    vmax = verts.max(axis=1).astype(np.float64)
    # subtree triangle ranges (leaf-ordered) for conservativeness checks
    sub_lo = {}
    sub_hi = {}

    def visit(k, depth):
        nd = nodes[k]
        scale = np.ldexp(1.0, nd["e"].astype(np.int64) - 127)
        lo = np.full(3, np.inf)
        hi = np.full(3, -np.inf)
        nch = 0
        ic = 0
        for s in range(8):
            internal = (nd["imask"] >> s) & 1
            first, cnt = leaf_of(nd, s)
            assert not (internal and cnt), "an internal slot has no triangles"
            if not internal and cnt == 0:
                continue
            nch += 1
            blo = nd["p"].astype(np.float64) + nd["qlo"][:, s] * scale
            bhi = nd["p"].astype(np.float64) + nd["qhi"][:, s] * scale
            if internal:
                c = int(nd["child_base"]) + ic
                ic += 1
                clo, chi, cd = visit(c, depth + 1)
                depth_c = cd
            else:
                idx = prims[first: first + cnt]
                clo = vmin[idx].min(0)
                chi = vmax[idx].max(0)
            assert (blo <= clo).all() and (bhi >= chi).all(), "quantised child box must contain its child"
            lo = np.minimum(lo, clo)
            hi = np.maximum(hi, chi)
        assert nch <= width
        return lo, hi, depth

    visit(0, 1)
    return nodes


def canonical_tree(nodes_raw, prims):
    """The tree as nested tuples: per node its quantisation frame and, per
    occupied slot in slot order, the child's boxes and subtree (or leaf prims)."""
    nodes = np.frombuffer(nodes_raw.tobytes(), dtype=NODE)

    def canon(k):
        nd = nodes[k]
        out = [tuple(nd["p"].tolist()), tuple(nd["e"].tolist())]
        ic = 0
        for s in range(8):
            internal = (nd["imask"] >> s) & 1
            b, cnt = leaf_of(nd, s)
            if not internal and cnt == 0:
                continue
            box = (tuple(nd["qlo"][:, s].tolist()), tuple(nd["qhi"][:, s].tolist()))
            if internal:
                out.append((s, box, canon(int(nd["child_base"]) + ic)))
                ic += 1
            else:
                out.append((s, box, tuple(prims[b:b + cnt].tolist())))
        return tuple(out)

    return canon(0)


@pytest.mark.parametrize("builder", ["ploc", "lbvh"])
@pytest.mark.parametrize("width", [8, 4])
def test_bvh8_structure(rn, width, builder):
    v = soup(20000, seed=3, dup=50)
    sc = rn.Scene.from_arrays(v, np.full((len(v), 3), 0.5, np.float32))
    sc.set_builder(builder)
    nodes, prims = sc.export_bvh(width)
    check_bvh(v, nodes, prims, width)
    assert sc.build(width)["builder"] == {"ploc": 0, "lbvh": 1}[builder]
    # both builders are deterministic; the collapse hands out node and leaf
    # slots by atomics, so compare the trees with slot addresses taken out
    nodes2, prims2 = sc.export_bvh(width)
    assert canonical_tree(nodes, prims) == canonical_tree(nodes2, prims2)


def test_flat_child_on_grid_line(rn):
    """A node's grid origin puts a flat child on a grid line: the triangle in
    the plane y = 0.3 keeps a zero-thickness box although its plane is not
    the node's lower y (0.1), which with the origin at the lower corner would
    lie 3.2 steps of 1/16 up and gain a step of thickness."""
    v = np.array([[[0.0, 0.3, 0.0], [5.0, 0.3, 0.0], [0.0, 0.3, 5.0]],
                  [[2.0, 0.1, 9.0], [10.0, 10.1, 3.0], [7.0, 4.0, 10.0]]], np.float32)
    sc = rn.Scene.from_arrays(v, np.full((2, 3), 0.5, np.float32))
    raw, prims = sc.export_bvh(8)
    check_bvh(v, raw, prims, 8)
    nd = np.frombuffer(raw.tobytes(), dtype=NODE)[0]
    assert nd["imask"] == 0
    sy = np.float32(2.0) ** (int(nd["e"][1]) - 127)
    assert sy == np.float32(1 / 16)
    flat = [s for s in range(8) if leaf_of(nd, s)[1] and prims[leaf_of(nd, s)[0]] == 0]
    assert len(flat) == 1
    s = flat[0]
    assert nd["qlo"][1, s] == nd["qhi"][1, s]
    assert nd["p"][1] + np.float32(nd["qlo"][1, s]) * sy == np.float32(0.3)
    # the ray along the plane's normal from just below does not reach the box
    prim, t, _, _ = sc.trace_rays(np.array([[1.0, 0.2, 1.0]], np.float32), np.array([[0, 1, 0]], np.float32))
    assert prim[0] == 0 and abs(t[0] - 0.1) < 1e-6


def test_non_finite_rays_miss(rn):
    """A ray with a NaN or infinite component misses every triangle (and is
    unoccluded), as its comparisons against every triangle fail; the
    traversal answers it at once instead of entering every box (NaN slab
    distances vanish in the min/max)."""
    v = soup(3000, seed=5)
    sc = rn.Scene.from_arrays(v, np.full((len(v), 3), 0.5, np.float32))
    nan, inf = np.float32(np.nan), np.float32(np.inf)
    o = np.array([[0, 0, 0], [nan, 0, 0], [0, inf, 0], [0, 0, 0], [0, 0, 0], [0, 0, -inf]], np.float32)
    d = np.array([[0, 0, 1], [0, 0, 1], [0, 0, 1], [nan, 0, 1], [0, -inf, 0], [0, 0, 1]], np.float32)
    prim, t, u, w = sc.trace_rays(o, d)
    assert (prim[1:] == -1).all()
    occ = sc.trace_rays(o, d, tmax=np.full(len(o), 100.0, np.float32), query=rn.QUERY_ANY)
    assert (occ[1:] == 0).all()
    # the finite ray of the batch is answered as usual
    ref_prim, _, _, _ = sc.trace_rays(o[:1], d[:1])
    assert prim[0] == ref_prim[0]


@pytest.mark.parametrize("builder", ["ploc", "lbvh"])
def test_closest_and_any_vs_brute(rn, oracle, builder):
    v = soup(3000, seed=5)
    sc = rn.Scene.from_arrays(v, np.full((len(v), 3), 0.5, np.float32))
    sc.set_builder(builder)
    rng = np.random.default_rng(9)
    nr = 4096
    o = rng.uniform(-8, 8, size=(nr, 3)).astype(np.float32)
    d = rng.normal(size=(nr, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    prim, t, u, w = sc.trace_rays(o, d)
    ps = oracle.PhysScene(v, np.full((len(v), 3), 0.5, np.float32), None, "f32")
    import ctypes as C
    mism = 0
    for i in range(nr):
        tt, uu, vv = C.c_float(), C.c_float(), C.c_float()
        bp = oracle.lib().or_phys_closest_f32(ps.h, np.ascontiguousarray(o[i]), np.ascontiguousarray(d[i]),
                                             np.float32(np.inf), 1, C.byref(tt), C.byref(uu), C.byref(vv))
        if bp != prim[i]:
            mism += 1
            continue
        if bp >= 0:
            assert t[i] == tt.value and u[i] == uu.value and w[i] == vv.value
    assert mism == 0, f"{mism} closest-hit mismatches"
    tmax = np.where(prim >= 0, t * 0.999, 100.0).astype(np.float32)
    occ = sc.trace_rays(o, d, tmax=tmax, query=rn.QUERY_ANY)
    # occluded strictly before the closest hit never happens; beyond a miss never
    assert (occ[prim < 0] == 0).all()
    tmax2 = np.where(prim >= 0, t * 1.001, 100.0).astype(np.float32)
    occ2 = sc.trace_rays(o, d, tmax=tmax2, query=rn.QUERY_ANY)
    assert (occ2[prim >= 0] == 1).all()


def _surface_rays_vs_brute(rn, oracle, sd, n_surf, n_in, seed):
    """Rays leaving the scene's surfaces the way the path tracer's secondary
    rays do (origin 0.05 off the surface, on the side the ray goes) and rays
    from the box interior give the brute-force closest hit bit for bit, and
    shadow rays toward the light are occluded exactly when the closest hit
    comes first."""
    import ctypes as C
    v = sd.verts
    sc = rn.Scene.from_arrays(v, sd.albedo, sd.emission)
    rng = np.random.default_rng(seed)
    tri = rng.integers(0, len(v), n_surf)
    b = rng.dirichlet((1, 1, 1), n_surf).astype(np.float32)
    p = (b[:, :1] * v[tri, 0] + b[:, 1:2] * v[tri, 1] + b[:, 2:] * v[tri, 2]).astype(np.float32)
    n = np.cross(v[tri, 1] - v[tri, 0], v[tri, 2] - v[tri, 0]).astype(np.float64)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    d = rng.normal(size=(n_surf, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    n = np.where((n * d).sum(1, keepdims=True) < 0, -n, n)
    o_surf = (p + 0.05 * n).astype(np.float32)
    o_in = rng.uniform([30, 30, 30], [525, 525, 525], size=(n_in, 3)).astype(np.float32)
    o = np.concatenate([o_surf, o_in])
    d = np.concatenate([d, rng.normal(size=(n_in, 3))])
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    prim, t, u, w = sc.trace_rays(o, d)
    ps = oracle.PhysScene(v, sd.albedo, sd.emission, "f32")
    mism = 0
    for i in range(len(o)):
        tt, uu, vv = C.c_float(), C.c_float(), C.c_float()
        bp = oracle.lib().or_phys_closest_f32(ps.h, np.ascontiguousarray(o[i]), np.ascontiguousarray(d[i]),
                                             np.float32(np.inf), 1, C.byref(tt), C.byref(uu), C.byref(vv))
        if bp != prim[i]:
            mism += 1
            continue
        if bp >= 0:
            assert t[i] == tt.value and u[i] == uu.value and w[i] == vv.value, i
    assert mism == 0, f"{mism} closest-hit mismatches"
    light = np.flatnonzero(sd.emission.max(1) > 0)
    lt = rng.choice(light, n_surf)
    lb = rng.dirichlet((1, 1, 1), n_surf).astype(np.float32)
    q = (lb[:, :1] * v[lt, 0] + lb[:, 1:2] * v[lt, 1] + lb[:, 2:] * v[lt, 2]).astype(np.float32)
    to = q - o_surf
    dist = np.linalg.norm(to, axis=1)
    sdir = (to / dist[:, None]).astype(np.float32)
    tmax = (dist * 0.999).astype(np.float32)
    occ = sc.trace_rays(o_surf, sdir, tmax=tmax, query=rn.QUERY_ANY)
    # against brute force over every triangle (any accepted hit before tmax)
    brute = oracle.all_hits(v, o_surf, sdir, "watertight", tmax=tmax, cap=1 << 12, phys=ps)
    want = np.array([len(p) > 0 for p, _ in brute])
    assert np.array_equal(occ.astype(bool), want), int((occ.astype(bool) != want).sum())
    assert 0 < want.sum() < len(want)  # both outcomes occur


def test_flat_scene_closest_and_any_vs_brute(rn, oracle):
    """The Cornell box: axis-aligned walls and the light lie on grid lines of
    their nodes, so their boxes have zero thickness."""
    from rednoise_amd import scenes
    _surface_rays_vs_brute(rn, oracle, scenes.cornell_box(), 3000, 1000, 21)


def test_bench_scene_closest_and_any_vs_brute(rn, oracle):
    """The 1,048,588-triangle scene of the throughput benchmark (displaced
    sphere in the Cornell walls): 384 surface rays and 128 interior rays
    against brute force over every triangle."""
    from rednoise_amd import scenes
    _surface_rays_vs_brute(rn, oracle, scenes.config1_scene(0), 384, 128, 22)


def test_tie_heavy_scene_closest_and_any_vs_brute(rn, oracle):
    """Runs of equal merged areas shape the clustering tree: a 150 x 150 grid
    of identical quads (a floor at y = 100) and a strip of 2,000 identical
    quads (at y = 300) inside the Cornell box, against brute force."""
    from rednoise_amd import scenes
    box = scenes.cornell_box()
    g = grid(150) * np.float32(3.0)
    g = np.stack([50 + g[..., 0], np.full_like(g[..., 0], 100.0), 50 + g[..., 1]], -1)
    st = strip(2000) * np.float32(0.2)
    st = np.stack([50 + st[..., 0], np.full_like(st[..., 0], 300.0), 250 + st[..., 1]], -1)
    extra = np.concatenate([g, st]).astype(np.float32)
    sd = type(box)(**{**box.__dict__,
                      "verts": np.concatenate([box.verts, extra]).astype(np.float32),
                      "albedo": np.concatenate([box.albedo, np.full((len(extra), 3), 0.5, np.float32)]),
                      "emission": np.concatenate([box.emission, np.zeros((len(extra), 3), np.float32)])})
    _surface_rays_vs_brute(rn, oracle, sd, 1000, 300, 23)


def _edge_band(u, v, eps=1e-6):
    return np.minimum(np.minimum(u, v), 1.0 - u - v) < eps


@pytest.mark.parametrize("asset,scale", [("cornell", 0.35), ("quad", 0.35), ("sphere", 1.0)])
def test_compat_primary_bit_exact(rn, asset, scale):
    """rn_render camera + reference triangle test on the GPU BVH8 against the
    reference BVH::intersect: identical triangle ids and bit-identical
    (t, u, v) outside the 1e-6 edge band."""
    g = np.load(os.path.join(GOLDEN, f"primary_{asset}.npz"))
    names = {"cornell": "cornell-box.obj", "quad": "textured-quad.obj", "sphere": "sphere.obj"}
    sc = rn.Scene.load_obj(os.path.join(ASSETS, names[asset]), scale)
    W, H = int(g["width"]), int(g["height"])
    prim, t, u, v, texel = sc.trace_primary_compat(W, H, float(g["cam_z"]))
    same = prim == g["prim"]
    band = _edge_band(g["u"], g["v"]) | _edge_band(u, v)
    bad = ~same & ~band
    assert bad.sum() == 0, f"{bad.sum()} id mismatches outside the edge band ({(~same).sum()} total)"
    hit = same & (prim >= 0)
    assert np.array_equal(t[hit], g["t"][hit]) and np.array_equal(u[hit], g["u"][hit]) and np.array_equal(
        v[hit], g["v"][hit])
    if asset == "quad":
        tc = sc.compat_attributes()["texcoords"]
        tex = sc.texture(0)
        th, tw = tex.shape
        from oracle import oracle as O
        exp = np.array([O.lib().or_compat_texel(float(g["u"][i]), float(g["v"][i]), np.ascontiguousarray(tc[p]),
                                                tw, th) if p >= 0 else -1 for i, p in enumerate(g["prim"])])
        assert np.array_equal(texel[same], exp[same])


@pytest.mark.parametrize("asset", ["cornell", "quad", "sphere"])
def test_compat_random_rays_vs_reference_bvh(rn, asset):
    g = np.load(os.path.join(GOLDEN, f"rays_{asset}.npz"))
    names = {"cornell": ("cornell-box.obj", 0.35), "quad": ("textured-quad.obj", 0.35), "sphere": ("sphere.obj", 1.0)}
    sc = rn.Scene.load_obj(os.path.join(ASSETS, names[asset][0]), names[asset][1])
    prim, t, u, v = sc.trace_rays(g["o"], g["d"], tri_test=rn.TRI_COMPAT)
    same = prim == g["prim"]
    band = _edge_band(g["u"], g["v"]) | _edge_band(u, v)
    assert (~same & ~band).sum() == 0
    hit = same & (prim >= 0)
    assert np.array_equal(t[hit], g["t"][hit])
    # any hit against BVH::occluded of the compiled reference: equal on every
    # ray except where the occluding hit lies in the edge band or within
    # 1e-6 t of maxd (none occur on these rays; the count is printed)
    occ = sc.trace_rays(g["o"], g["d"], tmax=g["maxd"], query=rn.QUERY_ANY, tri_test=rn.TRI_COMPAT)
    from oracle import oracle as O
    ob = np.load(os.path.join(GOLDEN, f"obj_{asset}.npz"))
    hits = O.all_hits(ob["verts"], g["o"], g["d"], "compat", cap=1 << 12)
    in_band = np.zeros(len(occ), bool)
    for i, (ps, ts) in enumerate(hits):
        near_maxd = np.abs(ts - g["maxd"][i]) <= 1e-6 * np.abs(ts)
        in_band[i] = bool(near_maxd.any())
    diff = occ != g["occluded"]
    print(f"{asset}: any-hit mismatches {int(diff.sum())} of {len(occ)} ({int((diff & in_band).sum())} in the band)")
    assert (diff & ~in_band).sum() == 0


def cornell_setup(rn, oracle, W, H, spp, depth=5, rr=3, eps=0.05, seed=0):
    from rednoise_amd import scenes
    sd = scenes.cornell_box()
    sc = rn.Scene.from_arrays(sd.verts, sd.albedo, sd.emission)
    cam = rn.Camera.make(**sd.camera)
    oc = oracle.camera_basis(sd.camera["eye"], sd.camera["target"], sd.camera["up"], sd.camera["vfov_deg"], W, H)
    p = rn.RenderParams.make(W, H, spp, depth, rr, seed=seed, ray_eps=eps)
    op = oracle.params(W, H, spp, depth, rr, 0.0, seed, 0, eps)
    return sd, sc, cam, p, oc, op


def test_megakernel_vs_oracle_fp32(rn, oracle):
    sd, sc, cam, p, oc, op = cornell_setup(rn, oracle, 80, 60, 8)
    gpu = sc.render_hdr(cam, p)
    ps = oracle.PhysScene(sd.verts, sd.albedo, sd.emission, "f32")
    cpu, _ = ps.render(oc, op)
    close = np.abs(gpu - cpu) <= 1e-3 * np.maximum(1.0, np.abs(cpu))
    assert close.mean() > 0.98, close.mean()
    st = rn.last_stats()
    assert st["paths"] == 80 * 60 * 8 and st["rays_closest"] > st["paths"]


def test_wavefront_matches_megakernel(rn, oracle):
    sd, sc, cam, p, oc, op = cornell_setup(rn, oracle, 96, 64, 16)
    mk = sc.render_hdr(cam, p)
    s_mk = rn.last_stats()
    p.integrator = rn.INTEGRATOR_WAVEFRONT
    p.queue_paths = 20000
    wf = sc.render_hdr(cam, p)
    s_wf = rn.last_stats()
    assert np.allclose(wf, mk, rtol=1e-4, atol=1e-5)
    assert s_wf["paths"] == s_mk["paths"]
    assert s_wf["rays_closest"] == s_mk["rays_closest"]
    assert s_wf["rays_shadow"] == s_mk["rays_shadow"]


@pytest.mark.parametrize("shape", [(96, 64, 16, 20000), (97, 61, 5, 6000), (1, 1, 64, 0), (7, 5, 8, 0),
                                   (65, 3, 4, 64)])
def test_wavefront_slots_are_the_megakernel_sums(rn, oracle, shape):
    """With radiance slots (the default) the wavefront adds a path's
    contributions in depth order and a pixel's paths in sample order, as the
    megakernel does: the same bits, on every run, for an image whose size is
    not a multiple of the 8x8 tile, with a queue that refills many times, and
    for images one tile (or one pixel) wide, where the work-index map divides
    by one.
    With RN_FLAG_FB_ATOMICS (float atomics, arbitrary order) the image agrees
    to float rounding."""
    W, H, spp, queue = shape
    sd, sc, cam, p, oc, op = cornell_setup(rn, oracle, W, H, spp, seed=3)
    mk = sc.render_hdr(cam, p)
    p.integrator = rn.INTEGRATOR_WAVEFRONT
    p.queue_paths = queue
    wf = sc.render_hdr(cam, p)
    assert np.isfinite(wf).all() and wf.mean() > 0
    assert np.array_equal(wf, mk)
    assert np.array_equal(sc.render_hdr(cam, p), wf)
    p.flags = rn.FLAG_FB_ATOMICS
    at = sc.render_hdr(cam, p)
    assert np.allclose(at, mk, rtol=1e-5, atol=1e-6)


def test_mode_r_matches_oracle_same_stream(rn, oracle):
    """GPU reference estimator vs the CPU restatement with the same Philox
    stream: most 8-bit pixels identical (libm sin/cos/pow/exp and FMA differ)."""
    sc = rn.Scene.load_obj(os.path.join(ASSETS, "cornell-box.obj"), 0.35)
    W, H, spp = 64, 48, 8
    img = sc.render_seeded(W, H, 4.0, spp, 11, 22).to_numpy()
    ca = sc.compat_attributes()
    v, _, _ = sc.triangles()
    import ctypes as C
    cs = oracle.CompatScene()
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt).reshape(-1)
        keep.append(a)
        return a.ctypes.data

    cs.n = sc.triangle_count
    cs.verts = arr(v, np.float32)
    cs.normals = arr(ca["normals"], np.float32)
    cs.colour = arr(ca["colour"], np.int32)
    cs.material = arr(ca["material"], np.int32)
    cs.texture_id = arr(ca["texture_id"], np.int32)
    cs.texcoords = arr(ca["texcoords"], np.float32)
    cs.n_textures = 0
    out = np.zeros(W * H * 4, np.uint8)
    oracle.lib().or_compat_render(C.byref(cs), W, H, 4.0, spp, 11, 22, 0, H, out)
    out = out.reshape(H, W, 4)
    eq = (img[..., :3].astype(int) - out[..., :3].astype(int))
    assert (np.abs(eq) <= 1).mean() > 0.95
    assert np.abs(eq).mean() < 1.0


@pytest.mark.parametrize("builder", ["ploc", "lbvh"])
@pytest.mark.parametrize("n,same", [(2, False), (3, False), (5, False), (17, True), (9000, False), (9000, True)])
def test_small_and_degenerate_builds(rn, oracle, builder, n, same):
    """Tiny scenes, and scenes whose triangles all share one centroid (every
    Morton code equal, every PLOC candidate pair tied): the tree is valid and
    the closest hits equal brute force."""
    v = soup(n, seed=n)
    if same:
        v = v - v.mean(axis=1, keepdims=True)  # every centroid at the origin
    sc = rn.Scene.from_arrays(v, np.full((n, 3), 0.5, np.float32))
    sc.set_builder(builder)
    nodes, prims = sc.export_bvh(8)
    check_bvh(v, nodes, prims, 8)
    rng = np.random.default_rng(n)
    o = rng.uniform(-8, 8, size=(512, 3)).astype(np.float32)
    tgt = v.reshape(-1, 3)[rng.integers(0, 3 * n, 512)]
    d = tgt - o
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    prim, t, _, _ = sc.trace_rays(o, d)
    ps = oracle.PhysScene(v, np.full((n, 3), 0.5, np.float32), None, "f32")
    import ctypes as C
    for i in range(len(o)):
        tt, uu, vv = C.c_float(), C.c_float(), C.c_float()
        bp = oracle.lib().or_phys_closest_f32(ps.h, np.ascontiguousarray(o[i]), np.ascontiguousarray(d[i]),
                                             np.float32(np.inf), 1, C.byref(tt), C.byref(uu), C.byref(vv))
        assert bp == prim[i] and (bp < 0 or t[i] == tt.value), (i, bp, prim[i])
