"""Write golden vectors from the compiled reference (oracle/_ref/librednoise_ref.so).

Run in the build container, where the reference sources exist:
    make -C oracle ref && python tests/golden/make_golden.py
Outputs (small .npz fixtures next to this script):
  obj_<asset>.npz      loadOBJ triangles/normals/colours/materials/texcoords
  primary_<asset>.npz  rn_render camera rays (160x120, cam_z 4; rays stored for
                       every 37th pixel) and the
                       BVH::intersect (prim, t, u, v) of each
  rays_<asset>.npz     seeded random rays: BVH::intersect and
                       getClosestIntersection results
  modeR_cornell.npz    rn_render(RN_PATHTRACED) RGBA8 images of the Cornell
                       asset (reference estimator, per-pixel mt19937)
  noise.npz            fractalNoise(p, 4) at seeded points
With --whitted (only these files):
  whitted_<asset>.npz  rn_render(RN_RAYTRACED) RGBA8 images (160x120, 96x64,
                       cam_z 4) and loadOBJ's smoothed vertex normals
With --overflow (only this file):
  modeR_overflow.npz   rn_render(RN_PATHTRACED) of the Cornell asset at
                       1024x432, 1 spp, cam_z 4: rows 184-247, columns
                       320-703, box pixels on both sides of y*W+x = 219,553,
                       past which the reference's int seed (y*W+x)*9781+1
                       (src/Renderer.cpp:1139) overflows
With --photon (only these files):
  photon_<asset>.npz   renderPhotonMapped (20000 photons; no final gather and
                       3 gather rays) and renderBidirectional (16 samples,
                       exposure 40) RGBA8 images at 64x48 with rn_render's
                       camera (cam_z 4)
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

ASSETS = os.path.join(HERE, "assets")


def load(ref, name, scale):
    path = os.path.join(ASSETS, name).encode()
    n = ref.ref_load_obj(path, scale, None, None, None, None, None, None)
    v = np.empty(9 * n, np.float32)
    nm = np.empty(3 * n, np.float32)
    col = np.empty(3 * n, np.int32)
    mat = np.empty(n, np.int32)
    ht = np.empty(n, np.int32)
    tc = np.empty(6 * n, np.float32)
    ref.ref_load_obj(path, scale, v.ctypes.data, nm.ctypes.data, col.ctypes.data, mat.ctypes.data, ht.ctypes.data,
                     tc.ctypes.data)
    return dict(verts=v, normals=nm, colour=col, material=mat, has_texture=ht, texcoords=tc, scale=np.float32(scale))


def bvh(ref, verts, o, d, ignore=None):
    n = verts.size // 9
    nr = o.size // 3
    prim = np.empty(nr, np.int32)
    t = np.empty(nr, np.float32)
    u = np.empty(nr, np.float32)
    v = np.empty(nr, np.float32)
    ig = None if ignore is None else np.ascontiguousarray(ignore, np.int32).ctypes.data
    ref.ref_bvh_intersect(verts, n, nr, o, d, ig, prim, t, u, v)
    return prim, t, u, v


def brute(ref, verts, o, d):
    n = verts.size // 9
    nr = o.size // 3
    prim = np.empty(nr, np.int32)
    t = np.empty(nr, np.float32)
    u = np.empty(nr, np.float32)
    v = np.empty(nr, np.float32)
    ref.ref_brute_closest(verts, n, nr, o, d, prim, t, u, v)
    return prim, t, u, v


WHITTED = {"cornell": ("cornell-box.obj", 0.35, 4.0), "quad": ("textured-quad.obj", 0.35, 4.0),
           "sphere": ("sphere.obj", 1.0, 4.0)}


def whitted(ref) -> None:
    """rn_render(RN_RAYTRACED) images of the compiled reference and loadOBJ's
    smoothed vertex normals, per asset: whitted_<asset>.npz."""
    ref.ref_load_obj_vnormals.argtypes = [C.c_char_p, C.c_float, C.c_void_p]
    ref.ref_load_obj_vnormals.restype = C.c_int
    for key, (name, scale, cz) in WHITTED.items():
        path = os.path.join(ASSETS, name).encode()
        n = ref.ref_load_obj_vnormals(path, scale, None)
        vn = np.empty(9 * n, np.float32)
        ref.ref_load_obj_vnormals(path, scale, vn.ctypes.data)
        sc = ref.rn_scene_load_obj(path, C.c_float(scale))
        out = {"vnormals": vn, "scale": np.float32(scale), "cam_z": np.float32(cz)}
        for (W, H) in [(160, 120), (96, 64)]:
            buf = np.empty(W * H * 4, np.uint8)
            assert ref.rn_render(sc, 2, W, H, C.c_float(cz), 1, buf) == 1
            out[f"rgba_{W}x{H}"] = buf.reshape(H, W, 4)
        ref.rn_scene_free(sc)
        np.savez_compressed(os.path.join(HERE, f"whitted_{key}.npz"), **out)
        print("whitted", key, {k: v.shape for k, v in out.items()})


PHOTON = {"cornell": ("cornell-box.obj", 0.35), "quad": ("textured-quad.obj", 0.35), "sphere": ("sphere.obj", 1.0)}


def photon(ref) -> None:
    """renderPhotonMapped / renderBidirectional images of the compiled
    reference per asset: photon_<asset>.npz."""
    W, H, cz, nph, spp, expo = 64, 48, 4.0, 20000, 16, 40.0
    for key, (name, scale) in PHOTON.items():
        path = os.path.join(ASSETS, name).encode()
        out = {"scale": np.float32(scale), "cam_z": np.float32(cz), "num_photons": np.int32(nph),
               "bdpt_samples": np.int32(spp), "bdpt_exposure": np.float32(expo)}
        for g in (0, 3):
            buf = np.empty(W * H * 4, np.uint8)
            assert ref.ref_render_photon(path, scale, W, H, cz, nph, g, buf) == 1
            out[f"photon_g{g}"] = buf.reshape(H, W, 4)
        buf = np.empty(W * H * 4, np.uint8)
        assert ref.ref_render_bidirectional(path, scale, W, H, cz, spp, expo, buf) == 1
        out["bdpt"] = buf.reshape(H, W, 4)
        np.savez_compressed(os.path.join(HERE, f"photon_{key}.npz"), **out)
        print("photon", key, {k: v.shape for k, v in out.items()})


def main() -> None:
    ref = O.ref()
    if ref is None:
        raise SystemExit("build the reference first: make -C oracle ref")
    assets = {"cornell": ("cornell-box.obj", 0.35), "quad": ("textured-quad.obj", 0.35),
              "sphere": ("sphere.obj", 1.0)}
    for key, (name, scale) in assets.items():
        ob = load(ref, name, scale)
        np.savez_compressed(os.path.join(HERE, f"obj_{key}.npz"), **ob)
        W, H, cz = 160, 120, 4.0
        o = np.empty(3 * W * H, np.float32)
        d = np.empty(3 * W * H, np.float32)
        ref.ref_primary_rays(W, H, C.c_float(cz), None, o, d)
        prim, t, u, v = bvh(ref, ob["verts"], o, d)
        sub = np.arange(0, W * H, 37)
        np.savez_compressed(os.path.join(HERE, f"primary_{key}.npz"), sub=sub, o_sub=o.reshape(-1, 3)[sub],
                            d_sub=d.reshape(-1, 3)[sub], prim=prim, t=t, u=u, v=v, width=W, height=H,
                            cam_z=np.float32(cz))
        rng = np.random.default_rng(12345)
        nr = 2048
        vv = ob["verts"].reshape(-1, 3)
        lo, hi = vv.min(0), vv.max(0)
        ro = rng.uniform(lo - 0.5 * (hi - lo), hi + 0.5 * (hi - lo), size=(nr, 3)).astype(np.float32)
        tgt = rng.uniform(lo, hi, size=(nr, 3)).astype(np.float32)
        rd = tgt - ro
        rd /= np.linalg.norm(rd, axis=1, keepdims=True)
        rd = rd.astype(np.float32)
        ro, rd = ro.reshape(-1), rd.reshape(-1)
        bp, bt, bu, bv = bvh(ref, ob["verts"], ro, rd)
        fp, ft, fu, fv = brute(ref, ob["verts"], ro, rd)
        maxd = rng.uniform(0.1, 1.5, nr).astype(np.float32) * np.float32(np.linalg.norm(hi - lo))
        occ = np.empty(nr, np.int32)
        ref.ref_bvh_occluded(ob["verts"], ob["verts"].size // 9, nr, ro, rd, maxd, None, occ)
        np.savez_compressed(os.path.join(HERE, f"rays_{key}.npz"), o=ro, d=rd, prim=bp, t=bt, u=bu, v=bv,
                            brute_prim=fp, brute_t=ft, brute_u=fu, brute_v=fv, maxd=maxd, occluded=occ)
        print(key, "tris", ob["verts"].size // 9, "primary hits", int((prim >= 0).sum()))
    # Reference-estimator images of the Cornell asset.
    sc = ref.rn_scene_load_obj(os.path.join(ASSETS, "cornell-box.obj").encode(), 0.35)
    imgs = {}
    for (W, H, spp) in [(160, 120, 64)]:
        buf = np.empty(W * H * 4, np.uint8)
        t0 = time.time()
        ok = ref.rn_render(sc, 3, W, H, C.c_float(4.0), spp, buf)
        assert ok == 1
        print(f"reference rn_render {W}x{H} spp={spp}: {time.time() - t0:.1f}s")
        imgs[f"rgba_{W}x{H}_{spp}"] = buf.reshape(H, W, 4)
    ref.rn_scene_free(sc)
    np.savez_compressed(os.path.join(HERE, "modeR_cornell.npz"), **imgs)
    rng = np.random.default_rng(7)
    pts = rng.uniform(-20, 20, size=(1024, 3)).astype(np.float32)
    out = np.empty(1024, np.float32)
    ref.ref_fractal_noise(1024, pts.reshape(-1), 4, out)
    np.savez_compressed(os.path.join(HERE, "noise.npz"), p=pts, fbm4=out)


def overflow(ref) -> None:
    # The asset keeps its pixel size (f = 300 px) centred in the image: at
    # 1024x432 the box's middle rows hold pixel 219,553 (row 214, x 417),
    # so the band kept here has box pixels on both sides of the overflow.
    W, H, spp = 1024, 432, 1
    y0, y1, x0, x1 = 184, 248, 320, 704
    assert y0 * W + x0 < 219_554 < (y1 - 1) * W + x1
    sc = ref.rn_scene_load_obj(os.path.join(ASSETS, "cornell-box.obj").encode(), 0.35)
    buf = np.empty(W * H * 4, np.uint8)
    assert ref.rn_render(sc, 3, W, H, C.c_float(4.0), spp, buf) == 1
    ref.rn_scene_free(sc)
    img = buf.reshape(H, W, 4)[y0:y1, x0:x1]
    np.savez_compressed(os.path.join(HERE, "modeR_overflow.npz"), rgba=img, width=W, height=H, spp=spp,
                        rows=np.array([y0, y1]), cols=np.array([x0, x1]), cam_z=np.float32(4.0),
                        scale=np.float32(0.35))
    print("overflow band", img.shape, "distinct colours", len(np.unique(img.reshape(-1, 4), axis=0)))


if __name__ == "__main__":
    if "--overflow" in sys.argv:
        r = O.ref()
        if r is None:
            raise SystemExit("build the reference first: make -C oracle ref")
        overflow(r)
    elif "--whitted" in sys.argv or "--photon" in sys.argv:
        r = O.ref()
        if r is None:
            raise SystemExit("build the reference first: make -C oracle ref")
        if "--whitted" in sys.argv:
            whitted(r)
        if "--photon" in sys.argv:
            photon(r)
    else:
        main()
