"""Concurrent calls (the reference's rn_render takes a const scene and may be
called from several threads, src/capi.cpp:44-79): threads rendering the same
scene, other scenes and other modes at once get the images a sequential run
gets. ctypes releases the GIL for the length of each call."""
import os
import threading

import numpy as np
import pytest

from conftest import ASSETS

pytestmark = pytest.mark.gpu


def test_concurrent_renders_match_sequential(rn):
    cornell = rn.Scene.load_obj(os.path.join(ASSETS, "cornell-box.obj"), 0.35)
    quad = rn.Scene.load_obj(os.path.join(ASSETS, "textured-quad.obj"), 0.35)
    jobs = [(cornell, "pathtraced", 64, 48, 4), (cornell, "raytraced", 80, 60, 1), (quad, "pathtraced", 48, 40, 8),
            (quad, "raytraced", 64, 64, 1)] * 3
    want = [sc.render(mode, w, h, 4.0, spp).to_numpy() for sc, mode, w, h, spp in jobs]
    got = [None] * len(jobs)
    errors = []

    def run(k):
        try:
            sc, mode, w, h, spp = jobs[k]
            got[k] = sc.render(mode, w, h, 4.0, spp).to_numpy()
        except Exception as e:  # reported below, with the job
            errors.append((k, repr(e)))

    threads = [threading.Thread(target=run, args=(k,)) for k in range(len(jobs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
    for k, (a, b) in enumerate(zip(got, want)):
        assert np.array_equal(a, b), k


def test_scene_freed_while_another_renders(rn):
    keep = rn.Scene.load_obj(os.path.join(ASSETS, "cornell-box.obj"), 0.35)
    want = keep.render("pathtraced", 64, 48, 4.0, 4).to_numpy()
    out = {}

    def render():
        out["img"] = keep.render("pathtraced", 64, 48, 4.0, 4).to_numpy()

    t = threading.Thread(target=render)
    t.start()
    for _ in range(8):
        tmp = rn.Scene.load_obj(os.path.join(ASSETS, "sphere.obj"), 1.0)
        tmp.render("raytraced", 32, 32, 4.0, 1)
        tmp.close()
    t.join(timeout=120)
    assert not t.is_alive()
    assert np.array_equal(out["img"], want)


def test_batch_and_set_materials_on_the_same_scene(rn):
    """A batch holds every scene it renders locked from upload to download:
    while another thread keeps replacing one scene's albedos, each batch sees
    one material set or the other, never a mix (two items of that scene in the
    same batch agree, and each equals one of the two sequential renders)."""
    from rednoise_amd import scenes
    sd = scenes.cornell_box()
    sc = rn.Scene.from_arrays(sd.verts, sd.albedo, sd.emission)
    other = rn.Scene.from_arrays(sd.verts, sd.albedo, sd.emission)
    cam = rn.Camera.make(**sd.camera)
    p = rn.RenderParams.make(48, 36, 4, 5, 3, seed=1, ray_eps=0.05)
    alb2 = (sd.albedo * 0.5).astype(np.float32)
    opts = rn.BatchOptions(0, 0, 8, 0, 8)
    items = [(sc, cam, p), (other, cam, p), (sc, cam, p)]
    want1 = rn.render_batch(items, opts)[0]
    sc.set_materials(albedo=alb2)
    want2 = rn.render_batch(items, opts)[0]
    sc.set_materials(albedo=sd.albedo)
    assert not np.allclose(want1, want2, rtol=1e-2)

    def close(a, b):
        return np.allclose(a, b, rtol=1e-4, atol=1e-6)

    stop = threading.Event()
    errors = []

    def flip():
        k = 0
        try:
            while not stop.is_set():
                sc.set_materials(albedo=alb2 if k % 2 == 0 else sd.albedo)
                k += 1
        except Exception as e:  # reported below
            errors.append(repr(e))

    t = threading.Thread(target=flip)
    t.start()
    try:
        for _ in range(16):
            a, _, b = rn.render_batch(items, opts)
            assert close(a, b)
            assert close(a, want1) or close(a, want2)
    finally:
        stop.set()
        t.join(timeout=60)
    assert not t.is_alive() and not errors, errors
