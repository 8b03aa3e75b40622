#!/usr/bin/env python3
"""Benchmark of the path-trace hot path on one MI355X.

Headline workload (config 1): a 1,048,576-triangle displaced sphere in
the Cornell box, 1920x1080, 64 spp, max depth 8, Russian roulette from depth 3;
every step rebuilds the BVH (PLOC tree, BVH8 collapse) on the GPU and renders the frame. `value` is
the wavefront integrator's throughput in Mrays/s (closest-hit + shadow
queries / wall time of the step, build included); the plain per-pixel
megakernel is timed over the same number of steps beside it. Both run the same
estimator with the same Philox streams, so their relative RMSE against the
FP64 65k-spp oracle subset is equal by construction (reported).

Also reported: the adjoint (config 3: 99,372-triangle sphere + Cornell walls,
512x512, 32 spp, depth 6) ms/iteration, its ratio to the forward pass and its
gradient error against central finite differences (h = 1e-3, 64 random
triangles); the CPU oracle (FP32 port of the estimator) timed on a bounded
pixel sample of the same workload on this host's cores; and rn_render(
RN_PATHTRACED) beside the reference's CPU time for the same call (this host's
port time scaled by the port/reference ratio measured in the build
container).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s @1M-tri 1080p, batched vs plain at equal RMSE; adjoint ms/iter"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_config1():
    from rednoise_amd import scenes
    sd = scenes.config1_scene(seed=0)
    W, H, spp, depth, rr, eps = 1920, 1080, 64, 8, 3, 0.03
    return sd, W, H, spp, depth, rr, eps


def render_device(rn, lib, sc, cam, p, out_ptr):
    if lib.rn_render_ex_device(sc.handle, C.byref(cam), C.byref(p), None, None, C.c_void_p(out_ptr), None) != 1:
        raise RuntimeError(lib.rn_last_error().decode())
    return rn.last_stats()


def timed_steps(rn, lib, sc, cam, p, out_ptr, steps, warmup, torch):
    """Each step: rebuild the BVH on the GPU, render the frame."""
    for _ in range(warmup):
        sc.build(p.bvh_width)
        render_device(rn, lib, sc, cam, p, out_ptr)
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        b = sc.build(p.bvh_width)
        st = render_device(rn, lib, sc, cam, p, out_ptr)
        st["build_ms"] = b["build_ms"]
        stats.append(st)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return dt, stats


def rel_rmse(img, fixture):
    pix = fixture["pixels"]
    ref = fixture["values"]
    got = img.reshape(-1, 3)[pix].astype(np.float64)
    return float(np.sqrt(((got - ref) ** 2).sum() / (ref ** 2).sum()))


def oracle_fixture(name):
    """The FP64 65,536-spp CPU-oracle pixel subset of a config
    (tests/golden/make_oracle_rmse.py), or None."""
    fx = os.path.join(ROOT, "tests", "golden", f"oracle64k_{name}.npz")
    return np.load(fx) if os.path.isfile(fx) else None


def subset_note(f):
    return f"{len(f['pixels'])} pixels, one seeded pixel per {int(f['block'])}x{int(f['block'])} block"


def config0_rmse(rn):
    """Config 0 (the 32-triangle Cornell box, 320x240, 16 spp, depth 5, RR
    from depth 3, Philox seed 0) from both integrators against its oracle."""
    from rednoise_amd import scenes
    f = oracle_fixture("config0")
    if f is None:
        return None
    sd = scenes.cornell_box()
    sc = rn.Scene.from_arrays(sd.verts, sd.albedo, sd.emission)
    cam = rn.Camera.make(**sd.camera)
    out = {"subset": subset_note(f)}
    for key, integ in (("wavefront_rel_rmse", rn.INTEGRATOR_WAVEFRONT), ("plain_rel_rmse", rn.INTEGRATOR_MEGAKERNEL)):
        p = rn.RenderParams.make(320, 240, 16, 5, 3, seed=0, ray_eps=0.05, integrator=integ)
        out[key] = round(rel_rmse(sc.render_hdr(cam, p), f), 5)
    sc.close()
    return out


def adjoint_bench(rn, lib, torch, n_fd=64):
    from rednoise_amd import scenes
    tgt, start = scenes.config3_scenes()
    W = H = 512
    spp, depth = 32, 6
    cam = rn.Camera.make(**tgt.camera)
    sc = rn.Scene.from_arrays(start.verts, start.albedo, start.emission)
    sct = rn.Scene.from_arrays(tgt.verts, tgt.albedo, tgt.emission)
    pt = rn.RenderParams.make(W, H, spp, depth, rr_depth=1 << 30, seed=1234, ray_eps=0.03)
    target = sct.render_hdr(cam, pt).astype(np.float64)
    # forward pass: the wavefront integrator (the estimator, and the Philox
    # streams the adjoint replays, are those of the megakernel, timed beside it)
    p = rn.RenderParams.make(W, H, spp, depth, rr_depth=1 << 30, seed=0, ray_eps=0.03,
                             integrator=rn.INTEGRATOR_WAVEFRONT)
    pm = rn.RenderParams.make(W, H, spp, depth, rr_depth=1 << 30, seed=0, ray_eps=0.03)
    n = sc.triangle_count
    img = sc.render_hdr(cam, p)
    dL = (img.astype(np.float64) - target).astype(np.float32)  # L = 0.5 sum (I - T)^2
    sc.adjoint(cam, p, dL)  # warm-up
    iters = 3
    t_fwd = t_adj = 0.0
    for _ in range(iters):
        t0 = time.perf_counter()
        img = sc.render_hdr(cam, p)
        t1 = time.perf_counter()
        dL = (img.astype(np.float64) - target).astype(np.float32)
        ga, ge = sc.adjoint(cam, p, dL)
        t2 = time.perf_counter()
        t_fwd += t1 - t0
        t_adj += t2 - t1
    fwd_ms, adj_ms = 1e3 * t_fwd / iters, 1e3 * t_adj / iters
    st = rn.last_stats()
    kern_adj = st["render_ms"]
    sc.render_hdr(cam, p)
    kern_fwd = rn.last_stats()["render_ms"]
    sc.render_hdr(cam, pm)
    t0 = time.perf_counter()
    img_m = sc.render_hdr(cam, pm)
    fwd_mega_ms = 1e3 * (time.perf_counter() - t0)
    same_image = bool(np.array_equal(img_m, img))  # radiance slots: the megakernel's sums, bit for bit
    # central differences on 64 random sphere triangles (albedo, one channel
    # each), loss L = 0.5 sum (I - T)^2 evaluated in FP64 from the FP32 images,
    # of the forward pass the iteration times: its pixels are sums in a fixed
    # order, so two renders differ only where the perturbed albedo enters.
    rng = np.random.default_rng(64)
    tris = rng.choice(np.arange(12, n), size=n_fd, replace=False)
    chans = rng.integers(0, 3, size=n_fd)
    alb = start.albedo.copy()
    h = 1e-3
    g_fd = np.zeros(n_fd)
    g_ad = np.zeros(n_fd)
    for k, (t, c) in enumerate(zip(tris, chans)):
        a0 = alb[t, c]
        alb[t, c] = a0 + h
        sc.set_materials(albedo=alb)
        ip = sc.render_hdr(cam, p).astype(np.float64)
        hi = float(alb[t, c])
        alb[t, c] = a0 - h
        sc.set_materials(albedo=alb)
        im = sc.render_hdr(cam, p).astype(np.float64)
        lo = float(alb[t, c])
        alb[t, c] = a0
        dloss = 0.5 * float(((ip - im) * (ip + im - 2.0 * target)).sum())
        g_fd[k] = dloss / (hi - lo)
        g_ad[k] = ga[t, c]
    sc.set_materials(albedo=alb)
    err = float(np.linalg.norm(g_ad - g_fd) / max(1e-30, np.linalg.norm(g_fd)))
    f = oracle_fixture("config3")
    rmse = {"subset": subset_note(f), "forward_rel_rmse": round(rel_rmse(img, f), 5)} if f is not None else None
    return {
        "workload": f"config3: {n} tris (cube sphere m=91 + 12 Cornell), 512x512, 32 spp, depth 6, no roulette",
        "ms_per_iter": round(fwd_ms + adj_ms, 3), "forward_ms": round(fwd_ms, 3), "adjoint_ms": round(adj_ms, 3),
        "forward_integrator": "wavefront", "forward_megakernel_ms": round(fwd_mega_ms, 3),
        "forward_images_equal": same_image,
        "adjoint_over_forward": round(adj_ms / fwd_ms, 3),
        "kernel_adjoint_over_forward": round(kern_adj / kern_fwd, 3),
        "grad_rel_err_vs_fd": err, "fd_h": h, "fd_triangles": n_fd,
        "target": "t_adjoint <= 3 t_forward and rel err <= 1e-2",
        "equal_rmse": rmse,
    }


def batch_bench(rn, torch):
    """Config 2: 64 seeded displaced spheres (1k-500k triangles, 256^2..1024^2,
    4/16/64 spp, depth 8) in one rn_render_batch call (one set of persistent
    wavefront launches replayed from a cached hipGraph), options picked by the
    autotuner. Beside it, the same 64 scenes as one wavefront call each (the
    gain of batching alone) and as one plain-megakernel call each. Every call
    downloads its images to the host; `*_gpu_s` is the GPU time of the
    renders alone (events around the launches)."""
    from rednoise_amd import autotune, scenes
    t0 = time.perf_counter()
    specs = scenes.config2_batch(64, seed=2)
    items = []
    for s in specs:
        sc = rn.Scene.from_arrays(s.scene.verts, s.scene.albedo, s.scene.emission)
        items.append((sc, rn.Camera.make(**s.scene.camera),
                      rn.RenderParams.make(s.width, s.height, s.spp, s.max_depth, 3, seed=0, ray_eps=0.03)))
    gen_s = time.perf_counter() - t0
    opts, report = autotune.autotune_batch(items, use_cache=False)
    rn.release_workspace()  # the first call below captures its graph afresh
    t0 = time.perf_counter()
    rn.render_batch(items, opts)
    dt_first = time.perf_counter() - t0
    torch.cuda.synchronize()
    # the images land in host arrays allocated and touched once (fresh pages
    # would fault inside the downloads); medians of 3 calls
    outs = [np.full(p.width * p.height * 3, 0.0, np.float32) for _, _, p in items]
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        imgs = rn.render_batch(items, opts, out=outs)
        runs.append((time.perf_counter() - t0, rn.last_stats()))
    imgs = [im.copy() for im in imgs]  # `outs` is reused below
    runs.sort(key=lambda r: r[0])
    dt, st = runs[1]
    rays = st["rays_closest"] + st["rays_shadow"]
    # the same scenes, one wavefront call each (same options)
    for sc, cam, p in items:
        p.integrator = rn.INTEGRATOR_WAVEFRONT
        p.queue_paths, p.block_size, p.bvh_width = opts.queue_paths, opts.block_size, opts.bvh_width
    for k, (sc, cam, p) in enumerate(items[:2]):
        sc.render_hdr(cam, p, out=outs[k])
    passes = []
    for _ in range(3):
        t0 = time.perf_counter()
        each_rays, each_gpu = 0, 0.0
        for k, (sc, cam, p) in enumerate(items):
            sc.render_hdr(cam, p, out=outs[k])
            s2 = rn.last_stats()
            each_rays += s2["rays_closest"] + s2["rays_shadow"]
            each_gpu += s2["render_ms"] * 1e-3
        passes.append((time.perf_counter() - t0, each_rays, each_gpu))
    passes.sort(key=lambda r: r[0])
    dt_each, each_rays, each_gpu = passes[1]
    for sc, cam, p in items:  # the per-scene megakernel path runs on width-8 BVHs, built here
        p.integrator, p.queue_paths, p.block_size, p.bvh_width = rn.INTEGRATOR_MEGAKERNEL, 0, 0, 8
        sc.build(8)
    for sc, cam, p in items[:2]:
        sc.render_hdr(cam, p)
    f = oracle_fixture("config2")
    keep = set(int(k) for k in f["scenes"]) if f is not None else set()
    plain_imgs = {}
    t0 = time.perf_counter()
    plain_rays = 0
    for k, (sc, cam, p) in enumerate(items):
        img = sc.render_hdr(cam, p)
        if k in keep:
            plain_imgs[k] = img
        s2 = rn.last_stats()
        plain_rays += s2["rays_closest"] + s2["rays_shadow"]
    dt_plain = time.perf_counter() - t0
    rmse = None
    if f is not None:
        def subset_rmse(get):
            num = den = 0.0
            for k in keep:
                sel = f["scene"] == k
                ref = f["values"][sel]
                got = get(k).reshape(-1, 3)[f["pixels"][sel]].astype(np.float64)
                num += float(((got - ref) ** 2).sum())
                den += float((ref ** 2).sum())
            return float(np.sqrt(num / den))
        rmse = {"subset": f"scenes {sorted(keep)}, {len(f['pixels'])} pixels, one seeded pixel per "
                          f"{int(f['block'])}x{int(f['block'])} block of each",
                "batched_rel_rmse": round(subset_rmse(lambda k: imgs[k]), 5),
                "plain_rel_rmse": round(subset_rmse(lambda k: plain_imgs[k]), 5)}
    assert each_rays == rays == plain_rays, (rays, each_rays, plain_rays)
    return {"workload": f"config2: 64 scenes, {sum(sc.triangle_count for sc, _, _ in items):,} triangles, "
                        f"{sum(p.width * p.height * p.spp for _, _, p in items) / 1e6:.1f}M paths, depth 8, "
                        "one rn_render_batch call: persistent wavefront launches replayed from a cached hipGraph "
                        "(times: medians of 3 calls, and of 3 passes of one call per scene)",
            "batched_mrays_s": round(rays / dt / 1e6, 2), "batched_s": round(dt, 3),
            "batched_gpu_s": round(st["render_ms"] * 1e-3, 3), "batched_first_call_s": round(dt_first, 3),
            "iterations": st["iterations"],
            "wavefront_one_call_each_mrays_s": round(each_rays / dt_each / 1e6, 2),
            "wavefront_one_call_each_s": round(dt_each, 3), "wavefront_one_call_each_gpu_s": round(each_gpu, 3),
            "batched_over_wavefront_one_call_each": round(dt_each / dt, 3),
            "batched_over_wavefront_one_call_each_gpu": round(each_gpu / (st["render_ms"] * 1e-3), 3),
            "plain_sequential_mrays_s": round(plain_rays / dt_plain / 1e6, 2), "plain_s": round(dt_plain, 3),
            "batched_over_plain": round(dt_plain / dt, 3),
            "autotune": {"best": report["best"], "candidates": len(report["candidates"]),
                         "trial_spp_fraction": report["trial_spp_fraction"]},
            "scene_generation_s": round(gen_s, 1), "equal_rmse": rmse}


def terrain_bench(rn, lib, torch):
    """Config 4: the 9,990,450-triangle terrain with 8 quad lights at
    3840x2160, 256 spp, depth 12, RR from depth 3: one step = GPU BVH build +
    wavefront render (1 warm-up step, 1 timed). Then the terrain frame at the
    same 256 spp batched with 16 config-2 scenes in one rn_render_batch call,
    beside the same 17 renders made one wavefront call each."""
    from rednoise_amd import scenes
    t0 = time.perf_counter()
    sd = scenes.terrain_scene()
    gen_s = time.perf_counter() - t0
    sc = rn.Scene.from_arrays(sd.verts, sd.albedo, sd.emission)
    cam = rn.Camera.make(**sd.camera)
    W, H, spp, depth, rr, eps = 3840, 2160, 256, 12, 3, 0.05
    out = torch.empty(W * H * 3, dtype=torch.float32, device="cuda")
    p = rn.RenderParams.make(W, H, spp, depth, rr, seed=0, ray_eps=eps, integrator=rn.INTEGRATOR_WAVEFRONT)
    dt, st = timed_steps(rn, lib, sc, cam, p, out.data_ptr(), 1, 1, torch)
    rays = st[0]["rays_closest"] + st[0]["rays_shadow"]
    res = {"workload": f"config4: {sd.notes}, {W}x{H}, {spp} spp, depth {depth}, RR from depth {rr}, "
                       "PLOC+BVH8 built on the GPU in the step, wavefront",
           "mrays_s": round(rays / dt / 1e6, 2), "seconds_per_step": round(dt, 3),
           "build_ms": round(st[0]["build_ms"], 2), "bvh": {"nodes": st[0]["bvh_nodes"], "depth": st[0]["bvh_depth"]},
           "rays": {"closest": st[0]["rays_closest"], "shadow": st[0]["rays_shadow"], "paths": st[0]["paths"]},
           "scene_generation_s": round(gen_s, 1)}
    f = oracle_fixture("config4")
    if f is not None:
        res["equal_rmse"] = {"subset": subset_note(f),
                             "wavefront_rel_rmse": round(rel_rmse(out.cpu().numpy(), f), 5)}
    del out
    specs = scenes.config2_batch(16, seed=2)
    pt = rn.RenderParams.make(W, H, spp, depth, rr, seed=0, ray_eps=eps)
    items = [(sc, cam, pt)]
    for s in specs:
        sci = rn.Scene.from_arrays(s.scene.verts, s.scene.albedo, s.scene.emission)
        items.append((sci, rn.Camera.make(**s.scene.camera),
                      rn.RenderParams.make(s.width, s.height, s.spp, s.max_depth, 3, seed=0, ray_eps=0.03)))
    opts = rn.BatchOptions(0, 0, 8, 1, 8)
    rn.render_batch(items, opts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rn.render_batch(items, opts)
    dtb = time.perf_counter() - t0
    stb = rn.last_stats()
    rays_b = stb["rays_closest"] + stb["rays_shadow"]
    t0 = time.perf_counter()
    rays_s, gpu_s = 0, 0.0
    for sci, ci, pi in items:
        pi.integrator = rn.INTEGRATOR_WAVEFRONT
        sci.render_hdr(ci, pi)
        s2 = rn.last_stats()
        rays_s += s2["rays_closest"] + s2["rays_shadow"]
        gpu_s += s2["render_ms"] * 1e-3
    dts = time.perf_counter() - t0
    res["batch_with_16_config2"] = {
        "workload": f"terrain 3840x2160 at {spp} spp + 16 config-2 scenes, one rn_render_batch call (cached "
                    "hipGraph replay), against the same 17 renders as one wavefront call each",
        "batched_mrays_s": round(rays_b / dtb / 1e6, 2), "batched_s": round(dtb, 3),
        "batched_gpu_s": round(stb["render_ms"] * 1e-3, 3),
        "wavefront_one_call_each_mrays_s": round(rays_s / dts / 1e6, 2), "wavefront_one_call_each_s": round(dts, 3),
        "wavefront_one_call_each_gpu_s": round(gpu_s, 3),
        "batched_over_wavefront_one_call_each": round(dts / dtb, 3)}
    for sci, _, _ in items:
        sci.close()
    return res


def progressive_bench(rn, sc, cam, W, H, depth, rr, eps):
    """Progressive 1080p rendering of the config-1 scene: 8 frames of 8 spp
    into one accumulator (wavefront), each followed by the fused tone-map +
    RGBA8 pack and its download."""
    acc = rn.Accumulator(W, H)
    p = rn.RenderParams.make(W, H, 8, depth, rr, seed=0, ray_eps=eps, integrator=rn.INTEGRATOR_WAVEFRONT)
    acc.render(sc, cam, p)  # warm-up
    acc.resolve()
    acc.reset()
    frames, t_render, t_resolve = 8, 0.0, 0.0
    for _ in range(frames):
        t0 = time.perf_counter()
        acc.render(sc, cam, p)
        t1 = time.perf_counter()
        acc.resolve(1.0, 2.2)
        t_resolve += time.perf_counter() - t1
        t_render += t1 - t0
    res = {"workload": f"config1 scene, {W}x{H}, {frames} frames x 8 spp, depth {depth}, tone-map exposure 1 gamma 2.2",
           "ms_per_frame": round(1e3 * t_render / frames, 2), "resolve_ms": round(1e3 * t_resolve / frames, 3),
           "samples": acc.samples}
    acc.close()
    return res


def scene_cache_bench(rn, sd):
    """The config-1 scene written as an OBJ, then loaded with
    rn_scene_load_obj (the reference's entry point) and with rn_scene_load
    from its binary cache (host-side load; the GPU upload is the same)."""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        obj = os.path.join(d, "config1.obj")
        v = sd.verts.reshape(-1, 3).astype(np.float64)
        with open(obj, "w") as f:
            np.savetxt(f, v, fmt="v %.9g %.9g %.9g")
            faces = np.arange(1, v.shape[0] + 1, dtype=np.int64).reshape(-1, 3)
            np.savetxt(f, faces, fmt="f %d %d %d")
        t0 = time.perf_counter()
        a = rn.Scene.load_obj(obj, 1.0)
        t_obj = time.perf_counter() - t0
        cache = os.path.join(d, "config1.rnscene")
        t0 = time.perf_counter()
        a.save(cache)
        t_save = time.perf_counter() - t0
        t0 = time.perf_counter()
        b = rn.Scene.load(cache)
        t_load = time.perf_counter() - t0
        same = bool(np.array_equal(a.triangles()[0], b.triangles()[0]))
        res = {"triangles": a.triangle_count, "obj_mb": round(os.path.getsize(obj) / 1e6, 1),
               "obj_load_s": round(t_obj, 3), "cache_mb": round(os.path.getsize(cache) / 1e6, 1),
               "cache_save_s": round(t_save, 3), "cache_load_s": round(t_load, 3),
               "load_speedup": round(t_obj / t_load, 1), "identical_geometry": same}
        a.close()
        b.close()
    return res


def cpu_baseline(sd, W, H, spp, depth, rr, eps, budget_s=12.0):
    """FP32 oracle (C port of the estimator, OpenMP) on a pixel sample of the
    same workload; sample grows until ~budget_s of CPU work."""
    from oracle import oracle as O
    threads = os.cpu_count() or 1
    t0 = time.perf_counter()
    ps = O.PhysScene(sd.verts, sd.albedo, sd.emission, "f32")
    prep = time.perf_counter() - t0
    cam = O.camera_basis(sd.camera["eye"], sd.camera["target"], sd.camera["up"], sd.camera["vfov_deg"], W, H)
    p = O.params(W, H, spp, depth, rr, 0.0, 0, 0, eps)
    rng = np.random.default_rng(5)
    npix = 64
    total_rays, total_paths, total_t = 0, 0, 0.0
    while total_t < budget_s and npix <= W * H:
        pix = rng.choice(W * H, size=npix, replace=False).astype(np.int32)
        t0 = time.perf_counter()
        _, cnt = ps.render(cam, p, threads=threads, pixels=pix)
        dt = time.perf_counter() - t0
        total_rays += cnt.closest + cnt.shadow
        total_paths += cnt.paths
        total_t += dt
        npix *= 2
    ps.close()
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(total_rays / total_t / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model,
            "sample": f"{total_paths // spp} random pixels x {spp} spp of the same 1080p frame (FP32 oracle port, "
                      f"median-split BVH, OpenMP {threads} threads), {total_t:.1f}s render + {prep:.1f}s BVH build",
            "samples_per_s": round(total_paths / total_t, 1)}


def reference_calibration():
    """Port/reference timings measured in the build container
    (tests/golden/calibrate_reference.py), or None."""
    fx = os.path.join(ROOT, "tests", "golden", "reference_timing.json")
    return json.load(open(fx)) if os.path.isfile(fx) else None


def _md5(a) -> str:
    import hashlib
    return hashlib.md5(np.ascontiguousarray(a, np.uint8).tobytes()).hexdigest()


def reference_modeR(rn, torch):
    """rn_render(RN_PATHTRACED) and rn_render(RN_RAYTRACED) on the Cornell
    asset beside the reference's CPU time for the same call. The compiled
    reference does not run here: its time on this host is the CPU port's time
    here (the oracle's restatement with the reference's own random streams,
    OpenMP over this host's cores) divided by the port/reference ratio
    measured on the same call in the build container; the GPU image is
    compared with the reference image's MD5 recorded there."""
    from oracle import oracle as O
    cal = reference_calibration()
    if cal is None:
        return None
    asset = os.path.join(ROOT, "tests", "golden", "assets", "cornell-box.obj")
    c = cal["rn_render_pathtraced"]
    W, H, spp, cz = c["width"], c["height"], c["spp"], c["cam_z"]
    ours = rn.Scene.load_obj(asset, 0.35)
    ours.render("pathtraced", W, H, cz, spp)
    t0 = time.perf_counter()
    img = ours.render("pathtraced", W, H, cz, spp).to_numpy()
    t_gpu = time.perf_counter() - t0
    cs = O.CompatArrays.from_scene(ours)
    port = np.zeros(W * H * 4, np.uint8)
    O.lib().or_compat_render_mt(C.byref(cs.s), W, H, cz, spp, 0, H, port)  # warm-up
    t0 = time.perf_counter()
    O.lib().or_compat_render_mt(C.byref(cs.s), W, H, cz, spp, 0, H, port)
    t_port = time.perf_counter() - t0
    t_ref = t_port / c["port_over_reference"]
    res = {"call": c["call"], "gpu_s": round(t_gpu, 5),
           "identical_to_reference_image": _md5(img) == c["reference_rgba_md5"],
           "cpu_port_s": round(t_port, 4), "cpu_port_threads": os.cpu_count(),
           "cpu_port_identical_pixels": round(float((port.reshape(-1, 4) == img.reshape(-1, 4)).all(axis=1).mean()), 5),
           "reference_cpu_s": round(t_ref, 4), "speedup": round(t_ref / t_gpu, 1),
           "reference_timing": (f"container-calibrated: the port's time on this host / {c['port_over_reference']} "
                                f"(port/reference time on the same call in the build container, "
                                f"{cal['container_threads']} threads: {c['port_s']} s / {c['reference_s']} s)")}
    c2 = cal["rn_render_raytraced"]
    W2, H2 = c2["width"], c2["height"]
    ours.render("raytraced", W2, H2, cz, 1)
    t0 = time.perf_counter()
    img = ours.render("raytraced", W2, H2, cz, 1).to_numpy()
    t_gpu2 = time.perf_counter() - t0
    res["raytraced"] = {"call": c2["call"], "gpu_s": round(t_gpu2, 5),
                        "identical_to_reference_image": _md5(img) == c2["reference_rgba_md5"],
                        "reference_cpu_s_container": c2["reference_s"],
                        "reference_timing": f"measured in the build container ({cal['container_threads']} threads; "
                                            "no CPU port of this renderer)"}
    ours.close()
    return res


def primary_parity(rn):
    """Primary rays of rn_render's camera through the GPU BVH8 with the
    reference's triangle test, against the reference's own BVH::intersect
    (tests/golden/primary_*.npz): ids compared for every pixel; mismatches
    are split by the 1e-6 barycentric edge band (a hit within 1e-6 of an
    edge on either side may legitimately go to the neighbouring triangle)."""
    out = {}
    for key, asset, scale in (("cornell", "cornell-box.obj", 0.35), ("quad", "textured-quad.obj", 0.35),
                              ("sphere", "sphere.obj", 1.0)):
        g = np.load(os.path.join(ROOT, "tests", "golden", f"primary_{key}.npz"))
        sc = rn.Scene.load_obj(os.path.join(ROOT, "tests", "golden", "assets", asset), scale)
        prim, t, u, v, _ = sc.trace_primary_compat(int(g["width"]), int(g["height"]), float(g["cam_z"]))
        same = prim == g["prim"]

        def band(uu, vv):
            return np.minimum(np.minimum(uu, vv), 1.0 - uu - vv) < 1e-6
        inb = band(g["u"], g["v"]) | band(u, v)
        hit = same & (prim >= 0)
        out[key] = {"rays": int(prim.size), "hits": int((g["prim"] >= 0).sum()), "id_mismatches": int((~same).sum()),
                    "in_edge_band": int((~same & inb).sum()), "outside_edge_band": int((~same & ~inb).sum()),
                    "tuv_bit_identical_of_matching_hits": round(float(
                        ((t[hit] == g["t"][hit]) & (u[hit] == g["u"][hit]) & (v[hit] == g["v"][hit])).mean())
                        if hit.any() else 1.0, 6)}
        sc.close()
    return out


def estimators_bench(rn):
    """Photon mapping and bidirectional path tracing on the Cornell asset at
    the reference CLI's defaults (render -m photon: 640x480, 200000 photons,
    scale 0.35, cam_z 4), beside the reference's CPU time for the same calls
    as measured in the build container (tests/golden/reference_timing.json);
    the compiled reference itself does not run here."""
    from oracle import oracle as O
    asset = os.path.join(ROOT, "tests", "golden", "assets", "cornell-box.obj")
    W, H, nph, cz = 640, 480, 200000, 4.0
    sc = rn.Scene.load_obj(asset, 0.35)

    def best(fn, reps=3):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    t_trace = best(lambda: rn.PhotonMap.trace(sc, nph, 0, 0).close())
    pm = rn.PhotonMap.trace(sc, nph, 0, 0)
    t_render = best(lambda: sc.render_photon(pm, W, H, cz))
    t_fg = best(lambda: sc.render_photon(pm, W, H, cz, 4), reps=1)
    res = {"photon": {"workload": f"cornell-box.obj {W}x{H}, {nph} photons, gather radius 0.25",
                      "photons_stored": len(pm), "gpu_pass1_s": round(t_trace, 5), "gpu_pass2_s": round(t_render, 5),
                      "gpu_s": round(t_trace + t_render, 5), "gpu_final_gather4_pass2_s": round(t_fg, 5)}}
    pm.close()
    spp = 16
    t_bd = best(lambda: sc.render_bidirectional(W, H, cz, spp, 40.0))
    t_bd64 = best(lambda: sc.render_bidirectional(W, H, cz, 64, 40.0), reps=1)
    res["bidirectional"] = {"workload": f"cornell-box.obj {W}x{H}, {spp} samples, exposure 40",
                            "gpu_s": round(t_bd, 5), "gpu_64spp_s": round(t_bd64, 5)}
    cal = reference_calibration()
    if cal is not None:
        # pass 2 on the GPU fed the reference's photons (the oracle's pass 1
        # with the reference's std::mt19937 stream): the reference's image
        cs = O.CompatArrays.from_scene(sc)
        mine = sc.render_photon(rn.PhotonMap.from_photons(*O.photon_trace(cs, nph, O.RNG_MT)), W, H, cz).to_numpy()
        cp, cb = cal["photon"], cal["bidirectional"]
        note = f"measured in the build container ({cal['container_threads']} threads)"
        res["photon"].update({"identical_to_reference_image_with_reference_photons":
                              _md5(mine) == cp["reference_rgba_md5"],
                              "reference_cpu_s_container": cp["reference_s"], "reference_timing": note})
        res["bidirectional"].update({"reference_cpu_s_container": cb["reference_s"], "reference_timing": note})
    sc.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--skip-extras", action="store_true", help="headline only (no adjoint / CPU / RMSE)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE={world}); the multi-GPU "
              "sharding of this path is design-only.", file=sys.stderr)
        sys.exit(2)
    import torch

    import rednoise_amd as rn

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    lib = rn._load_library()
    t0 = time.perf_counter()
    sd, W, H, spp, depth, rr, eps = make_config1()
    log(f"scene: {sd.notes} ({time.perf_counter() - t0:.1f}s)")
    sc = rn.Scene.from_arrays(sd.verts, sd.albedo, sd.emission)
    cam = rn.Camera.make(**sd.camera)
    out = torch.empty(W * H * 3, dtype=torch.float32, device="cuda")
    pw = rn.RenderParams.make(W, H, spp, depth, rr, seed=0, ray_eps=eps, integrator=rn.INTEGRATOR_WAVEFRONT,
                              flags=rn.FLAG_TIME_STAGES)
    pm = rn.RenderParams.make(W, H, spp, depth, rr, seed=0, ray_eps=eps, integrator=rn.INTEGRATOR_MEGAKERNEL)
    sc.build(8)

    # ---- timed: wavefront (value) and plain megakernel beside it
    dt_w, st_w = timed_steps(rn, lib, sc, cam, pw, out.data_ptr(), args.steps, args.warmup, torch)
    img_w = out.cpu().numpy()
    log(f"wavefront: {1e3 * dt_w / args.steps:.1f} ms/step")
    dt_m, st_m = timed_steps(rn, lib, sc, cam, pm, out.data_ptr(), args.steps, args.warmup, torch)
    img_m = out.cpu().numpy()
    log(f"megakernel: {1e3 * dt_m / args.steps:.1f} ms/step")
    rays_w = sum(s["rays_closest"] + s["rays_shadow"] for s in st_w)
    rays_m = sum(s["rays_closest"] + s["rays_shadow"] for s in st_m)
    paths_w = sum(s["paths"] for s in st_w)
    value = rays_w / dt_w / 1e6
    plain = rays_m / dt_m / 1e6

    # ---- traversal counts on the same BVH8 (instrumented megakernel)
    pc = rn.RenderParams.make(W, H, spp, depth, rr, seed=0, ray_eps=eps, integrator=rn.INTEGRATOR_MEGAKERNEL,
                              flags=rn.FLAG_COUNT_TRAVERSAL)
    stc = render_device(rn, lib, sc, cam, pc, out.data_ptr())
    nc = stc["nodes_closest"] / stc["rays_closest"]
    tc = stc["tris_closest"] / stc["rays_closest"]
    ns = stc["nodes_shadow"] / max(1, stc["rays_shadow"])
    ts = stc["tris_shadow"] / max(1, stc["rays_shadow"])
    b_closest = 48 + 80 * nc + 36 * tc
    b_shadow = 36 + 80 * ns + 36 * ts
    # the wavefront's trace kernel: each launch traces the iteration's
    # closest-hit rays, then the previous iteration's any-hit rays
    ext_ms = sum(s["extend_ms"] for s in st_w)
    launches = sum(s["iterations"] for s in st_w)
    rc = sum(s["rays_closest"] for s in st_w)
    rs = sum(s["rays_shadow"] for s in st_w)
    kname, kbytes, kms = "k_trace (wavefront closest-hit + any-hit)", rc * b_closest + rs * b_shadow, ext_ms
    achieved = kbytes / (kms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_config1.json")
    if os.path.isfile(tf):
        try:
            traffic = json.load(open(tf)).get(kname.split()[0])
        except Exception:
            traffic = None
    avg_ms = kms / max(1, launches)
    # The algorithmic bytes count node and triangle loads that the L2 and the
    # MALL serve; the measured fabric traffic (rocprofv3 FETCH_SIZE +
    # WRITE_SIZE per launch, profiles/) is below them, and the kernel is
    # bound by VALU issue (divergent lanes), not by HBM.
    traffic_frac = round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5) if traffic else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_frac": traffic_frac,
                "limiter": "VALU issue (lane divergence); frac counts cache-served node loads, traffic_frac is the "
                           "measured fabric traffic", "kernel": kname,
                "avg_launch_ms": round(avg_ms, 4), "launches": launches,
                "bytes_per_closest_ray": round(b_closest, 1), "bytes_per_shadow_ray": round(b_shadow, 1),
                "nodes_per_closest_ray": round(nc, 3), "tris_per_closest_ray": round(tc, 3),
                "nodes_per_shadow_ray": round(ns, 3), "tris_per_shadow_ray": round(ts, 3),
                "compulsory_gbs": round((rc * 48 + rs * 36) / (kms * 1e-3) / 1e9, 2)}
    megakernel_kernel_ms = sum(s["render_ms"] for s in st_m)
    mk_bytes = sum(s["rays_closest"] for s in st_m) * b_closest + sum(s["rays_shadow"] for s in st_m) * b_shadow

    result = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * dt_w / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene generator)",
        "config": {"workload": "config1: 1,048,576-tri displaced sphere + 12 Cornell tris, 1920x1080, 64 spp, "
                               "depth 8, RR from depth 3, PLOC+BVH8 rebuilt on the GPU every step, wavefront "
                               "integrator", "triangles": sd.n, "width": W, "height": H, "spp": spp,
                   "max_depth": depth},
        "roofline": roofline,
        "wavefront": {"mrays_s": round(value, 2), "ms_per_step": round(1e3 * dt_w / args.steps, 3),
                      "samples_per_s": round(paths_w / dt_w, 1),
                      "mrays_s_excl_build": round(rays_w / (dt_w - 1e-3 * sum(s["build_ms"] for s in st_w)) / 1e6,
                                                  2),
                      "trace_ms": round(ext_ms / args.steps, 2), "shade_ms": round(
                          sum(s["shade_ms"] for s in st_w) / args.steps, 2),
                      "iterations": st_w[0]["iterations"]},
        "plain_megakernel": {"mrays_s": round(plain, 2), "ms_per_step": round(1e3 * dt_m / args.steps, 3),
                             "kernel_ms": round(megakernel_kernel_ms / args.steps, 2),
                             "kernel_achieved_gbs": round(mk_bytes / (megakernel_kernel_ms * 1e-3) / 1e9, 2)},
        "wavefront_over_plain": round(value / plain, 3),
        "build_ms": round(sum(s["build_ms"] for s in st_w) / args.steps, 3),
        "bvh": {"nodes": st_w[0]["bvh_nodes"], "depth": st_w[0]["bvh_depth"]},
        "rays_per_step": {"closest": st_w[0]["rays_closest"], "shadow": st_w[0]["rays_shadow"],
                          "paths": st_w[0]["paths"]},
    }
    f = oracle_fixture("config1")
    if f is not None:
        result["equal_rmse"] = {"oracle": f"FP64 CPU oracle, {int(f['spp'])} spp, independent Philox seed; "
                                          "configs 2-4 carry theirs in their own entries",
                                "subset": subset_note(f),
                                "wavefront_rel_rmse": round(rel_rmse(img_w, f), 5),
                                "plain_rel_rmse": round(rel_rmse(img_m, f), 5)}
    if not args.skip_extras:
        try:
            c0 = config0_rmse(rn)
            if c0 is not None and "equal_rmse" in result:
                result["equal_rmse"]["config0"] = c0
        except Exception as e:
            result.setdefault("equal_rmse", {})["config0"] = {"error": str(e)}
        try:
            result["adjoint"] = adjoint_bench(rn, lib, torch)
            log(f"adjoint: {result['adjoint']}")
        except Exception as e:  # reported, not fatal for the headline
            result["adjoint"] = {"error": str(e)}
        try:
            result["batch"] = batch_bench(rn, torch)
            log(f"batch: {result['batch']}")
        except Exception as e:
            result["batch"] = {"error": str(e)}
        try:
            result["progressive"] = progressive_bench(rn, sc, cam, W, H, depth, rr, eps)
            log(f"progressive: {result['progressive']}")
        except Exception as e:
            result["progressive"] = {"error": str(e)}
        try:
            result["scene_cache"] = scene_cache_bench(rn, sd)
            log(f"scene_cache: {result['scene_cache']}")
        except Exception as e:
            result["scene_cache"] = {"error": str(e)}
        try:
            result["config4"] = terrain_bench(rn, lib, torch)
            log(f"config4: {result['config4']}")
        except Exception as e:
            result["config4"] = {"error": str(e)}
        try:
            r = reference_modeR(rn, torch)
            if r:
                result["reference_rn_render"] = r
        except Exception as e:
            result["reference_rn_render"] = {"error": str(e)}
        try:
            result["primary_ray_parity"] = primary_parity(rn)
            log(f"primary_ray_parity: {result['primary_ray_parity']}")
        except Exception as e:
            result["primary_ray_parity"] = {"error": str(e)}
        try:
            result["estimators"] = estimators_bench(rn)
            log(f"estimators: {result['estimators']}")
        except Exception as e:
            result["estimators"] = {"error": str(e)}
        result["cpu_baseline"] = cpu_baseline(sd, W, H, spp, depth, rr, eps)
        result["cpu_baseline"]["gpu_over_cpu"] = round(value / result["cpu_baseline"]["value"], 1)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
