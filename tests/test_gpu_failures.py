"""Error paths of the C ABI, forced from the host with RN_FAULT_INJECT (the
reference's contract: every entry point returns 0 or NULL on failure,
src/capi.cpp:22-34,46-47; here rn_last_error() also says why). Each failure is
raised before a kernel launch or between synchronised launches, so no kernel
is interrupted; after the switch is cleared the same scenes render exactly
as scenes that never failed.

- alloc: every device and pinned allocation fails. Every entry point that
  has to allocate reports it; rn_render_ex_device_noalloc, which allocates
  nothing, still renders.
- bvh_depth: the guard that keeps a too-deep BVH from the traversal kernels
  (their stack holds the LDS entries plus the spill stack) rejects the build.
- wavefront_iterations: the wavefront's bound on its iterations (a queue
  accounting error would otherwise loop) fires after one launch group.
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import ASSETS

pytestmark = pytest.mark.gpu


@pytest.fixture
def cornell(rn):
    from rednoise_amd import scenes
    sd = scenes.cornell_box()
    return sd, rn.Camera.make(**sd.camera)


def _params(rn, integ=0, **kw):
    return rn.RenderParams.make(40, 30, 4, 5, 3, seed=2, ray_eps=0.05, integrator=integ, **kw)


def test_injected_allocation_failures(rn, cornell, monkeypatch):
    import torch
    lib = rn._load_library()
    sd, cam = cornell
    # reference images, and a built scene with device outputs for the
    # allocation-free entry point, before anything fails
    ok_scene = rn.Scene.from_arrays(sd.verts, sd.albedo, sd.emission)
    ok_scene.build(8)
    want_mk = ok_scene.render_hdr(cam, _params(rn))
    want_wf = ok_scene.render_hdr(cam, _params(rn, rn.INTEGRATOR_WAVEFRONT))
    obj = rn.Scene.load_obj(os.path.join(ASSETS, "cornell-box.obj"), 0.35)
    want_rgba = obj.render("pathtraced", 32, 24, 4.0, 4).to_numpy()
    out_mk = torch.zeros((30, 40, 3), device="cuda")
    out_wf = torch.zeros((30, 40, 3), device="cuda")
    obj.close()
    rn.release_workspace()  # the next calls must allocate their work buffers

    fresh = rn.Scene.from_arrays(sd.verts, sd.albedo, sd.emission)  # host data only
    fresh_obj = rn.Scene.load_obj(os.path.join(ASSETS, "cornell-box.obj"), 0.35)
    monkeypatch.setenv("RN_FAULT_INJECT", "alloc")
    calls = {
        "rn_scene_build": lambda: fresh.build(8),
        "rn_render_ex (megakernel)": lambda: fresh.render_hdr(cam, _params(rn)),
        "rn_render_ex (wavefront)": lambda: fresh.render_hdr(cam, _params(rn, rn.INTEGRATOR_WAVEFRONT)),
        "rn_render_ex (built scene)": lambda: ok_scene.render_hdr(cam, _params(rn)),
        "rn_render": lambda: fresh_obj.render("pathtraced", 32, 24, 4.0, 4),
        "rn_render RAYTRACED": lambda: fresh_obj.render("raytraced", 32, 24, 4.0, 1),
        "rn_render_batch": lambda: rn.render_batch([(fresh, cam, _params(rn))], rn.BatchOptions(0, 0, 8, 0, 8)),
        "rn_adjoint": lambda: fresh.adjoint(cam, _params(rn), np.ones((30, 40, 3), np.float32)),
        "rn_trace_rays": lambda: fresh.trace_rays(np.zeros((4, 3), np.float32), np.ones((4, 3), np.float32)),
        "rn_accum_create": lambda: rn.Accumulator(40, 30),
        "rn_photon_map_trace": lambda: rn.PhotonMap.trace(fresh_obj, 1000),
        "rn_render_bidirectional": lambda: fresh_obj.render_bidirectional(32, 24, 4.0, 2),
    }
    for name, call in calls.items():
        with pytest.raises(RuntimeError, match="injected allocation failure"):
            call()
        print(f"{name}: {lib.rn_last_error().decode()}")
    keys = np.zeros(4, np.uint32)
    assert lib.rn_radix_sort_pairs(keys.ctypes.data, keys.ctypes.data, 4, 32) == 0
    assert b"injected" in lib.rn_last_error()
    # the allocation-free entry point renders regardless
    for integ, out, want in ((rn.INTEGRATOR_MEGAKERNEL, out_mk, want_mk), (rn.INTEGRATOR_WAVEFRONT, out_wf, want_wf)):
        p = _params(rn, integ)
        assert lib.rn_render_ex_device_noalloc(ok_scene.handle, C.byref(cam), C.byref(p), None, None,
                                               C.c_void_p(out.data_ptr()), None) == 1, lib.rn_last_error()
        torch.cuda.synchronize()
        assert np.allclose(out.cpu().numpy(), want, rtol=1e-5, atol=1e-6)

    monkeypatch.delenv("RN_FAULT_INJECT")
    # nothing was left half-built: the same scenes now render as usual
    fresh.build(8)
    assert np.array_equal(fresh.render_hdr(cam, _params(rn)), want_mk)
    assert np.allclose(fresh.render_hdr(cam, _params(rn, rn.INTEGRATOR_WAVEFRONT)), want_wf, rtol=1e-5, atol=1e-6)
    assert np.array_equal(fresh_obj.render("pathtraced", 32, 24, 4.0, 4).to_numpy(), want_rgba)
    assert np.array_equal(ok_scene.render_hdr(cam, _params(rn)), want_mk)
    for s in (fresh, fresh_obj, ok_scene):
        s.close()


def test_bvh_depth_guard(rn, cornell, monkeypatch):
    sd, cam = cornell
    sc = rn.Scene.from_arrays(sd.verts, sd.albedo, sd.emission)
    monkeypatch.setenv("RN_FAULT_INJECT", "bvh_depth")
    with pytest.raises(RuntimeError, match="deeper than the traversal stack"):
        sc.build(8)
    for call in (lambda: sc.render_hdr(cam, _params(rn)), lambda: sc.trace_rays(np.zeros((1, 3), np.float32),
                                                                                 np.ones((1, 3), np.float32))):
        with pytest.raises(RuntimeError, match="deeper than the traversal stack"):
            call()
    monkeypatch.delenv("RN_FAULT_INJECT")
    assert sc.build(8)["depth"] >= 1
    assert np.isfinite(sc.render_hdr(cam, _params(rn))).all()
    sc.close()


def test_wavefront_iteration_bound(rn, cornell, monkeypatch):
    sd, cam = cornell
    sc = rn.Scene.from_arrays(sd.verts, sd.albedo, sd.emission)
    # 40 x 30 x 4 = 4,800 paths through a queue of 256: many refills, so the
    # render is not done after its first launch group
    p = _params(rn, rn.INTEGRATOR_WAVEFRONT, queue_paths=256)
    want = sc.render_hdr(cam, _params(rn))
    monkeypatch.setenv("RN_FAULT_INJECT", "wavefront_iterations")
    with pytest.raises(RuntimeError, match="iteration limit exceeded"):
        sc.render_hdr(cam, p)
    monkeypatch.delenv("RN_FAULT_INJECT")
    got = sc.render_hdr(cam, p)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-6)
    sc.close()
