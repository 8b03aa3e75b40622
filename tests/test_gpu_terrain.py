"""The 10M-triangle terrain (2236 x 2236 vertex grid, 9,990,450 triangles
plus 8 quad lights): the largest scene the configurations name. The GPU
build and render at full size are checked against the FP32 oracle on the
same Philox streams at a seeded pixel subset, and the wavefront against the
megakernel (same paths, same rays)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def terrain():
    from rednoise_amd import scenes
    return scenes.terrain_scene()


def test_terrain_full_size_vs_oracle(rn, oracle, terrain):
    sd = terrain
    assert sd.n == 9_990_450 + 16
    sc = rn.Scene.from_arrays(sd.verts, sd.albedo, sd.emission)
    info = sc.build(8)
    assert info["nodes"] > sd.n // 24 and info["depth"] >= 8
    W, H, spp, depth, rr, eps = 384, 216, 4, 12, 3, 0.05
    cam = rn.Camera.make(**sd.camera)
    p = rn.RenderParams.make(W, H, spp, depth, rr, seed=0, ray_eps=eps, integrator=rn.INTEGRATOR_WAVEFRONT)
    img = sc.render_hdr(cam, p)
    st = rn.last_stats()
    assert st["paths"] == W * H * spp
    ps = oracle.PhysScene(sd.verts, sd.albedo, sd.emission, "f32")
    oc = oracle.camera_basis(sd.camera["eye"], sd.camera["target"], sd.camera["up"], sd.camera["vfov_deg"], W, H)
    op = oracle.params(W, H, spp, depth, rr, 0.0, 0, 0, eps)
    pix = np.random.default_rng(1).choice(W * H, size=2048, replace=False).astype(np.int32)
    cpu, _ = ps.render(oc, op, pixels=pix)
    ps.close()
    gpu = img.reshape(-1, 3)[pix]
    close = np.abs(gpu - cpu) <= 1e-3 * np.maximum(1.0, np.abs(cpu))
    assert close.mean() > 0.98, close.mean()
    assert cpu.mean() > 0.01  # the subset sees lit terrain, not only sky


def test_terrain_wavefront_matches_megakernel(rn, terrain):
    sd = terrain
    sc = rn.Scene.from_arrays(sd.verts, sd.albedo, sd.emission)
    cam = rn.Camera.make(**sd.camera)
    p = rn.RenderParams.make(256, 144, 2, 12, 3, seed=3, ray_eps=0.05)
    mk = sc.render_hdr(cam, p)
    s_mk = rn.last_stats()
    p.integrator = rn.INTEGRATOR_WAVEFRONT
    wf = sc.render_hdr(cam, p)
    s_wf = rn.last_stats()
    assert np.allclose(wf, mk, rtol=1e-4, atol=1e-5)
    assert (s_wf["rays_closest"], s_wf["rays_shadow"]) == (s_mk["rays_closest"], s_mk["rays_shadow"])
