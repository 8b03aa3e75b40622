"""The oracle's photon mapper and bidirectional path tracer (oracle/photon.c)
against the reference (CPU only).

With the reference's own random streams (std::mt19937, OR_RNG_MT) the
oracle's images must equal the compiled reference's renderPhotonMapped and
renderBidirectional images pixel for pixel (tests/golden/photon_<asset>.npz,
written by tests/golden/make_golden.py --photon): that pins photon emission
and tracing, the photon map, photonShade with and without the final gather,
directLight, both subpaths and the connection sum. With the Philox streams the
GPU uses (OR_RNG_PHILOX) the estimator is the same: per-tile means agree with
the mt19937 image within the spread of several Philox keys."""
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASSETS = os.path.join(ROOT, "tests", "golden", "assets")
CASES = [("cornell", "cornell-box.obj", 0.35), ("quad", "textured-quad.obj", 0.35), ("sphere", "sphere.obj", 1.0)]


def compat(rn, oracle, asset, scale):
    sc = rn.Scene.load_obj(os.path.join(ASSETS, asset), scale)
    return oracle.CompatArrays.from_scene(sc)


@pytest.mark.parametrize("key,asset,scale", CASES)
def test_oracle_photon_and_bdpt_equal_reference_images(rn, oracle, key, asset, scale):
    g = np.load(os.path.join(ROOT, "tests", "golden", f"photon_{key}.npz"))
    cs = compat(rn, oracle, asset, scale)
    H, W = g["photon_g0"].shape[:2]
    cz = float(g["cam_z"])
    ph = oracle.photon_trace(cs, int(g["num_photons"]), oracle.RNG_MT)
    for gr in (0, 3):
        img = oracle.photon_render(cs, ph, W, H, cz, gr, oracle.RNG_MT)
        assert np.array_equal(img, g[f"photon_g{gr}"]), (key, gr)
    img = oracle.bdpt_render(cs, W, H, cz, int(g["bdpt_samples"]), float(g["bdpt_exposure"]), oracle.RNG_MT)
    assert np.array_equal(img, g["bdpt"]), key


def test_oracle_gather_equals_reference_photon_map(oracle):
    ref = oracle.ref()
    if ref is None:
        pytest.skip("compiled reference not built (make -C oracle ref)")
    rng = np.random.default_rng(5)
    n, m = 20000, 4000
    pos = rng.uniform(-1.2, 1.2, (n, 3)).astype(np.float32)
    pos[::7] = np.round(pos[::7] * 4) / 4  # photons on cell boundaries
    pw = rng.uniform(0, 1e-3, (n, 3)).astype(np.float32)
    dr = rng.normal(size=(n, 3)).astype(np.float32)
    pts = rng.uniform(-1.3, 1.3, (m, 3)).astype(np.float32)
    nrm = rng.normal(size=(m, 3)).astype(np.float32)
    for cell, radius in ((0.25, 0.25), (0.25, 0.1), (0.4, 0.3)):
        mine = oracle.photon_gather((pos, pw, dr), cell, pts, nrm, radius)
        theirs = np.empty(3 * m, np.float32)
        ref.ref_photon_gather(n, pos.reshape(-1), pw.reshape(-1), dr.reshape(-1), cell, m, pts.reshape(-1),
                              nrm.reshape(-1), radius, theirs)
        assert np.array_equal(mine, theirs.reshape(-1, 3)), (cell, radius)
        assert (mine > 0).any()


def tile_means(img, t=8):
    H, W = img.shape[:2]
    a = img[: H // t * t, : W // t * t, :3].astype(np.float64)
    return a.reshape(H // t, t, W // t, t, 3).mean(axis=(1, 3))


def test_philox_streams_estimate_what_mt19937_does(rn, oracle):
    """The substitution of the random streams changes the noise, not the
    image: tile means of the reference's (mt19937) image lie within the
    spread of Philox renders with different keys."""
    cs = compat(rn, oracle, "cornell-box.obj", 0.35)
    W, H, cz = 48, 32, 4.0
    mt = oracle.photon_render(cs, oracle.photon_trace(cs, 20000, oracle.RNG_MT), W, H, cz, 0, oracle.RNG_MT)
    px = [oracle.photon_render(cs, oracle.photon_trace(cs, 20000, oracle.RNG_PHILOX, (k, 1)), W, H, cz, 0,
                               oracle.RNG_PHILOX, (k, 1)) for k in range(6)]
    tm = np.stack([tile_means(p) for p in px])
    mu, sd = tm.mean(0), tm.std(0, ddof=1)
    z = np.abs(tile_means(mt) - mu) / (sd + 0.5)
    assert z.max() < 6.0 and np.median(z) < 2.0, (z.max(), np.median(z))
    assert abs(mt[..., :3].mean() - mu.mean()) < 0.02 * mu.mean() + 0.5

    mtb = oracle.bdpt_render(cs, W, H, cz, 8, 40.0, oracle.RNG_MT)
    bx = [oracle.bdpt_render(cs, W, H, cz, 8, 40.0, oracle.RNG_PHILOX, (k, 2)) for k in range(6)]
    tb = np.stack([tile_means(b) for b in bx])
    mu, sd = tb.mean(0), tb.std(0, ddof=1)
    z = np.abs(tile_means(mtb) - mu) / (sd + 0.5)
    assert z.max() < 6.0 and np.median(z) < 2.0, (z.max(), np.median(z))


def test_photon_list_shape_and_energy(rn, oracle):
    """Pass 1 deposits only on diffuse surfaces, with the power the light
    hands each photon (40 * 1.8 / N) scaled by albedos along the way."""
    cs = compat(rn, oracle, "cornell-box.obj", 0.35)
    n = 5000
    pos, pw, dr = oracle.photon_trace(cs, n, oracle.RNG_PHILOX, (3, 4))
    assert 0 < len(pos) <= 6 * n
    e = np.float32(40.0) * np.float32(1.8) / np.float32(n)
    assert pw.max() <= e * (1 + 1e-6)
    assert np.all(pw >= 0) and np.isclose(np.linalg.norm(dr, axis=1), 1.0, atol=1e-3).all()
    first = np.isclose(pw, e).all(axis=1)
    assert first.mean() > 0.3  # the first diffuse hit carries the emitted power unchanged


def test_oracle_path_tracer_with_mt19937_equals_reference_image(rn, oracle):
    """renderPathTraced restated with the reference's per-pixel std::mt19937
    streams and pathTrace's recursive sums equals the compiled reference's
    image pixel for pixel (160x120, 64 spp)."""
    import ctypes as C
    g = np.load(os.path.join(ROOT, "tests", "golden", "modeR_cornell.npz"))
    ref = g["rgba_160x120_64"]
    H, W = ref.shape[:2]
    cs = compat(rn, oracle, "cornell-box.obj", 0.35)
    out = np.zeros(W * H * 4, np.uint8)
    oracle.lib().or_compat_render_mt(C.byref(cs.s), W, H, 4.0, 64, 0, H, out)
    assert np.array_equal(out.reshape(H, W, 4), ref)
