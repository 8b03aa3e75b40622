"""FP64 CPU-oracle references at 65,536 spp for the equal-RMSE protocol.

For each workload, one seeded pixel per B x B block of the image is rendered
by the FP64 oracle with an independent Philox seed; the GPU renders are
scored against these values (relative RMSE over the subset). The subset and
seed are stored with the values.

    python tests/golden/make_oracle_rmse.py config0 config1
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from rednoise_amd import scenes  # noqa: E402

ORACLE_SEED = 0x0DDBA11
SPP = 65536


def pixel_subset(W, H, block, seed):
    rng = np.random.default_rng(seed)
    pix = []
    for by in range(0, H, block):
        for bx in range(0, W, block):
            x = bx + int(rng.integers(0, min(block, W - bx)))
            y = by + int(rng.integers(0, min(block, H - by)))
            pix.append(y * W + x)
    return np.array(pix, np.int32)


WORKLOADS = {
    # name: (scene builder, W, H, max_depth, rr_depth, ray_eps, block)
    "config0": (lambda: scenes.cornell_box(), 320, 240, 5, 3, 0.05, 8),
    "config1": (lambda: scenes.config1_scene(), 1920, 1080, 8, 3, 0.03, 16),
    # the adjoint's forward render: the start scene (seed-2 albedos, emission
    # x 0.5), no roulette
    "config3": (lambda: scenes.config3_scenes()[1], 512, 512, 6, 1 << 30, 0.03, 16),
    # 10M triangles: one pixel per 32 x 32 block (8,160 pixels) keeps the CPU
    # reference to minutes
    "config4": (lambda: scenes.terrain_scene(), 3840, 2160, 12, 3, 0.05, 32),
}
# config 2: every 8th of the 64 batch scenes (scenes.config2_batch(64, seed=2)),
# one pixel per 64 x 64 block of each
CONFIG2_SCENES = list(range(0, 64, 8))
CONFIG2_BLOCK = 64


def render_subset(sd, W, H, D, rr, eps, block, seed=17):
    cam = O.camera_basis(sd.camera["eye"], sd.camera["target"], sd.camera["up"], sd.camera["vfov_deg"], W, H)
    pix = pixel_subset(W, H, block, seed=seed)
    ps = O.PhysScene(sd.verts, sd.albedo, sd.emission, "f64")
    p = O.params(W, H, SPP, D, rr, 0.0, ORACLE_SEED, 0, eps)
    out, cnt = ps.render(cam, p, pixels=pix)
    return pix, out, cnt


def main(names):
    for name in names:
        t0 = time.time()
        if name == "config2":
            specs = scenes.config2_batch(64, seed=2)
            idx, pixs, vals, widths = [], [], [], []
            for k in CONFIG2_SCENES:
                s = specs[k]
                pix, out, _ = render_subset(s.scene, s.width, s.height, s.max_depth, 3, 0.03, CONFIG2_BLOCK)
                idx.append(np.full(len(pix), k, np.int32))
                pixs.append(pix)
                vals.append(out)
                widths.append(s.width)
                print(f"  config2 scene {k}: {s.width}x{s.height}, {len(pix)} pixels, {time.time() - t0:.0f}s",
                      flush=True)
            dt = time.time() - t0
            np.savez_compressed(os.path.join(HERE, "oracle64k_config2.npz"), scene=np.concatenate(idx),
                                pixels=np.concatenate(pixs), values=np.concatenate(vals), spp=SPP, seed=ORACLE_SEED,
                                scenes=np.array(CONFIG2_SCENES, np.int32), block=CONFIG2_BLOCK, max_depth=8,
                                rr_depth=3, ray_eps=0.03, seconds=dt, threads=os.cpu_count())
            print(f"config2: {sum(len(x) for x in pixs)} pixels x {SPP} spp in {dt:.0f}s", flush=True)
            continue
        build, W, H, D, rr, eps, block = WORKLOADS[name]
        sd = build()
        pix, out, cnt = render_subset(sd, W, H, D, rr, eps, block)
        dt = time.time() - t0
        np.savez_compressed(os.path.join(HERE, f"oracle64k_{name}.npz"), pixels=pix, values=out, spp=SPP,
                            seed=ORACLE_SEED, width=W, height=H, max_depth=D, rr_depth=rr, ray_eps=eps, block=block,
                            seconds=dt, threads=os.cpu_count())
        print(f"{name}: {len(pix)} pixels x {SPP} spp in {dt:.0f}s ({cnt.closest + cnt.shadow} rays)", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or list(WORKLOADS) + ["config2"])
