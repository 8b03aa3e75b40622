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
    "config1": (lambda: scenes.config1_scene(), 1920, 1080, 8, 3, 0.03, 32),
}


def main(names):
    for name in names:
        build, W, H, D, rr, eps, block = WORKLOADS[name]
        sd = build()
        cam = O.camera_basis(sd.camera["eye"], sd.camera["target"], sd.camera["up"], sd.camera["vfov_deg"], W, H)
        pix = pixel_subset(W, H, block, seed=17)
        ps = O.PhysScene(sd.verts, sd.albedo, sd.emission, "f64")
        p = O.params(W, H, SPP, D, rr, 0.0, ORACLE_SEED, 0, eps)
        t0 = time.time()
        out, cnt = ps.render(cam, p, pixels=pix)
        dt = time.time() - t0
        np.savez_compressed(os.path.join(HERE, f"oracle64k_{name}.npz"), pixels=pix, values=out, spp=SPP,
                            seed=ORACLE_SEED, width=W, height=H, max_depth=D, rr_depth=rr, ray_eps=eps, block=block,
                            seconds=dt, threads=os.cpu_count())
        print(f"{name}: {len(pix)} pixels x {SPP} spp in {dt:.0f}s ({cnt.closest + cnt.shadow} rays)")


if __name__ == "__main__":
    main(sys.argv[1:] or list(WORKLOADS))
