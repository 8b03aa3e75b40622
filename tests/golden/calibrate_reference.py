"""Time the compiled reference against the oracle's CPU port, in this build
container, and write tests/golden/reference_timing.json.

The compiled reference (oracle/_ref/librednoise_ref.so, built from the
reference's own sources by `make -C oracle ref`) never travels to the GPU
machine. What the benchmark reports there as "reference CPU time" is the CPU
port's time on that machine divided by the port/reference time ratio measured
here on the same call; the reference images' MD5 digests let the benchmark
state whether the GPU image of a deterministic call is the reference's.

Calls (the reference CLI's defaults where it has them):
  rn_render(RN_PATHTRACED)  cornell-box.obj (scale 0.35), 320x240, 16 spp,
                            cam_z 4 -- port: or_compat_render_mt (the same
                            per-pixel std::mt19937 streams; OpenMP rows)
  rn_render(RN_RAYTRACED)   cornell-box.obj, 1280x960, cam_z 4 -- no port:
                            the reference time here is reported as measured
  renderPhotonMapped        cornell-box.obj, 640x480, 200,000 photons
  renderBidirectional       cornell-box.obj, 640x480, 16 samples, exposure 40

usage: make -C oracle ref && python tests/golden/calibrate_reference.py
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import platform
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

ASSET = os.path.join(HERE, "assets", "cornell-box.obj")
OUT = os.path.join(HERE, "reference_timing.json")


def best_of(fn, reps=3):
    fn()  # warm-up (page-in, OpenMP pool)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return min(ts)


def md5(buf: np.ndarray) -> str:
    return hashlib.md5(np.ascontiguousarray(buf, np.uint8).tobytes()).hexdigest()


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            return next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), platform.processor())
    except OSError:
        return platform.processor()


def main() -> None:
    ref = O.ref()
    if ref is None:
        raise SystemExit("build the reference first: make -C oracle ref")
    import rednoise_amd as rn  # host-side scene loading only (no GPU call)

    threads = os.cpu_count() or 1
    out = {"container_cpu": cpu_model(), "container_threads": threads,
           "note": "reference = oracle/_ref (the reference's sources compiled with g++ -O2 -fopenmp); port = the "
                   "oracle's C restatement; both timed here, best of 3 after a warm-up"}

    # rn_render(RN_PATHTRACED): the reference estimator with its own streams
    W, H, spp, cz = 320, 240, 16, 4.0
    sc = ref.rn_scene_load_obj(ASSET.encode(), C.c_float(0.35))
    buf = np.empty(W * H * 4, np.uint8)
    t_ref = best_of(lambda: ref.rn_render(sc, 3, W, H, C.c_float(cz), spp, buf))
    ref.rn_scene_free(sc)
    ours = rn.Scene.load_obj(ASSET, 0.35)
    cs = O.CompatArrays.from_scene(ours)
    port = np.zeros(W * H * 4, np.uint8)
    t_port = best_of(lambda: O.lib().or_compat_render_mt(C.byref(cs.s), W, H, cz, spp, 0, H, port))
    assert np.array_equal(port, buf), "the port's image differs from the reference's"
    out["rn_render_pathtraced"] = {
        "call": f"rn_render(RN_PATHTRACED) cornell-box.obj scale 0.35 {W}x{H} spp={spp} cam_z={cz:g}",
        "width": W, "height": H, "spp": spp, "cam_z": cz,
        "reference_s": round(t_ref, 5), "port_s": round(t_port, 5), "port_over_reference": round(t_port / t_ref, 4),
        "reference_rgba_md5": md5(buf)}

    # rn_render(RN_RAYTRACED): deterministic; no CPU port
    W2, H2 = 1280, 960
    sc = ref.rn_scene_load_obj(ASSET.encode(), C.c_float(0.35))
    buf2 = np.empty(W2 * H2 * 4, np.uint8)
    t_ref2 = best_of(lambda: ref.rn_render(sc, 2, W2, H2, C.c_float(cz), 1, buf2))
    ref.rn_scene_free(sc)
    out["rn_render_raytraced"] = {"call": f"rn_render(RN_RAYTRACED) cornell-box.obj scale 0.35 {W2}x{H2} cam_z={cz:g}",
                                  "width": W2, "height": H2, "cam_z": cz, "reference_s": round(t_ref2, 5),
                                  "reference_rgba_md5": md5(buf2)}

    # photon mapping: pass 1 + pass 2 at the CLI defaults
    W3, H3, nph = 640, 480, 200000
    buf3 = np.empty(W3 * H3 * 4, np.uint8)
    t_ref3 = best_of(lambda: ref.ref_render_photon(ASSET.encode(), 0.35, W3, H3, cz, nph, 0, buf3), reps=2)
    out["photon"] = {"call": f"renderPhotonMapped cornell-box.obj scale 0.35 {W3}x{H3} {nph} photons cam_z={cz:g}",
                     "width": W3, "height": H3, "photons": nph, "cam_z": cz, "reference_s": round(t_ref3, 4),
                     "reference_rgba_md5": md5(buf3)}

    # bidirectional path tracing
    bspp = 16
    buf4 = np.empty(W3 * H3 * 4, np.uint8)
    t_ref4 = best_of(lambda: ref.ref_render_bidirectional(ASSET.encode(), 0.35, W3, H3, cz, bspp, 40.0, buf4), reps=2)
    out["bidirectional"] = {"call": f"renderBidirectional cornell-box.obj scale 0.35 {W3}x{H3} {bspp} samples "
                                    f"exposure 40 cam_z={cz:g}",
                            "reference_s": round(t_ref4, 4)}
    ours.close()
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
