"""Autotuner for batched wavefront rendering.

Picks the wavefront queue size (paths in flight), the BVH width (4 or 8
children) and the block size of the trace kernels for a batch of scenes by
timing each candidate on a reduced-spp copy of the batch. Results are cached
as JSON keyed by the GPU architecture and a signature of the workload
(scene count, total triangles, pixels and samples, each rounded to a power of
two), in ``$RN_AUTOTUNE_CACHE`` or ``~/.cache/rednoise_amd/autotune.json``.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time
from typing import List, Optional, Sequence, Tuple

from . import BatchOptions, Camera, RenderParams, Scene, device_name, last_stats, render_batch

QUEUE_SIZES = (1 << 23, 1 << 25, 1 << 27)
BVH_WIDTHS = (8, 4)
BLOCKS = (256, 128)


def _cache_path() -> str:
    p = os.environ.get("RN_AUTOTUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "rednoise_amd", "autotune.json")


def _bucket(x: float) -> int:
    return int(round(math.log2(max(1.0, float(x)))))


def signature(items: Sequence[Tuple[Scene, Camera, RenderParams]]) -> str:
    tris = sum(sc.triangle_count for sc, _, _ in items)
    pix = sum(p.width * p.height for _, _, p in items)
    paths = sum(p.width * p.height * p.spp for _, _, p in items)
    depth = max(p.max_depth for _, _, p in items)
    key = f"{device_name()}|n={len(items)}|t={_bucket(tris)}|p={_bucket(pix)}|s={_bucket(paths)}|d={depth}"
    return key


def _reduced(items, frac: float):
    out = []
    for sc, cam, p in items:
        q = RenderParams.from_buffer_copy(p)
        q.spp = max(1, int(round(p.spp * frac)))
        out.append((sc, cam, q))
    return out


def autotune_batch(items: Sequence[Tuple[Scene, Camera, RenderParams]], frac: float = 0.0625,
                   use_cache: bool = True, candidates: Optional[List[dict]] = None) -> Tuple[BatchOptions, dict]:
    """Return (best BatchOptions, report). The report lists every candidate's
    Mrays/s on the reduced batch."""
    key = signature(items)
    path = _cache_path()
    cache = {}
    if use_cache and os.path.isfile(path):
        try:
            cache = json.load(open(path))
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache:
        c = cache[key]["best"]
        return BatchOptions(c["queue_paths"], c["block_size"], c["bvh_width"], 1, c["iters_per_graph"]), {
            "key": key, "cached": True, "best": c}
    total_paths = sum(p.width * p.height * p.spp for _, _, p in items)
    if candidates is None:
        candidates = [dict(queue_paths=q, bvh_width=w, block_size=b, iters_per_graph=8)
                      for w in BVH_WIDTHS for q in QUEUE_SIZES for b in BLOCKS if q <= 2 * total_paths]
    trial = _reduced(items, frac)
    results = []
    for c in candidates:
        opts = BatchOptions(c["queue_paths"], c["block_size"], c["bvh_width"], 1, c["iters_per_graph"])
        render_batch(trial, opts)  # warm-up (and BVH build for this width)
        t0 = time.perf_counter()
        render_batch(trial, opts)
        dt = time.perf_counter() - t0
        st = last_stats()
        rays = st["rays_closest"] + st["rays_shadow"]
        results.append(dict(c, mrays_s=round(rays / dt / 1e6, 1), seconds=round(dt, 4)))
    best = max(results, key=lambda r: r["mrays_s"])
    report = {"key": key, "cached": False, "trial_spp_fraction": frac, "candidates": results, "best": best}
    if use_cache:
        cache[key] = report
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1)
        os.replace(tmp, path)
    return BatchOptions(best["queue_paths"], best["block_size"], best["bvh_width"], 1, best["iters_per_graph"]), report
