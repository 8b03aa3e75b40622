"""Seeded, deterministic scene generators for the benchmark workloads.

All arrays are float32, triangles as (n, 3, 3). Randomness comes from
numpy's PCG64 (``np.random.default_rng(seed)``), which is platform-stable.

Noise follows the reference's value-noise and Perlin fBm semantics
(src/Noise.cpp:197-244 value noise with the Perlin permutation hash,
src/Noise.cpp:44-61,246-255 gradient fBm), vectorised over points.
Displacement moves shared vertices, so displaced meshes stay watertight
(unlike the per-face src/Transform.cpp:50-65).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

# Ken Perlin's improved-noise permutation (the table src/Noise.cpp:7-30 uses).
_PERM = np.array([
    151, 160, 137, 91, 90, 15, 131, 13, 201, 95, 96, 53, 194, 233, 7, 225, 140, 36, 103, 30, 69, 142, 8, 99, 37, 240,
    21, 10, 23, 190, 6, 148, 247, 120, 234, 75, 0, 26, 197, 62, 94, 252, 219, 203, 117, 35, 11, 32, 57, 177, 33, 88,
    237, 149, 56, 87, 174, 20, 125, 136, 171, 168, 68, 175, 74, 165, 71, 134, 139, 48, 27, 166, 77, 146, 158, 231, 83,
    111, 229, 122, 60, 211, 133, 230, 220, 105, 92, 41, 55, 46, 245, 40, 244, 102, 143, 54, 65, 25, 63, 161, 1, 216,
    80, 73, 209, 76, 132, 187, 208, 89, 18, 169, 200, 196, 135, 130, 116, 188, 159, 86, 164, 100, 109, 198, 173, 186,
    3, 64, 52, 217, 226, 250, 124, 123, 5, 202, 38, 147, 118, 126, 255, 82, 85, 212, 207, 206, 59, 227, 47, 16, 58, 17,
    182, 189, 28, 42, 223, 183, 170, 213, 119, 248, 152, 2, 44, 154, 163, 70, 221, 153, 101, 155, 167, 43, 172, 9, 129,
    22, 39, 253, 19, 98, 108, 110, 79, 113, 224, 232, 178, 185, 112, 104, 218, 246, 97, 228, 251, 34, 242, 193, 238,
    210, 144, 12, 191, 179, 162, 241, 81, 51, 145, 235, 249, 14, 239, 107, 49, 192, 214, 31, 181, 199, 106, 157, 184,
    84, 204, 176, 115, 121, 50, 45, 127, 4, 150, 254, 138, 236, 205, 93, 222, 114, 67, 29, 24, 72, 243, 141, 128, 195,
    78, 66, 215, 61, 156, 180], dtype=np.int64)
_P = np.concatenate([_PERM, _PERM])


def _fade(t):
    return t * t * t * (t * (t * 6.0 - 15.0) + 10.0)


def value_noise(p: np.ndarray) -> np.ndarray:
    """Value noise in [0,1] at points p (m,3) (src/Noise.cpp:197-230)."""
    p = np.asarray(p, dtype=np.float64)
    fl = np.floor(p)
    X, Y, Z = (fl[:, k].astype(np.int64) for k in range(3))
    x, y, z = (p[:, k] - fl[:, k] for k in range(3))
    u, v, w = _fade(x), _fade(y), _fade(z)

    def lat(a, b, c):
        return _P[_P[_P[a & 255] + (b & 255)] + (c & 255)] / 255.0

    x00 = lat(X, Y, Z) + u * (lat(X + 1, Y, Z) - lat(X, Y, Z))
    x10 = lat(X, Y + 1, Z) + u * (lat(X + 1, Y + 1, Z) - lat(X, Y + 1, Z))
    x01 = lat(X, Y, Z + 1) + u * (lat(X + 1, Y, Z + 1) - lat(X, Y, Z + 1))
    x11 = lat(X, Y + 1, Z + 1) + u * (lat(X + 1, Y + 1, Z + 1) - lat(X, Y + 1, Z + 1))
    y0 = x00 + v * (x10 - x00)
    y1 = x01 + v * (x11 - x01)
    return y0 + w * (y1 - y0)


def fbm_value(p: np.ndarray, octaves: int) -> np.ndarray:
    """fBm of value noise, normalised to [0,1] (src/Noise.cpp:232-244)."""
    s = np.zeros(len(p))
    amp, freq, norm = 1.0, 1.0, 0.0
    for _ in range(octaves):
        s += amp * value_noise(p * freq)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return s / norm


def _grad(h, x, y, z):
    h = h & 15
    u = np.where(h < 8, x, y)
    v = np.where(h < 4, y, np.where((h == 12) | (h == 14), x, z))
    return np.where((h & 1) == 0, u, -u) + np.where((h & 2) == 0, v, -v)


def perlin(p: np.ndarray) -> np.ndarray:
    """Improved Perlin noise (src/Noise.cpp:44-61), vectorised."""
    p = np.asarray(p, dtype=np.float64)
    fl = np.floor(p)
    X, Y, Z = ((fl[:, k].astype(np.int64)) & 255 for k in range(3))
    x, y, z = (p[:, k] - fl[:, k] for k in range(3))
    u, v, w = _fade(x), _fade(y), _fade(z)
    A = _P[X] + Y
    AA = _P[A] + Z
    AB = _P[A + 1] + Z
    B = _P[X + 1] + Y
    BA = _P[B] + Z
    BB = _P[B + 1] + Z

    def lerp(t, a, b):
        return a + t * (b - a)

    return lerp(w,
                lerp(v, lerp(u, _grad(_P[AA], x, y, z), _grad(_P[BA], x - 1, y, z)),
                     lerp(u, _grad(_P[AB], x, y - 1, z), _grad(_P[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, _grad(_P[AA + 1], x, y, z - 1), _grad(_P[BA + 1], x - 1, y, z - 1)),
                     lerp(u, _grad(_P[AB + 1], x, y - 1, z - 1), _grad(_P[BB + 1], x - 1, y - 1, z - 1))))


def fbm_perlin(p: np.ndarray, octaves: int) -> np.ndarray:
    """Gradient fBm mapped to ~[0,1] (src/Noise.cpp:246-255)."""
    s = np.zeros(len(p))
    amp, freq, norm = 1.0, 1.0, 0.0
    for _ in range(octaves):
        s += amp * perlin(p * freq)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return 0.5 + 0.5 * (s / norm)


# ----------------------------------------------------------------- scenes --

@dataclass
class SceneData:
    verts: np.ndarray     # (n,3,3) float32
    albedo: np.ndarray    # (n,3) float32
    emission: np.ndarray  # (n,3) float32
    camera: dict = field(default_factory=dict)
    notes: str = ""

    @property
    def n(self) -> int:
        return int(self.verts.shape[0])


def _quad(a, b, c, d):
    """Two triangles (a,b,c), (a,c,d)."""
    return [[a, b, c], [a, c, d]]


def _box(center, size, angle_deg):
    """Five faces (no bottom) of an axis box rotated about +y, 10 triangles."""
    cx, cy, cz = center
    sx, sy, sz = size
    th = math.radians(angle_deg)
    c, s = math.cos(th), math.sin(th)

    def P(x, y, z):
        x, z = x - cx, z - cz
        return [cx + c * x + s * z, y, cz - s * x + c * z]

    x0, x1 = cx - sx / 2, cx + sx / 2
    z0, z1 = cz - sz / 2, cz + sz / 2
    y0, y1 = cy - sy / 2, cy + sy / 2
    t = []
    t += _quad(P(x0, y1, z0), P(x1, y1, z0), P(x1, y1, z1), P(x0, y1, z1))  # top
    t += _quad(P(x0, y0, z0), P(x1, y0, z0), P(x1, y1, z0), P(x0, y1, z0))  # front
    t += _quad(P(x1, y0, z0), P(x1, y0, z1), P(x1, y1, z1), P(x1, y1, z0))  # right
    t += _quad(P(x1, y0, z1), P(x0, y0, z1), P(x0, y1, z1), P(x1, y1, z1))  # back
    t += _quad(P(x0, y0, z1), P(x0, y0, z0), P(x0, y1, z0), P(x0, y1, z1))  # left
    return t


CORNELL_CAMERA = {"eye": (278.0, 273.0, -800.0), "target": (278.0, 273.0, 0.0), "up": (0.0, 1.0, 0.0),
                  "vfov_deg": 40.0}
WHITE = (0.73, 0.73, 0.73)
RED = (0.65, 0.05, 0.05)
GREEN = (0.12, 0.45, 0.15)
LIGHT_RADIANCE = (17.0, 12.0, 4.0)


def cornell_box(with_boxes: bool = True) -> SceneData:
    """Config-0 Cornell box: 5 walls on a 555-unit cube (2 triangles each),
    boxes 165^3 and 165x330x165 rotated -18/+18 degrees, a 130x105 quad
    light 0.1 below the ceiling with radiance (17,12,4); 32 triangles.
    Camera (278,273,-800) looking +z, vertical fov 40 degrees."""
    L = 555.0
    tris, alb = [], []

    def add(ts, a):
        tris.extend(ts)
        alb.extend([a] * len(ts))

    add(_quad([0, 0, 0], [L, 0, 0], [L, 0, L], [0, 0, L]), WHITE)            # floor
    add(_quad([0, L, 0], [0, L, L], [L, L, L], [L, L, 0]), WHITE)            # ceiling
    add(_quad([0, 0, L], [L, 0, L], [L, L, L], [0, L, L]), WHITE)            # back
    add(_quad([L, 0, 0], [L, L, 0], [L, L, L], [L, 0, L]), RED)              # left (x = 555)
    add(_quad([0, 0, 0], [0, 0, L], [0, L, L], [0, L, 0]), GREEN)            # right (x = 0)
    if with_boxes:
        add(_box((185.5, 82.5, 169.0), (165.0, 165.0, 165.0), -18.0), WHITE)
        add(_box((368.5, 165.0, 351.5), (165.0, 330.0, 165.0), 18.0), WHITE)
    ly = L - 0.1
    lt = _quad([213.0, ly, 227.0], [343.0, ly, 227.0], [343.0, ly, 332.0], [213.0, ly, 332.0])
    n_before = len(tris)
    add(lt, WHITE)
    v = np.asarray(tris, dtype=np.float32)
    a = np.asarray(alb, dtype=np.float32)
    e = np.zeros_like(a)
    e[n_before:] = LIGHT_RADIANCE
    return SceneData(v, a, e, dict(CORNELL_CAMERA), "cornell box, 32 triangles" if with_boxes else "cornell walls")


def _subdivide(verts: np.ndarray, faces: np.ndarray, levels: int):
    """Loop-free 1:4 midpoint subdivision with shared edge vertices."""
    for _ in range(levels):
        e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e[:, 0].astype(np.int64) * (1 << 32) + e[:, 1], return_inverse=True)
        a = (uniq >> 32).astype(np.int64)
        b = (uniq & 0xFFFFFFFF).astype(np.int64)
        mids = 0.5 * (verts[a] + verts[b])
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        base = len(verts)
        verts = np.concatenate([verts, mids])
        m = inv.reshape(3, -1) + base
        m01, m12, m20 = m[0], m[1], m[2]
        f0, f1, f2 = faces[:, 0], faces[:, 1], faces[:, 2]
        faces = np.concatenate([
            np.stack([f0, m01, m20], 1), np.stack([m01, f1, m12], 1), np.stack([m20, m12, f2], 1),
            np.stack([m01, m12, m20], 1)])
    return verts, faces


def geodesic_sphere(levels: int):
    """Tetrahedron-seeded geodesic sphere: 4 * 4^levels triangles
    (levels = 9 gives exactly 1,048,576). Unit radius."""
    v = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 1, 2], [0, 3, 1], [0, 2, 3], [1, 3, 2]], dtype=np.int64)
    return _subdivide(v, f, levels)


def cube_sphere(m: int):
    """Cube sphere with an m x m grid per face: 12 m^2 triangles. Vertices on
    shared cube edges are merged by their integer lattice key, so the mesh is
    watertight. Unit radius."""
    ij = np.arange(m + 1)
    I, J = np.meshgrid(ij, ij, indexing="ij")
    keys, tris = [], []
    axes = [(0, 1, 2, 1), (0, 1, 2, -1), (1, 2, 0, 1), (1, 2, 0, -1), (2, 0, 1, 1), (2, 0, 1, -1)]
    for f, (a, b, c, sgn) in enumerate(axes):
        lat = np.zeros((m + 1, m + 1, 3), np.int64)  # lattice coords in [-m, m]
        lat[..., a] = 2 * I - m
        lat[..., b] = 2 * J - m
        lat[..., c] = sgn * m
        keys.append(lat.reshape(-1, 3))
        base = f * (m + 1) * (m + 1)
        idx = base + (I * (m + 1) + J)
        q00, q10, q01, q11 = idx[:-1, :-1], idx[1:, :-1], idx[:-1, 1:], idx[1:, 1:]
        if sgn > 0:
            tris += [np.stack([q00, q10, q11], -1).reshape(-1, 3), np.stack([q00, q11, q01], -1).reshape(-1, 3)]
        else:
            tris += [np.stack([q00, q11, q10], -1).reshape(-1, 3), np.stack([q00, q01, q11], -1).reshape(-1, 3)]
    allk = np.concatenate(keys)
    uniq, inv = np.unique(allk, axis=0, return_inverse=True)
    v = uniq.astype(np.float64) / m
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    faces = inv.reshape(-1)[np.concatenate(tris)]
    return v, faces


def displace(unit_verts: np.ndarray, radius: float, amplitude: float, seed: int, octaves: int = 4,
             frequency: float = 4.0) -> np.ndarray:
    """Radial displacement r = R (1 + A (2 fbm - 1)) with value-noise fBm
    sampled at frequency * direction + seed offset."""
    off = np.array([17.31, 31.77, 47.13]) * (seed + 1)
    n = fbm_value(unit_verts * frequency + off, octaves)
    r = radius * (1.0 + amplitude * (2.0 * n - 1.0))
    return unit_verts * r[:, None]


SPHERE_CENTER = (278.0, 278.0, 279.5)
SPHERE_RADIUS = 150.0


def _sphere_in_box(unit_v, faces, seed, albedo_lo, albedo_hi, albedo_seed=None):
    rng = np.random.default_rng(seed if albedo_seed is None else albedo_seed)
    pts = displace(unit_v, SPHERE_RADIUS, 0.15, seed) + np.array(SPHERE_CENTER)
    tv = pts[faces].astype(np.float32)
    alb = rng.uniform(albedo_lo, albedo_hi, size=(len(faces), 3)).astype(np.float32)
    box = cornell_box(with_boxes=False)
    v = np.concatenate([box.verts, tv])
    a = np.concatenate([box.albedo, alb])
    e = np.concatenate([box.emission, np.zeros_like(alb)])
    return SceneData(v, a, e, dict(CORNELL_CAMERA))


def config1_scene(seed: int = 0) -> SceneData:
    """1,048,576-triangle displaced geodesic sphere (radius 150, value-noise
    amplitude 0.15 r, 4 octaves) in the Cornell walls + light (12 triangles;
    the two inner boxes are left out because the sphere would intersect
    them). Per-triangle albedo U(0.2, 0.8)."""
    uv, f = geodesic_sphere(9)
    s = _sphere_in_box(uv, f, seed, 0.2, 0.8)
    s.notes = f"displaced geodesic sphere 1,048,576 tris + 12 Cornell tris, seed {seed}"
    return s


def sphere_scene_tris(target_tris: int, seed: int, albedo_lo=0.2, albedo_hi=0.8, albedo_seed=None) -> SceneData:
    """Displaced cube sphere with m = round(sqrt(target/12)) in the Cornell
    walls + light."""
    m = max(1, int(round(math.sqrt(target_tris / 12.0))))
    uv, f = cube_sphere(m)
    s = _sphere_in_box(uv, f, seed, albedo_lo, albedo_hi, albedo_seed)
    s.notes = f"displaced cube sphere {len(f)} tris (m={m}) + 12 Cornell tris, seed {seed}"
    return s


def config3_scenes():
    """Differentiable render: 100k-class sphere (cube sphere m=91: 99,372
    tris) + Cornell walls + light. Target albedos from seed 1, start from
    seed-2 albedos and emission * 0.5. Returns (target, start)."""
    tgt = sphere_scene_tris(100_000, seed=3, albedo_seed=1)
    start = sphere_scene_tris(100_000, seed=3, albedo_seed=2)
    start.emission = (start.emission * 0.5).astype(np.float32)
    # The Cornell walls keep their fixed albedos in both.
    return tgt, start


@dataclass
class BatchSpec:
    scene: SceneData
    width: int
    height: int
    spp: int
    max_depth: int


def config2_batch(count: int = 64, seed: int = 2, scale_spp: float = 1.0) -> List[BatchSpec]:
    """Batched workload: `count` seeded displaced spheres in the Cornell box,
    triangles log-uniform in [1k, 500k], resolution uniform over
    {256, 512, 768, 1024}^2, spp uniform over {4, 16, 64}, depth 8."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        tris = int(round(math.exp(rng.uniform(math.log(1000), math.log(500_000)))))
        res = int(rng.choice([256, 512, 768, 1024]))
        spp = int(rng.choice([4, 16, 64]))
        sc = sphere_scene_tris(tris, seed=1000 + i)
        out.append(BatchSpec(sc, res, res, max(1, int(round(spp * scale_spp))), 8))
    return out


def terrain_scene(n_side: int = 2236, seed: int = 4, n_lights: int = 8) -> SceneData:
    """Fractal heightfield: an n_side x n_side vertex grid ((n_side-1)^2 * 2
    triangles; 2236 gives 9,990,450), 6-octave Perlin fBm heights, albedo
    U(0.3, 0.7), `n_lights` random quad lights with radiance U(5, 20)."""
    rng = np.random.default_rng(seed)
    ext = 1000.0
    xs = np.linspace(0.0, ext, n_side, dtype=np.float64)
    X, Z = np.meshgrid(xs, xs, indexing="xy")
    pts = np.stack([X.ravel() / ext * 6.0 + 0.37, Z.ravel() / ext * 6.0 + 0.73, np.full(X.size, 0.5 + seed)], 1)
    H = (fbm_perlin(pts, 6) - 0.5) * 300.0
    V = np.stack([X.ravel(), H, Z.ravel()], 1).astype(np.float32)
    idx = np.arange(n_side * n_side, dtype=np.int64).reshape(n_side, n_side)
    a, b, c, d = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel(), idx[1:, :-1].ravel(), idx[1:, 1:].ravel()
    faces = np.concatenate([np.stack([a, c, b], 1), np.stack([b, c, d], 1)])
    tv = V[faces]
    alb = rng.uniform(0.3, 0.7, size=(len(faces), 3)).astype(np.float32)
    em = np.zeros_like(alb)
    lights, lalb, lem = [], [], []
    for _ in range(n_lights):
        cx, cz = rng.uniform(100, 900, 2)
        y = 450.0 + rng.uniform(0, 100)
        s = rng.uniform(30, 80)
        q = _quad([cx - s, y, cz - s], [cx + s, y, cz - s], [cx + s, y, cz + s], [cx - s, y, cz + s])
        L = rng.uniform(5, 20, 3)
        lights += q
        lalb += [WHITE] * 2
        lem += [L] * 2
    v = np.concatenate([tv, np.asarray(lights, np.float32)])
    al = np.concatenate([alb, np.asarray(lalb, np.float32)])
    e = np.concatenate([em, np.asarray(lem, np.float32)])
    cam = {"eye": (500.0, 420.0, -350.0), "target": (500.0, 0.0, 500.0), "up": (0.0, 1.0, 0.0), "vfov_deg": 50.0}
    return SceneData(v.astype(np.float32), al.astype(np.float32), e.astype(np.float32), cam,
                     f"terrain {len(faces)} tris + {2 * n_lights} light tris, seed {seed}")
