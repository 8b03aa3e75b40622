"""rednoise_amd: MI355X-native path tracer behind the rednoise C ABI.

The public surface mirrors the reference Python binding
(bindings/python/rednoise/__init__.py: ``RenderMode``, ``Scene.load_obj``,
``Scene.render``, ``Scene.triangle_count``, ``Scene.close``, ``Image``,
``version``) over ``librednoise.so``, which this package builds in place
(``rednoise_amd/librednoise.so``). Rendering runs on the GPU only: there is no
CPU fallback, and every call raises when the library or the GPU is missing.

Additions beyond the reference binding (the extended ABI, rednoise_ex.h):
``Scene.from_arrays``, ``Scene.render_hdr`` (physical estimator),
``Scene.adjoint``, ``Scene.trace_rays``, ``Scene.trace_primary_compat``,
``Scene.build``, ``render_batch`` and ``last_stats``.
"""
from __future__ import annotations

import ctypes as C
import os
from enum import IntEnum
from typing import Optional, Sequence, Tuple, Union

import numpy as np

__all__ = [
    "RenderMode", "Scene", "Image", "version", "Camera", "RenderParams", "BatchOptions", "render_batch",
    "last_stats", "library_path", "device_name", "INTEGRATOR_MEGAKERNEL", "INTEGRATOR_WAVEFRONT",
]

__version__ = "0.1.2"

INTEGRATOR_MEGAKERNEL = 0
INTEGRATOR_WAVEFRONT = 1
FLAG_COUNT_TRAVERSAL = 1
FLAG_TIME_STAGES = 2
FLAG_USE_GRAPH = 4
FLAG_FB_ATOMICS = 8
QUERY_CLOSEST = 0
QUERY_ANY = 1
TRI_WATERTIGHT = 0
TRI_COMPAT = 1

_HERE = os.path.dirname(os.path.abspath(__file__))


class RenderMode(IntEnum):
    """Which renderer to run (the C ``rn_render_mode`` enum)."""

    WIREFRAME = 0
    RASTERISED = 1
    RAYTRACED = 2
    PATHTRACED = 3

    @classmethod
    def coerce(cls, value: Union["RenderMode", int, str]) -> "RenderMode":
        """The member named by a member, its integer value or its name.

        Names are looked up in a table of folded spellings: case, surrounding
        blanks and the separators '-', '_' and ' ' do not matter, so
        "path_traced" and "Path-Traced" both give PATHTRACED. Unknown names
        and values raise ValueError; a bool (an int subclass) and any other
        type raise TypeError. The accepted inputs and exception types are
        those of bindings/python/rednoise/__init__.py:51-73."""
        if isinstance(value, cls):
            return value
        if isinstance(value, str):
            member = _MODES_BY_FOLDED_NAME.get(_fold_mode_name(value))
            if member is None:
                raise ValueError(f"no render mode is called {value!r} (known: {', '.join(_MODES_BY_FOLDED_NAME)})")
            return member
        if isinstance(value, int) and not isinstance(value, bool):
            return cls(value)
        raise TypeError(f"a render mode is a RenderMode, an int or a name; got {type(value).__name__}")


def _fold_mode_name(name: str) -> str:
    return "".join(ch for ch in name.strip().lower() if ch not in "-_ ")


_MODES_BY_FOLDED_NAME = {_fold_mode_name(m.name): m for m in RenderMode}


# ----------------------------------------------------------------- structs --

class Camera(C.Structure):
    """rn_camera: eye, look-at target, up, vertical field of view (degrees)."""

    _fields_ = [("eye", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3), ("vfov_deg", C.c_float)]

    @classmethod
    def make(cls, eye, target, up=(0.0, 1.0, 0.0), vfov_deg=40.0) -> "Camera":
        return cls((C.c_float * 3)(*eye), (C.c_float * 3)(*target), (C.c_float * 3)(*up), float(vfov_deg))


class RenderParams(C.Structure):
    """rn_render_params."""

    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("spp", C.c_int), ("max_depth", C.c_int), ("rr_depth", C.c_int),
        ("rr_prob", C.c_float), ("seed", C.c_uint32), ("ray_eps", C.c_float), ("integrator", C.c_int),
        ("queue_paths", C.c_int), ("block_size", C.c_int), ("bvh_width", C.c_int), ("flags", C.c_int),
    ]

    @classmethod
    def make(cls, width, height, spp, max_depth, rr_depth=1 << 30, rr_prob=0.0, seed=0, ray_eps=0.0,
             integrator=INTEGRATOR_MEGAKERNEL, queue_paths=0, block_size=0, bvh_width=8, flags=0) -> "RenderParams":
        return cls(int(width), int(height), int(spp), int(max_depth), int(min(rr_depth, 1 << 30)), float(rr_prob),
                   int(seed) & 0xFFFFFFFF, float(ray_eps), int(integrator), int(queue_paths), int(block_size),
                   int(bvh_width), int(flags))


class BatchItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", Camera), ("params", RenderParams), ("hdr_rgb", C.c_void_p)]


class BatchOptions(C.Structure):
    """rn_batch_options."""

    _fields_ = [("queue_paths", C.c_int), ("block_size", C.c_int), ("bvh_width", C.c_int), ("use_graph", C.c_int),
                ("iters_per_graph", C.c_int)]


class Stats(C.Structure):
    """rn_stats of the last call on this thread."""

    _fields_ = [
        ("build_ms", C.c_double), ("render_ms", C.c_double), ("total_ms", C.c_double),
        ("rays_closest", C.c_uint64), ("rays_shadow", C.c_uint64), ("paths", C.c_uint64),
        ("trace_kernel_ms", C.c_double), ("trace_kernel_launches", C.c_uint64), ("bvh_nodes", C.c_int),
        ("bvh_depth", C.c_int), ("nodes_closest", C.c_uint64), ("tris_closest", C.c_uint64),
        ("nodes_shadow", C.c_uint64), ("tris_shadow", C.c_uint64), ("extend_ms", C.c_double),
        ("shade_ms", C.c_double), ("shadow_ms", C.c_double), ("iterations", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class BvhInfo(C.Structure):
    _fields_ = [("nodes", C.c_int), ("depth", C.c_int), ("leaf_tris", C.c_int), ("width", C.c_int),
                ("build_ms", C.c_double), ("builder", C.c_int)]


# ------------------------------------------------------------ the library --

_LIB = None


def library_path() -> str:
    """Path of the in-tree librednoise.so (REDNOISE_LIBRARY overrides)."""
    env = os.environ.get("REDNOISE_LIBRARY")
    if env:
        if not os.path.isfile(env):
            raise OSError(f"REDNOISE_LIBRARY is set to {env!r} but that file does not exist.")
        return env
    return os.path.join(_HERE, "librednoise.so")


def _load_library() -> C.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    # PyTorch wheels ship their own libamdhip64.so under the same soname as
    # the system one; whichever is loaded first serves the whole process.
    # Loading torch first keeps a single HIP runtime when both are used.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    path = library_path()
    if not os.path.isfile(path):
        raise OSError(f"librednoise.so not found at {path!r}; build it with `python -c \"import __graft_entry__ as g; "
                      "g.build()\"` or `make -C rednoise_amd/csrc`.")
    lib = C.CDLL(path)
    _configure(lib)
    _LIB = lib
    return lib


def _configure(lib: C.CDLL) -> None:
    vp, i, f, cp = C.c_void_p, C.c_int, C.c_float, C.c_char_p
    sig = {
        "rn_scene_load_obj": ([cp, f], vp), "rn_scene_free": ([vp], None), "rn_scene_triangle_count": ([vp], i),
        "rn_render": ([vp, i, i, i, f, i, vp], i), "rn_render_seeded": ([vp, i, i, f, i, C.c_uint32, C.c_uint32, vp], i), "rn_save_png": ([cp, i, i, vp], i), "rn_version": ([], cp),
        "rn_last_error": ([], cp), "rn_device_name": ([], cp), "rn_release_workspace": ([], i),
        "rn_scene_save": ([vp, cp], i), "rn_scene_load": ([cp], vp),
        "rn_scene_get_vertex_normals": ([vp, vp], i),
        "rn_accum_create": ([i, i], vp), "rn_accum_free": ([vp], None), "rn_accum_reset": ([vp], i),
        "rn_accum_render": ([vp, vp, C.POINTER(Camera), C.POINTER(RenderParams)], i),
        "rn_accum_samples": ([vp], i), "rn_accum_resolve_rgba8": ([vp, f, f, vp], i),
        "rn_accum_mean": ([vp, vp], i),
        "rn_scene_from_arrays": ([vp, i, vp, vp], vp), "rn_scene_set_materials": ([vp, vp, vp], i),
        "rn_scene_get_triangles": ([vp, vp, vp, vp], i), "rn_scene_get_compat": ([vp, vp, vp, vp, vp, vp], i),
        "rn_scene_get_texture": ([vp, i, vp, vp, vp], i),
        "rn_render_ex": ([vp, C.POINTER(Camera), C.POINTER(RenderParams), vp], i),
        "rn_render_ex_device": ([vp, C.POINTER(Camera), C.POINTER(RenderParams), vp, vp, vp, vp], i),
        "rn_render_ex_device_noalloc": ([vp, C.POINTER(Camera), C.POINTER(RenderParams), vp, vp, vp, vp], i),
        "rn_adjoint": ([vp, C.POINTER(Camera), C.POINTER(RenderParams), vp, vp, vp], i),
        "rn_adjoint_device": ([vp, C.POINTER(Camera), C.POINTER(RenderParams), vp, vp, vp, vp, vp, vp], i),
        "rn_render_batch": ([C.POINTER(BatchItem), i, C.POINTER(BatchOptions)], i),
        "rn_get_last_stats": ([C.POINTER(Stats)], i),
        "rn_trace_rays": ([vp, i, vp, vp, vp, vp, i, i, vp, vp, vp, vp], i),
        "rn_trace_primary_compat": ([vp, i, i, f, vp, vp, vp, vp, vp], i),
        "rn_scene_build": ([vp, i, C.POINTER(BvhInfo)], i), "rn_scene_export_bvh": ([vp, vp, vp], i),
        "rn_scene_set_builder": ([vp, i], i),
        "rn_lbvh_stages": ([vp, i, vp, vp, vp, vp], i), "rn_ploc_stages": ([vp, i, vp, vp, vp], i),
        "rn_radix_sort_pairs": ([vp, vp, i, i], i),
        "rn_philox4x32_10": ([i, vp, vp, vp], i),
        "rn_photon_map_trace": ([vp, i, C.c_uint32, C.c_uint32], vp),
        "rn_photon_map_create": ([i, vp, vp, vp, f], vp), "rn_photon_map_free": ([vp], None),
        "rn_photon_map_count": ([vp], i), "rn_photon_map_get": ([vp, vp, vp, vp], i),
        "rn_photon_map_gather": ([vp, i, vp, vp, f, vp], i),
        "rn_render_photon": ([vp, vp, i, i, f, i, C.c_uint32, C.c_uint32, vp], i),
        "rn_render_bidirectional": ([vp, i, i, f, i, f, C.c_uint32, C.c_uint32, vp], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _err(what: str) -> RuntimeError:
    msg = _load_library().rn_last_error().decode(errors="replace")
    return RuntimeError(f"{what} failed: {msg}")


def version() -> str:
    """Library version string (``rn_version``)."""
    return _load_library().rn_version().decode()


def device_name() -> str:
    return _load_library().rn_device_name().decode()


class Accumulator:
    """Progressive renderer state (``rn_accum_*``): a float64 running sum on
    the GPU. ``render`` adds ``params.spp`` samples per pixel continuing the
    sample numbering, ``resolve`` tone-maps and packs RGBA8 in one GPU pass,
    ``mean`` returns the HDR running mean."""

    def __init__(self, width: int, height: int):
        lib = _load_library()
        self._lib = lib
        self._h = lib.rn_accum_create(int(width), int(height))
        if not self._h:
            raise _err("rn_accum_create")
        self.width, self.height = int(width), int(height)

    @property
    def samples(self) -> int:
        return int(self._lib.rn_accum_samples(self._h))

    def reset(self) -> None:
        if self._lib.rn_accum_reset(self._h) != 1:
            raise _err("rn_accum_reset")

    def render(self, scene: "Scene", camera: "Camera", params: "RenderParams") -> None:
        if self._lib.rn_accum_render(self._h, scene.handle, C.byref(camera), C.byref(params)) != 1:
            raise _err("rn_accum_render")

    def resolve(self, exposure: float = 1.0, gamma: float = 1.0) -> np.ndarray:
        out = np.empty(self.width * self.height * 4, np.uint8)
        if self._lib.rn_accum_resolve_rgba8(self._h, C.c_float(exposure), C.c_float(gamma), _ptr(out)) != 1:
            raise _err("rn_accum_resolve_rgba8")
        return out.reshape(self.height, self.width, 4)

    def mean(self) -> np.ndarray:
        out = np.empty(self.width * self.height * 3, np.float32)
        if self._lib.rn_accum_mean(self._h, _ptr(out)) != 1:
            raise _err("rn_accum_mean")
        return out.reshape(self.height, self.width, 3)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.rn_accum_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PhotonMap:
    """A photon map on the GPU (``rn_photon_map_*``), the PhotonMap of the
    reference's photon mapper: ``PhotonMap.trace(scene, n)`` runs pass 1 of
    renderPhotonMapped, ``PhotonMap.from_photons`` wraps given photons;
    ``photons()`` returns them in emission order and ``gather`` evaluates
    PhotonMap::gather."""

    def __init__(self, handle):
        self._lib = _load_library()
        self._h = handle

    @classmethod
    def trace(cls, scene: "Scene", num_photons: int, key0: int = 0, key1: int = 0) -> "PhotonMap":
        h = _load_library().rn_photon_map_trace(scene.handle, int(num_photons), int(key0) & 0xFFFFFFFF,
                                                 int(key1) & 0xFFFFFFFF)
        if not h:
            raise _err("rn_photon_map_trace")
        return cls(h)

    @classmethod
    def from_photons(cls, position, power, direction, cell: float = 0.25) -> "PhotonMap":
        pos = _f32(position)
        n = pos.size // 3
        pw, dr = _f32(power, 3 * n), _f32(direction, 3 * n)
        h = _load_library().rn_photon_map_create(n, _ptr(pos), _ptr(pw), _ptr(dr), C.c_float(cell))
        if not h:
            raise _err("rn_photon_map_create")
        return cls(h)

    @property
    def handle(self):
        return self._h

    def __len__(self) -> int:
        return int(self._lib.rn_photon_map_count(self._h))

    def photons(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        n = len(self)
        pos, pw, dr = (np.empty(3 * n, np.float32) for _ in range(3))
        if self._lib.rn_photon_map_get(self._h, _ptr(pos), _ptr(pw), _ptr(dr)) != 1:
            raise _err("rn_photon_map_get")
        return pos.reshape(-1, 3), pw.reshape(-1, 3), dr.reshape(-1, 3)

    def gather(self, points, normals, radius: float = 0.25) -> np.ndarray:
        p = _f32(points)
        m = p.size // 3
        q = _f32(normals, 3 * m)
        out = np.empty(3 * m, np.float32)
        if self._lib.rn_photon_map_gather(self._h, m, _ptr(p), _ptr(q), C.c_float(radius), _ptr(out)) != 1:
            raise _err("rn_photon_map_gather")
        return out.reshape(-1, 3)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.rn_photon_map_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def release_workspace() -> None:
    """Free the device work buffers kept between renders (the wavefront keeps
    up to 2^27 paths in flight by default, about 21 GB); the next render
    allocates them again."""
    lib = _load_library()
    if lib.rn_release_workspace() != 1:
        raise RuntimeError(lib.rn_last_error().decode())


def last_stats() -> dict:
    s = Stats()
    _load_library().rn_get_last_stats(C.byref(s))
    return s.as_dict()


def _f32(a, n=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} floats, got {a.size}")
    return a


def _out_f32(a, n: int) -> np.ndarray:
    """A caller's output array (C-contiguous float32 of n values, viewed flat),
    or a new one."""
    if a is None:
        return np.empty(n, np.float32)
    if not isinstance(a, np.ndarray) or a.dtype != np.float32 or not a.flags.c_contiguous or a.size != n:
        raise ValueError(f"out must be a C-contiguous float32 array of {n} values")
    return a.reshape(-1)


# ------------------------------------------------------------------ image --

class Image:
    """An RGBA8 image (bindings/python/rednoise/__init__.py:240-325)."""

    __slots__ = ("_rgba", "_width", "_height")

    def __init__(self, rgba: bytes, width: int, height: int):
        expected = width * height * 4
        if len(rgba) != expected:
            raise ValueError(f"rgba buffer has {len(rgba)} bytes, expected {expected} for a {width}x{height} "
                             "RGBA8 image")
        self._rgba = bytes(rgba)
        self._width = int(width)
        self._height = int(height)

    @property
    def rgba(self) -> bytes:
        return self._rgba

    @property
    def width(self) -> int:
        return self._width

    @property
    def height(self) -> int:
        return self._height

    @property
    def size(self) -> Tuple[int, int]:
        return (self._width, self._height)

    def to_numpy(self) -> np.ndarray:
        return np.frombuffer(self._rgba, dtype=np.uint8).reshape(self._height, self._width, 4)

    def save(self, path: Union[str, os.PathLike]) -> None:
        """Write the image. A name ending in .png (or with no extension) is
        written by the library's own PNG encoder (``rn_save_png``); any other
        extension is handed to Pillow when it is installed, as the reference
        binding does, and raises ValueError when it is not."""
        path = os.fspath(path)
        ext = os.path.splitext(path)[1].lower()
        if ext and ext != ".png":
            try:
                pil = self.to_pillow()
            except ImportError:
                raise ValueError(f"{ext!r} needs Pillow, which is not installed; the native writer only "
                                 "writes .png") from None
            pil.save(path)
            return
        lib = _load_library()
        buf = np.frombuffer(self._rgba, dtype=np.uint8)
        if lib.rn_save_png(os.fsencode(path), self._width, self._height, _ptr(buf)) != 1:
            raise RuntimeError(f"rn_save_png failed to write {path!r}")

    def to_pillow(self):
        """The image as a ``PIL.Image.Image`` in RGBA mode. Pillow is optional
        and imported only here; ImportError says so when it is missing."""
        try:
            from PIL import Image as _PIL
        except ImportError as exc:
            raise ImportError("Image.to_pillow() needs Pillow (the PIL package)") from exc
        return _PIL.frombytes("RGBA", (self._width, self._height), self._rgba)

    def __repr__(self) -> str:
        return f"<rednoise.Image {self._width}x{self._height} RGBA8>"


# ------------------------------------------------------------------ scene --

class Scene:
    """A scene backed by a native ``rn_scene`` handle."""

    __slots__ = ("_handle", "_n", "__weakref__")

    def __init__(self, handle: int):
        if not handle:
            raise ValueError("Scene requires a non-NULL native handle")
        self._handle = handle
        self._n = int(_load_library().rn_scene_triangle_count(handle))

    @classmethod
    def load_obj(cls, path: Union[str, os.PathLike], scale: float = 0.35) -> "Scene":
        """Load a Wavefront OBJ (+ .mtl). Raises FileNotFoundError / RuntimeError."""
        path = os.fspath(path)
        if not os.path.isfile(path):
            raise FileNotFoundError(f"OBJ file not found: {path!r}")
        lib = _load_library()
        h = lib.rn_scene_load_obj(os.fsencode(path), C.c_float(scale))
        if not h:
            raise _err(f"rn_scene_load_obj({path!r})")
        return cls(h)

    @classmethod
    def load(cls, path: Union[str, os.PathLike]) -> "Scene":
        """Load a binary scene cache written by :meth:`save` (``rn_scene_load``)."""
        path = os.fspath(path)
        if not os.path.isfile(path):
            raise FileNotFoundError(f"scene cache not found: {path!r}")
        h = _load_library().rn_scene_load(os.fsencode(path))
        if not h:
            raise _err(f"rn_scene_load({path!r})")
        return cls(h)

    def save(self, path: Union[str, os.PathLike]) -> None:
        """Write this scene as a binary scene cache (``rn_scene_save``)."""
        if _load_library().rn_scene_save(self.handle, os.fsencode(os.fspath(path))) != 1:
            raise _err(f"rn_scene_save({os.fspath(path)!r})")

    @classmethod
    def from_arrays(cls, verts, albedo, emission=None) -> "Scene":
        """Scene from (n,3,3) vertices, (n,3) albedo and optional (n,3) emission."""
        v = _f32(verts)
        n = v.size // 9
        a = _f32(albedo, 3 * n)
        e = None if emission is None else _f32(emission, 3 * n)
        h = _load_library().rn_scene_from_arrays(_ptr(v), n, _ptr(a), _ptr(e))
        if not h:
            raise _err("rn_scene_from_arrays")
        return cls(h)

    @property
    def handle(self) -> int:
        self._check_open()
        return self._handle

    @property
    def triangle_count(self) -> int:
        self._check_open()
        return int(_load_library().rn_scene_triangle_count(self._handle))

    def render(self, mode: Union[RenderMode, int, str] = "pathtraced", width: int = 640, height: int = 480,
               cam_z: float = 4.0, samples: int = 64) -> Image:
        """``rn_render``: RGBA8 image of the fixed (0,0,cam_z) camera."""
        self._check_open()
        m = RenderMode.coerce(mode)
        width, height = int(width), int(height)
        if width <= 0 or height <= 0:
            raise ValueError(f"width and height must be positive, got {width}x{height}")
        buf = np.empty(width * height * 4, np.uint8)
        ok = _load_library().rn_render(self._handle, int(m), width, height, C.c_float(cam_z), int(samples), _ptr(buf))
        if ok != 1:
            raise _err(f"rn_render(mode={m.name}, {width}x{height})")
        return Image(buf.tobytes(), width, height)

    def render_seeded(self, width: int, height: int, cam_z: float, samples: int, key0: int, key1: int = 0) -> Image:
        """``rn_render_seeded``: the reference estimator with an explicit Philox key."""
        self._check_open()
        buf = np.empty(width * height * 4, np.uint8)
        if _load_library().rn_render_seeded(self._handle, width, height, C.c_float(cam_z), int(samples),
                                            int(key0) & 0xFFFFFFFF, int(key1) & 0xFFFFFFFF, _ptr(buf)) != 1:
            raise _err("rn_render_seeded")
        return Image(buf.tobytes(), width, height)

    def render_photon(self, photon_map: "PhotonMap", width: int, height: int, cam_z: float, gather_rays: int = 0,
                      key0: int = 0, key1: int = 0) -> Image:
        """``rn_render_photon``: pass 2 of the reference's photon mapper."""
        self._check_open()
        buf = np.empty(width * height * 4, np.uint8)
        if _load_library().rn_render_photon(self._handle, photon_map.handle, width, height, C.c_float(cam_z),
                                            int(gather_rays), int(key0) & 0xFFFFFFFF, int(key1) & 0xFFFFFFFF,
                                            _ptr(buf)) != 1:
            raise _err("rn_render_photon")
        return Image(buf.tobytes(), width, height)

    def render_bidirectional(self, width: int, height: int, cam_z: float, samples: int = 64, exposure: float = 40.0,
                             key0: int = 0, key1: int = 0) -> Image:
        """``rn_render_bidirectional``: the reference's bidirectional path tracer."""
        self._check_open()
        buf = np.empty(width * height * 4, np.uint8)
        if _load_library().rn_render_bidirectional(self._handle, width, height, C.c_float(cam_z), int(samples),
                                                   C.c_float(exposure), int(key0) & 0xFFFFFFFF,
                                                   int(key1) & 0xFFFFFFFF, _ptr(buf)) != 1:
            raise _err("rn_render_bidirectional")
        return Image(buf.tobytes(), width, height)

    def render_hdr(self, camera: Camera, params: RenderParams, out: Optional[np.ndarray] = None) -> np.ndarray:
        """``rn_render_ex``: (H, W, 3) float32 mean radiance, into `out` when
        given (a C-contiguous float32 array of width*height*3 values, reused
        across calls without new host pages)."""
        self._check_open()
        out = _out_f32(out, params.width * params.height * 3)
        if _load_library().rn_render_ex(self._handle, C.byref(camera), C.byref(params), _ptr(out)) != 1:
            raise _err("rn_render_ex")
        return out.reshape(params.height, params.width, 3)

    def adjoint(self, camera: Camera, params: RenderParams, dl_dimage) -> Tuple[np.ndarray, np.ndarray]:
        """``rn_adjoint``: (dL/d albedo, dL/d emission), each (n, 3)."""
        self._check_open()
        g = _f32(dl_dimage, params.width * params.height * 3)
        ga = np.empty(3 * self._n, np.float32)
        ge = np.empty(3 * self._n, np.float32)
        if _load_library().rn_adjoint(self._handle, C.byref(camera), C.byref(params), _ptr(g), _ptr(ga),
                                      _ptr(ge)) != 1:
            raise _err("rn_adjoint")
        return ga.reshape(-1, 3), ge.reshape(-1, 3)

    def set_materials(self, albedo=None, emission=None) -> None:
        a = None if albedo is None else _f32(albedo, 3 * self._n)
        e = None if emission is None else _f32(emission, 3 * self._n)
        if _load_library().rn_scene_set_materials(self._handle, _ptr(a), _ptr(e)) != 1:
            raise _err("rn_scene_set_materials")

    def triangles(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        v = np.empty(9 * self._n, np.float32)
        a = np.empty(3 * self._n, np.float32)
        e = np.empty(3 * self._n, np.float32)
        _load_library().rn_scene_get_triangles(self._handle, _ptr(v), _ptr(a), _ptr(e))
        return v.reshape(-1, 3, 3), a.reshape(-1, 3), e.reshape(-1, 3)

    def compat_attributes(self) -> dict:
        n = self._n
        nm = np.empty(3 * n, np.float32)
        col = np.empty(3 * n, np.int32)
        mat = np.empty(n, np.int32)
        tid = np.empty(n, np.int32)
        tc = np.empty(6 * n, np.float32)
        _load_library().rn_scene_get_compat(self._handle, _ptr(nm), _ptr(col), _ptr(mat), _ptr(tid), _ptr(tc))
        return {"normals": nm.reshape(-1, 3), "colour": col.reshape(-1, 3), "material": mat, "texture_id": tid,
                "texcoords": tc.reshape(-1, 6)}

    def vertex_normals(self) -> np.ndarray:
        """(n, 3, 3) smoothed vertex normals (face normals for array scenes)."""
        out = np.empty(9 * self._n, np.float32)
        if _load_library().rn_scene_get_vertex_normals(self._handle, _ptr(out)) != 1:
            raise _err("rn_scene_get_vertex_normals")
        return out.reshape(-1, 3, 3)

    def texture(self, tex_id: int) -> np.ndarray:
        w, h = C.c_int(), C.c_int()
        if _load_library().rn_scene_get_texture(self._handle, tex_id, C.byref(w), C.byref(h), None) != 1:
            raise _err("rn_scene_get_texture")
        px = np.empty(w.value * h.value, np.uint32)
        _load_library().rn_scene_get_texture(self._handle, tex_id, C.byref(w), C.byref(h), _ptr(px))
        return px.reshape(h.value, w.value)

    def set_builder(self, builder: str) -> None:
        """Select the BVH builder: "ploc" (default) or "lbvh". Drops the built BVH."""
        code = {"ploc": 0, "lbvh": 1}.get(builder)
        if code is None:
            raise ValueError(f"unknown BVH builder {builder!r}")
        if _load_library().rn_scene_set_builder(self._handle, code) != 1:
            raise _err("rn_scene_set_builder")

    def build(self, bvh_width: int = 8) -> dict:
        info = BvhInfo()
        if _load_library().rn_scene_build(self._handle, int(bvh_width), C.byref(info)) != 1:
            raise _err("rn_scene_build")
        return {k: getattr(info, k) for k, _ in BvhInfo._fields_}

    def export_bvh(self, bvh_width: int = 8) -> Tuple[np.ndarray, np.ndarray]:
        """Rebuild the BVH and return (nodes as (m, 80) bytes, leaf-ordered prim ids)."""
        info = self.build(bvh_width)
        nodes = np.empty(info["nodes"] * 80, np.uint8)
        prims = np.empty(self._n, np.int32)
        if _load_library().rn_scene_export_bvh(self._handle, _ptr(nodes), _ptr(prims)) != 1:
            raise _err("rn_scene_export_bvh")
        return nodes.reshape(-1, 80), prims

    def trace_rays(self, origins, dirs, tmax=None, ignore=None, query=QUERY_CLOSEST, tri_test=TRI_WATERTIGHT):
        """``rn_trace_rays``: returns (prim, t, u, v) for closest, occluded flags for any."""
        self._check_open()
        o = _f32(origins)
        d = _f32(dirs, o.size)
        n = o.size // 3
        tm = None if tmax is None else _f32(tmax, n)
        ig = None if ignore is None else np.ascontiguousarray(ignore, dtype=np.int32)
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        u = np.empty(n, np.float32)
        v = np.empty(n, np.float32)
        if _load_library().rn_trace_rays(self._handle, n, _ptr(o), _ptr(d), _ptr(tm), _ptr(ig), int(query),
                                         int(tri_test), _ptr(prim), _ptr(t), _ptr(u), _ptr(v)) != 1:
            raise _err("rn_trace_rays")
        if query == QUERY_ANY:
            return prim
        return prim, t, u, v

    def trace_primary_compat(self, width: int, height: int, cam_z: float = 4.0):
        n = width * height
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        u = np.empty(n, np.float32)
        v = np.empty(n, np.float32)
        tx = np.empty(n, np.int32)
        if _load_library().rn_trace_primary_compat(self._handle, width, height, C.c_float(cam_z), _ptr(prim), _ptr(t),
                                                   _ptr(u), _ptr(v), _ptr(tx)) != 1:
            raise _err("rn_trace_primary_compat")
        return prim, t, u, v, tx

    def close(self) -> None:
        """Release the native scene; safe to call more than once."""
        h = self._handle
        if h:
            self._handle = 0
            try:
                if _LIB is not None:
                    _LIB.rn_scene_free(h)
            except Exception:  # pragma: no cover - interpreter shutdown
                pass

    def _check_open(self) -> None:
        if not self._handle:
            raise ValueError("scene is closed")

    def __enter__(self) -> "Scene":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        self.close()

    def __repr__(self) -> str:
        return f"<rednoise.Scene {self._n} triangles>" if self._handle else "<rednoise.Scene (closed)>"


def render_batch(items: Sequence[Tuple[Scene, Camera, RenderParams]], options: Optional[BatchOptions] = None,
                 out: Optional[Sequence[np.ndarray]] = None):
    """``rn_render_batch``: list of (H, W, 3) images, one per item, into the
    arrays of `out` when given (as for ``Scene.render_hdr``)."""
    lib = _load_library()
    if out is not None and len(out) != len(items):
        raise ValueError("out must hold one array per item")
    outs = []
    arr = (BatchItem * len(items))()
    for k, (sc, cam, p) in enumerate(items):
        o = _out_f32(None if out is None else out[k], p.width * p.height * 3)
        outs.append(o)
        arr[k].scene = sc.handle
        arr[k].camera = cam
        arr[k].params = p
        arr[k].hdr_rgb = o.ctypes.data
    opts = options if options is not None else BatchOptions(0, 0, 8, 0, 0)
    if lib.rn_render_batch(arr, len(items), C.byref(opts)) != 1:
        raise _err("rn_render_batch")
    return [o.reshape(p.height, p.width, 3) for o, (_, _, p) in zip(outs, items)]


def lbvh_stages(verts):
    v = _f32(verts)
    n = v.size // 9
    codes = np.empty(n, np.uint32)
    order = np.empty(n, np.int32)
    left = np.empty(max(n - 1, 1), np.int32)
    right = np.empty(max(n - 1, 1), np.int32)
    if _load_library().rn_lbvh_stages(_ptr(v), n, _ptr(codes), _ptr(order), _ptr(left), _ptr(right)) != 1:
        raise _err("rn_lbvh_stages")
    return codes, order, left[: n - 1], right[: n - 1]


def ploc_stages(verts):
    """Sorted order and PLOC binary tree of the default builder (left/right:
    >= 0 internal node, < 0 leaf ~sorted position)."""
    v = _f32(verts)
    n = v.size // 9
    order = np.empty(n, np.int32)
    left = np.empty(max(n - 1, 1), np.int32)
    right = np.empty(max(n - 1, 1), np.int32)
    if _load_library().rn_ploc_stages(_ptr(v), n, _ptr(order), _ptr(left), _ptr(right)) != 1:
        raise _err("rn_ploc_stages")
    return order, left[: n - 1], right[: n - 1]


def radix_sort_pairs(keys, values, bits=32):
    k = np.ascontiguousarray(keys, dtype=np.uint32).copy()
    v = np.ascontiguousarray(values, dtype=np.uint32).copy()
    if _load_library().rn_radix_sort_pairs(_ptr(k), _ptr(v), k.size, int(bits)) != 1:
        raise _err("rn_radix_sort_pairs")
    return k, v


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1)
    k = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1)
    n = c.size // 4
    out = np.empty(4 * n, np.uint32)
    if _load_library().rn_philox4x32_10(n, _ptr(c), _ptr(k), _ptr(out)) != 1:
        raise _err("rn_philox4x32_10")
    return out.reshape(-1, 4)
