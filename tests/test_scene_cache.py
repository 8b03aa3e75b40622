"""Binary scene cache (rn_scene_save / rn_scene_load): round trips of every
asset and of array scenes, the on-disk layout read back independently with
numpy, and the error paths (corrupt, truncated, foreign, trailing bytes)."""
import os
import struct

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASSETS = os.path.join(ROOT, "tests", "golden", "assets")


@pytest.fixture(scope="module")
def rn():
    import rednoise_amd
    return rednoise_amd


def _same(a, b):
    va, aa, ea = a.triangles()
    vb, ab, eb = b.triangles()
    assert np.array_equal(va, vb) and np.array_equal(aa, ab) and np.array_equal(ea, eb)
    ca, cb = a.compat_attributes(), b.compat_attributes()
    assert ca.keys() == cb.keys()
    for k in ca:
        assert np.array_equal(np.asarray(ca[k]), np.asarray(cb[k])), k


@pytest.mark.parametrize("asset", ["cornell-box.obj", "textured-quad.obj", "sphere.obj"])
def test_round_trip_assets(rn, tmp_path, asset):
    a = rn.Scene.load_obj(os.path.join(ASSETS, asset), 0.35)
    p = tmp_path / "s.rnscene"
    a.save(p)
    b = rn.Scene.load(p)
    assert b.triangle_count == a.triangle_count
    _same(a, b)
    ids = set(np.asarray(a.compat_attributes()["texture_id"]).tolist()) - {-1}
    for t in ids:
        assert np.array_equal(a.texture(t), b.texture(t))


def test_round_trip_arrays_and_layout(rn, tmp_path):
    rng = np.random.default_rng(3)
    n = 1001
    v = rng.normal(size=(n, 3, 3)).astype(np.float32)
    al = rng.uniform(0, 1, size=(n, 3)).astype(np.float32)
    em = np.where(rng.uniform(size=(n, 1)) < 0.1, rng.uniform(0, 9, size=(n, 3)), 0).astype(np.float32)
    a = rn.Scene.from_arrays(v, al, em)
    p = tmp_path / "a.rnscene"
    a.save(p)
    _same(a, rn.Scene.load(p))
    # independent reader of the documented layout
    raw = p.read_bytes()
    assert raw[:8] == b"RNSCENE1"
    ver, cnt, ntex, _ = struct.unpack_from("<4I", raw, 8)
    (stored,) = struct.unpack_from("<Q", raw, 24)
    assert (ver, cnt, ntex) == (1, n, 0)
    body = raw[32:]
    off = 0
    arrays = {}
    for name, dt, k in [("verts", "<f4", 9), ("albedo", "<f4", 3), ("emission", "<f4", 3), ("normals", "<f4", 3),
                        ("vnormals", "<f4", 9), ("colour", "<i4", 3), ("material", "<i4", 1),
                        ("texture_id", "<i4", 1), ("texcoords", "<f4", 6)]:
        arrays[name] = np.frombuffer(body, dt, k * n, off)
        off += 4 * k * n
    assert off == len(body)
    assert np.array_equal(arrays["verts"], v.reshape(-1))
    assert np.array_equal(arrays["albedo"], al.reshape(-1))
    assert np.array_equal(arrays["emission"], em.reshape(-1))
    assert np.all(arrays["texture_id"] == -1)
    # a scene from arrays has no shared vertices: its vertex normals are its face normals
    assert np.array_equal(arrays["vnormals"].reshape(n, 3, 3), np.repeat(arrays["normals"].reshape(n, 1, 3), 3, 1))
    # checksum: fold 64-bit little-endian words (tail zero-padded), then the byte count
    mask = (1 << 64) - 1
    h = 0
    padded = body + b"\0" * (-len(body) % 8)
    for (w,) in struct.iter_unpack("<Q", padded):
        h = ((h ^ w) << 29 | (h ^ w) >> 35) & mask
        h = (h * 0x9E3779B97F4A7C15) & mask
    h ^= len(body)
    h = ((h << 29) | (h >> 35)) & mask
    h = (h * 0x9E3779B97F4A7C15) & mask
    assert h == stored


def test_corrupt_truncated_foreign(rn, tmp_path):
    a = rn.Scene.load_obj(os.path.join(ASSETS, "cornell-box.obj"), 0.35)
    p = tmp_path / "c.rnscene"
    a.save(p)
    raw = bytearray(p.read_bytes())
    bad = bytearray(raw)
    bad[100] ^= 0x40
    (tmp_path / "bad").write_bytes(bytes(bad))
    with pytest.raises(RuntimeError, match="checksum"):
        rn.Scene.load(tmp_path / "bad")
    (tmp_path / "short").write_bytes(bytes(raw[:-7]))
    with pytest.raises(RuntimeError, match="truncated"):
        rn.Scene.load(tmp_path / "short")
    (tmp_path / "long").write_bytes(bytes(raw) + b"x")
    with pytest.raises(RuntimeError, match="trailing"):
        rn.Scene.load(tmp_path / "long")
    (tmp_path / "obj").write_bytes(open(os.path.join(ASSETS, "cornell-box.obj"), "rb").read())
    with pytest.raises(RuntimeError, match="not a scene cache"):
        rn.Scene.load(tmp_path / "obj")
    with pytest.raises(FileNotFoundError):
        rn.Scene.load(tmp_path / "missing")
    # a header that claims 2^29 triangles over a few bytes of payload is
    # rejected from the file size, before any array is sized for it
    huge = bytearray(raw[:32]) + bytearray(64)
    struct.pack_into("<I", huge, 12, 1 << 29)
    (tmp_path / "huge").write_bytes(bytes(huge))
    with pytest.raises(RuntimeError, match="truncated"):
        rn.Scene.load(tmp_path / "huge")


@pytest.mark.parametrize("w,h", [(0, 395), (480, 0), (-1, 4)])
def test_empty_texture_rejected(rn, tmp_path, w, h):
    """A texture record of zero (or negative) width or height is refused: the
    texel lookup of a textured triangle needs at least one texel."""
    a = rn.Scene.load_obj(os.path.join(ASSETS, "textured-quad.obj"), 0.35)
    p = tmp_path / "q.rnscene"
    a.save(p)
    raw = bytearray(p.read_bytes())
    n, ntex = struct.unpack_from("<2I", raw, 12)
    assert ntex == 1
    off = 32 + n * (9 + 3 + 3 + 3 + 9 + 3 + 1 + 1 + 6) * 4
    assert struct.unpack_from("<2i", raw, off) == (480, 395)
    struct.pack_into("<2i", raw, off, w, h)
    (tmp_path / "e").write_bytes(bytes(raw))
    with pytest.raises(RuntimeError, match="texture size"):
        rn.Scene.load(tmp_path / "e")


@pytest.mark.parametrize("key,asset,scale", [("cornell", "cornell-box.obj", 0.35), ("quad", "textured-quad.obj", 0.35),
                                             ("sphere", "sphere.obj", 1.0)])
def test_vertex_normals_match_reference_loader(rn, key, asset, scale):
    """Smoothed vertex normals bit-identical to loadOBJ's
    (src/ObjLoader.cpp:178-205), golden from the compiled reference."""
    g = np.load(os.path.join(ROOT, "tests", "golden", f"whitted_{key}.npz"))
    sc = rn.Scene.load_obj(os.path.join(ASSETS, asset), scale)
    ours = sc.vertex_normals().reshape(-1)
    assert np.array_equal(ours.view(np.uint32), g["vnormals"].view(np.uint32))
