"""rn_render(RN_RAYTRACED): the reference's Whitted renderer against images
of the compiled reference on the three assets. The renderer is
deterministic and combines colours in the reference's order; pixels could
differ only where the device powf (specular, Fresnel) rounds differently from
the host's.

The centre column and row (x = W/2, y = H/2) are compared separately: there
the primary ray has an exactly zero direction component, and the
reference's slab test computes 0 * inf = NaN for every BVH node with a face
on the ray's plane (src/BVH.cpp:66-75), so it misses geometry that the ray
hits (on sphere.obj it shows sky across the sphere's centre column). This
library's traversal does not drop those hits."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASSETS = os.path.join(ROOT, "tests", "golden", "assets")
CASES = [("cornell", "cornell-box.obj", 0.35), ("quad", "textured-quad.obj", 0.35), ("sphere", "sphere.obj", 1.0)]


@pytest.mark.parametrize("key,asset,scale", CASES)
@pytest.mark.parametrize("size", [(160, 120), (96, 64)])
def test_whitted_matches_reference(rn, key, asset, scale, size):
    W, H = size
    g = np.load(os.path.join(ROOT, "tests", "golden", f"whitted_{key}.npz"))
    ref = g[f"rgba_{W}x{H}"].astype(np.int32)
    sc = rn.Scene.load_obj(os.path.join(ASSETS, asset), scale)
    img = sc.render("raytraced", W, H, float(g["cam_z"]), 1).to_numpy().astype(np.int32)
    d = np.abs(img - ref).max(axis=2)
    axis = np.zeros((H, W), bool)
    axis[:, W // 2] = True
    axis[H // 2, :] = True
    same = (d[~axis] == 0).mean()
    near = (d[~axis] <= 2).mean()
    assert np.all(img[..., 3] == 255)
    assert same > 0.9995 and near == 1.0, (key, W, H, same, near, d[~axis].max())
    if key != "sphere":  # no node face lies on an axis plane of these rays: the lines agree too
        assert (d[axis] == 0).mean() > 0.99


def test_whitted_deterministic_and_sized(rn):
    sc = rn.Scene.load_obj(os.path.join(ASSETS, "cornell-box.obj"), 0.35)
    a = sc.render("raytraced", 64, 48, 4.0, 1).to_numpy()
    b = sc.render("raytraced", 64, 48, 4.0, 7).to_numpy()  # samples is ignored, as in the reference
    assert np.array_equal(a, b)
    assert a.shape == (48, 64, 4)
