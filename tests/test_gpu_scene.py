"""The reference's own scene end to end: scenes/obj-model.txt (326 shapes: OBJ models, an emissive
quad, a mirror and a glass sphere) read by ptgpu_scene_read, built as an LBVH and a KD tree, rendered
on the GPU, checked against the oracle's brute-force render (RMSE <= 1e-3), and written as the
reference's PPM.
"""
from pathlib import Path

import numpy as np
import pytest

import ptgpu_pc_amd as pt
import pyoracle as O

from test_gpu_parity import image_checks, render_gpu, render_oracle, say

pytestmark = pytest.mark.gpu
SCENE_DIR = Path(__file__).resolve().parent / "golden" / "reference_scene"


def test_reference_scene_renders_like_the_oracle(ctx, tmp_path):
    w, h, spp, depth = 64, 48, 4, 6
    shapes, lc, cam = pt.read_scene(SCENE_DIR / "scenes" / "obj-model.txt", SCENE_DIR, w, h)
    pt.update_camera(cam)
    say("case reference scene obj-model.txt", len(shapes), "shapes")
    s = shapes.copy()
    bvh = pt.CLBVH(ctx, s)
    bvh.buildRadixTree()
    bvh.buildBVHTree()
    c_g, s_g, p_g = render_gpu(ctx, s, lc, cam, w, h, spp, depth, None, bvh=bvh)[:3]
    c_o, s_o, p_o = render_oracle(s, lc, cam, w, h, spp, depth, O.accel_none())
    image_checks(c_g, s_g, p_g, c_o, s_o, p_o)
    assert c_g.mean() > 0
    kd = pt.KDTree(ctx, shapes)
    c_k, s_k, p_k = render_gpu(ctx, shapes, lc, cam, w, h, spp, depth, None, kd=kd)[:3]
    assert c_k.tobytes() == c_g.tobytes() and p_k.tobytes() == p_g.tobytes()
    out = tmp_path / "image.ppm"
    pt.write_ppm(out, p_g, w, h)
    tok = out.read_text().split()
    assert tok[:4] == ["P3", str(w), str(h), "255"]
    rgb = np.array(tok[4:], dtype=np.int64).reshape(h, w, 3)[::-1]     # file rows run top-down
    px = p_g.reshape(h, w)
    assert (rgb[..., 0] == (px & 0xff)).all() and (rgb[..., 2] == ((px >> 16) & 0xff)).all()
    bvh.close()
    kd.close()
