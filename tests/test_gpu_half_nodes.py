"""Four-wide nodes with binary16 boxes rounded outward (tuning "wide" 2).

Each binary16 box contains its fp32 box, so every ray that hits the fp32 box hits the binary16 one;
the walk may test a few more boxes, and the closest hits, shadow-ray outcomes and so the images are
the same bits.  Checked against the fp32 four-wide nodes ("wide" 1) on scenes whose coordinates
land in every binary16 regime: normal numbers, subnormals and signed zeros (a scene straddling 0), beyond the
binary16 range (scaled by 1e5: bounds become +-65504 or infinities), and coarse ulps far from 0,
through ptgpu_render and ptgpu_render_batch, on LBVH, treelet-restructured and KD trees.
"""
import numpy as np
import pytest

import ptgpu_pc_amd as pt
from ptgpu_pc_amd import scenes

from test_gpu_batch import _run
from test_gpu_parity import render_gpu, say

pytestmark = pytest.mark.gpu


def _transform(shapes, cam, scale, shift):
    """Every position times `scale` plus `shift` (areas by scale^2), camera with it."""
    s = shapes.copy()
    for f in ("p1", "p2", "p3"):
        tri = s["type"] == 1
        s[f][tri] = s[f][tri] * np.float32(scale) + np.float32(shift)
    sph = s["type"] == 0
    s["p"][sph] = s["p"][sph] * np.float32(scale) + np.float32(shift)
    s["rad"][sph] = s["rad"][sph] * np.float32(scale)
    s["area"] = s["area"] * np.float32(scale * scale)
    c = cam.copy()
    c["orig"] = c["orig"] * np.float32(scale) + np.float32(shift)
    c["target"] = c["target"] * np.float32(scale) + np.float32(shift)
    return s, c


def _both(ctx, s, lc, cam, w, h, spp, bvh=None, kd=None):
    """fp32 nodes, binary16 nodes one lane per query, binary16 nodes with quads in sparse waves."""
    d_sh = ctx.to_device(s)
    out = []
    for wide, quad in ((1, 16), (2, 0), (2, 16)):
        ctx.set_tuning("wide", wide)
        ctx.set_tuning("quad", quad)
        ctx.set_tuning("blocks_per_cu", 1)   # few waves: many pixels per lane, long sparse tails
        out.append(render_gpu(ctx, s, lc, cam, w, h, spp, 5, None, bvh=bvh, kd=kd, d_shapes=d_sh)[:3])
    ctx.set_tuning("wide", 2)
    ctx.set_tuning("quad", 16)
    ctx.set_tuning("blocks_per_cu", 2)
    assert all(x.tobytes() == y.tobytes() for x, y in zip(out[1], out[2]))
    return out[:2]


@pytest.mark.parametrize("scene,scale,shift", [
    ("soup", 1.0, 0.0), ("soup", 1.0, -0.5), ("soup", 1e5, 0.0), ("soup", 3.0, -1000.0),
    ("ico", 1.0, 0.0), ("ico", 1e5, -7e4), ("cornell", 1.0, 0.0)])
def test_half_nodes_same_image(ctx, scene, scale, shift):
    say("case binary16 nodes", scene, scale, shift)
    w, h = 48, 40
    if scene == "soup":
        shapes, lc, cam = scenes.soup(6000, seed=21, width=w, height=h)
    elif scene == "ico":
        shapes, lc, cam = scenes.icosphere_in_box(4, seed=5, width=w, height=h)
    else:
        shapes, lc, cam = scenes.cornell(w, h)
    s, cam = _transform(shapes, cam, scale, shift)
    pt.update_camera(cam)
    bvh = pt.CLBVH(ctx, s)
    bvh.buildRadixTree()
    bvh.buildBVHTree()
    a, b = _both(ctx, s, lc, cam, w, h, 3, bvh=bvh)
    for x, y in zip(a, b):
        assert x.tobytes() == y.tobytes()
    assert a[0].any()
    if scene != "cornell":
        bvh.optimize(5)   # a restructured tree keeps one primitive per leaf
        a2, b2 = _both(ctx, s, lc, cam, w, h, 3, bvh=bvh)
        for x, y in zip(a2, b2):
            assert x.tobytes() == y.tobytes()
    bvh.close()
    kd = pt.KDTree(ctx, s, sah=True, max_depth=20, max_leaf=3)   # leaves with several primitives
    a3, b3 = _both(ctx, s, lc, cam, w, h, 2, kd=kd)
    for x, y in zip(a3, b3):
        assert x.tobytes() == y.tobytes()
    kd.close()


def test_half_nodes_batch(ctx):
    say("case binary16 nodes, batch")
    from test_gpu_batch import _jobs
    jobs, cache = _jobs(ctx)
    res = {}
    for wide, quad in ((1, 16), (2, 0), (2, 16)):
        ctx.set_tuning("wide", wide)
        ctx.set_tuning("quad", quad)
        res[(wide, quad)] = _run(ctx, jobs, batch=True)
    ctx.set_tuning("wide", 2)
    ctx.set_tuning("quad", 16)
    for other in ((2, 0), (2, 16)):
        for a, b in zip(res[(1, 16)], res[other]):
            for x, y in zip(a, b):
                assert x.tobytes() == y.tobytes()
    for _, bvh, kd, _ in cache.values():
        (bvh if bvh is not None else kd).close()


@pytest.mark.parametrize("n_tri", [0, 1, 3, 30])
def test_smallest_trees(ctx, n_tri):
    """Trees of 2 to 32 shapes (a single internal node upward).  A binary16 walk step has every lane
    load a primitive and a node, lanes with none reading primitive 0 and node 0; those exist in any
    tree with binary16 nodes.  All node formats give the same bits, and the oracle's image."""
    from test_gpu_parity import image_checks, render_oracle
    import pyoracle as O
    w, h = 40, 32
    shapes, lc, cam = scenes.soup(n_tri, seed=70 + n_tri, width=w, height=h)
    pt.update_camera(cam)
    say("case smallest tree", len(shapes), "shapes")
    bvh = pt.CLBVH(ctx, shapes)
    bvh.buildRadixTree()
    bvh.buildBVHTree()
    d_sh = ctx.to_device(shapes)
    out = []
    try:
        for wide, quad in ((2, 16), (2, 0), (1, 16), (0, 16)):
            ctx.set_tuning("wide", wide)
            ctx.set_tuning("quad", quad)
            out.append(render_gpu(ctx, shapes, lc, cam, w, h, 4, 5, None, bvh=bvh, d_shapes=d_sh)[:3])
    finally:
        ctx.set_tuning("wide", 2)
        ctx.set_tuning("quad", 16)
        bvh.close()
    for o in out[1:]:
        for x, y in zip(out[0], o):
            assert x.tobytes() == y.tobytes()
    _, _, btn, btl = O.build_lbvh(shapes)
    ref = render_oracle(shapes, lc, cam, w, h, 4, 5, O.accel_bvh(btn, btl))
    image_checks(*out[0], *ref)
