"""ptgpu_radiance_gpu: the RadianceGPU argument list over reference-layout trees (rendering_kernel.cl:1072-1150).

* Each accel mode (BVH btn/btl, KD kng/kn heap, none) is compared with the oracle directly, one
  RadianceGPU launch per sample as DrawFrame issues them (smallptGPU.cpp:1010-1087).
* A caller's tree may be much deeper than ceil(log2 N) (getTrees after optimize, or another builder):
  the traversal stack is sized from the tree's real height, so a deep tree renders the same image as a
  shallow tree over the same shapes.
* Malformed trees (child or shape index out of range, a node with two parents, an unreachable leaf)
  are refused before any kernel follows their links.
"""
import numpy as np
import pytest

import ptgpu_pc_amd as pt
import pyoracle as O
from ptgpu_pc_amd import scenes
from ptgpu_pc_amd.layout import BVHNODE_DTYPE, RAY_DTYPE, concat_shapes, default_seeds

from test_gpu_parity import gpu_lbvh, image_checks, render_gpu, render_oracle, same, say

pytestmark = pytest.mark.gpu
F32 = np.float32


def radiance_frames(ctx, accel, shapes_dev, n, lc, cam, w, h, spp, **trees):
    """`spp` ptgpu_radiance_gpu launches (currentSample 0..spp-1) into fresh frame buffers."""
    col = ctx.alloc(w * h * 12, F32).zero()
    sd = ctx.to_device(default_seeds(w, h))
    pix = ctx.alloc(w * h * 4, np.uint32)
    d_cam = ctx.to_device(cam)
    for k in range(spp):
        pt.radiance_gpu(ctx, accel, col, sd, d_cam, shapes_dev, n, lc, width=w, height=h, currentSample=k,
                        pixels=pix, **trees)
    return col.download(F32), sd.download(np.uint32), pix.download(np.uint32)


def test_radiance_gpu_bvh_vs_oracle(ctx):
    shapes, lc, cam = scenes.soup(3000, seed=4, width=48, height=40)
    pt.update_camera(cam)
    say("case radiance_gpu bvh soup3002 48x40x4 vs oracle")
    s, bvh = gpu_lbvh(ctx, shapes)
    btn, btl = bvh.getTrees()
    ctx.set_max_depth(5)
    try:
        got = radiance_frames(ctx, pt.ACCEL_BVH, ctx.to_device(s), len(s), lc, cam, 48, 40, 4,
                              btn=ctx.to_device(btn), btl=ctx.to_device(btl))
    finally:
        ctx.set_max_depth(6)
    _, _, obtn, obtl = O.build_lbvh(shapes)
    image_checks(*got, *render_oracle(s, lc, cam, 48, 40, 4, 5, O.accel_bvh(obtn, obtl)))
    bvh.close()


def test_radiance_gpu_kd_heap_vs_oracle(ctx):
    shapes, lc, cam = scenes.soup(600, seed=9, width=40, height=32)
    pt.update_camera(cam)
    say("case radiance_gpu kd heap soup602 40x32x4 vs oracle")
    kd = pt.KDTree(ctx, shapes)
    kng, kn = kd.getTrees()
    ctx.set_max_depth(5)
    try:
        got = radiance_frames(ctx, pt.ACCEL_KD, ctx.to_device(shapes), len(shapes), lc, cam, 40, 32, 4,
                              kng=ctx.to_device(kng), kngCnt=len(kng), kn=ctx.to_device(kn), knCnt=len(kn))
    finally:
        ctx.set_max_depth(6)
    okng, okn = O.KD(shapes).export_heap()
    assert same(okng, kng) and same(okn, kn)
    image_checks(*got, *render_oracle(shapes, lc, cam, 40, 32, 4, 5, O.accel_kd_heap(okng, okn)))


def test_radiance_gpu_none_vs_oracle(ctx):
    shapes, lc, cam = scenes.cornell(64, 48)
    pt.update_camera(cam)
    say("case radiance_gpu none cornell 64x48x3 vs oracle")
    ctx.set_max_depth(5)
    try:
        got = radiance_frames(ctx, pt.ACCEL_NONE, ctx.to_device(shapes), len(shapes), lc, cam, 64, 48, 3)
    finally:
        ctx.set_max_depth(6)
    image_checks(*got, *render_oracle(shapes, lc, cam, 64, 48, 3, 5, O.accel_none()))


# ---------------------------------------------------------------------------------------------
def karras_export(shapes, order, split):
    """Reference-layout arrays of the binary tree that split(lo, hi) describes over the leaves in
    `order`, in the numbering the reference's arrays use (BVH.cl:141-208): the node over leaves
    [lo, hi] split after g has children g and g+1, a child being leaf g when g == lo (leaf g+1 when
    g+1 == hi).  Bounds are unions of the leaves' boxes.  Returns (btn, btl, height)."""
    n = len(shapes)
    btn = np.zeros(n - 1, dtype=BVHNODE_DTYPE)
    btl = np.zeros(n, dtype=BVHNODE_DTYPE)
    btl["nShape"] = order
    btl["min"] = btl["max"] = np.arange(n)
    btl["bound"] = shapes["b"][order]
    visit, stack, height = [], [(0, 0, n - 1, 0)], 0
    while stack:
        i, lo, hi, d = stack.pop()
        g = split(lo, hi)
        assert lo <= g < hi
        btn["min"][i], btn["max"][i], btn["nLeft"][i], btn["nRight"][i] = lo, hi, g, g + 1
        btn["leaf"][i] = 2 if (g == lo or g + 1 == hi) else 0
        visit.append(i)
        for c, clo, chi in ((g, lo, g), (g + 1, g + 1, hi)):
            if clo == chi:
                btl["nParent"][c] = i
                height = max(height, d + 1)
            else:
                btn["nParent"][c] = i
                stack.append((c, clo, chi, d + 1))
    bl = btl["bound"]
    bn = btn["bound"]
    for i in reversed(visit):   # children before parents
        g, lo, hi = int(btn["nLeft"][i]), int(btn["min"][i]), int(btn["max"][i])
        a = bl[g] if g == lo else bn[g]
        b = bl[g + 1] if g + 1 == hi else bn[g + 1]
        bn[i] = [min(a[0], b[0]), max(a[1], b[1]), min(a[2], b[2]), max(a[3], b[3]), min(a[4], b[4]),
                 max(a[5], b[5])]
    return btn, btl, height


def test_radiance_gpu_deep_tree_same_image(ctx):
    """A caterpillar tree (a spine taking a fifth of the remaining leaves at every level) is far deeper
    than ceil(log2 N); the image must equal the one rendered over the library's own shallow export."""
    n_tri = 20000
    shapes, lc, cam = scenes.soup(n_tri, seed=12, width=64, height=48)
    pt.update_camera(cam)
    s, bvh = gpu_lbvh(ctx, shapes)
    n = len(s)
    btn0, btl0 = bvh.getTrees()

    def spine(lo, hi):
        size = hi - lo + 1
        if hi == n - 1 and size > 64:
            return lo + max(1, size // 5) - 1
        return lo + size // 2 - 1

    btn, btl, height = karras_export(s, bvh.sorted_indices(), spine)
    log2n = int(np.ceil(np.log2(n)))
    say(f"case radiance_gpu deep tree: height {height} (ceil log2 N = {log2n})")
    assert height > log2n + 10
    d_sh = ctx.to_device(s)
    ctx.set_max_depth(5)
    try:
        shallow = radiance_frames(ctx, pt.ACCEL_BVH, d_sh, n, lc, cam, 64, 48, 3, btn=ctx.to_device(btn0),
                                  btl=ctx.to_device(btl0))
        deep = radiance_frames(ctx, pt.ACCEL_BVH, d_sh, n, lc, cam, 64, 48, 3, btn=ctx.to_device(btn),
                               btl=ctx.to_device(btl))
    finally:
        ctx.set_max_depth(6)
    for a, b in zip(shallow, deep):
        assert same(a, b)
    # the oracle's traversal of the deep arrays agrees with brute force on camera-like rays
    rng = np.random.default_rng(3)
    rays = np.zeros(2000, dtype=RAY_DTYPE)
    rays["o"] = (rng.random((2000, 3)) * 1.2 - 0.1).astype(F32)
    d = rng.normal(size=(2000, 3))
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(F32)
    t_b, id_b = O.trace_closest(s, O.accel_none(), rays)
    t_e, id_e = O.trace_closest(s, O.accel_bvh(btn, btl), rays)
    assert (id_b == id_e).all() and (t_b == t_e).all()
    bvh.close()


def test_radiance_gpu_optimised_export_clustered(ctx):
    """getTrees after optimize on a clustered scene (duplicate Morton codes in the cluster): the
    exported tree renders through ptgpu_radiance_gpu the image the handle renders."""
    rng = np.random.default_rng(7)
    m = 12000
    cen = np.concatenate([rng.random((m // 2, 3)) * 1e-3 + 0.5,                     # one Morton cell
                          0.5 + 0.5 * (rng.random((m // 2, 3)) - 0.5) ** 3 * 8]).astype(F32)
    tri = cen[:, None, :] + (rng.random((m, 3, 3)) * 2e-4 - 1e-4).astype(F32)
    body = scenes.triangles_to_shapes(tri.astype(F32))
    quad = scenes.soup(2, seed=1, width=64, height=48)[0][-2:]   # the emissive quad of the soup scenes
    shapes = concat_shapes([body, quad])
    lc = scenes.light_count(shapes)
    cam = scenes.soup(2, seed=1, width=64, height=48)[2]
    pt.update_camera(cam)
    s, bvh = gpu_lbvh(ctx, shapes)
    d_sh = ctx.to_device(s)
    img = render_gpu(ctx, s, lc, cam, 64, 48, 3, 5, None, bvh=bvh, d_shapes=d_sh)[:3]
    bvh.optimize(7)
    btn, btl = bvh.getTrees()
    n = len(s)
    # height of the exported tree from its parent links
    par = np.concatenate([btn["nParent"], btl["nParent"]])
    depth = np.zeros(2 * n - 1, dtype=np.int64)
    for k in range(n - 1, 2 * n - 1):
        v, d = k, 0
        while v != 0:
            v = par[v]
            d += 1
            assert d < 4 * n
        depth[k] = d
    height = int(depth.max())
    say(f"case radiance_gpu optimised clustered export: height {height} (ceil log2 N = {int(np.ceil(np.log2(n)))})")
    ctx.set_max_depth(5)
    try:
        got = radiance_frames(ctx, pt.ACCEL_BVH, d_sh, n, lc, cam, 64, 48, 3, btn=ctx.to_device(btn),
                              btl=ctx.to_device(btl))
    finally:
        ctx.set_max_depth(6)
    for a, b in zip(img, got):
        assert same(a, b)
    bvh.close()


@pytest.mark.parametrize("fault", ["child_range", "two_parents", "unreachable", "nshape_range"])
def test_radiance_gpu_refuses_malformed_tree(ctx, fault):
    shapes, lc, cam = scenes.soup(500, seed=2, width=16, height=16)
    pt.update_camera(cam)
    s, bvh = gpu_lbvh(ctx, shapes)
    btn, btl = bvh.getTrees()
    n = len(s)
    internal = [i for i in range(n - 1) if btn["nLeft"][i] != btn["min"][i] and btn["nRight"][i] != btn["max"][i]]
    i = internal[len(internal) // 2]
    if fault == "child_range":
        btn["nLeft"][i] = n + 7          # not a leaf (min differs), and past the internal nodes
    elif fault == "two_parents":
        btn["nRight"][i] = btn["nLeft"][i]
    elif fault == "unreachable":
        btn["nLeft"][0] = btn["min"][0]  # the root's left child becomes leaf 0: its subtree is cut off
    else:
        btl["nShape"][n // 3] = n
    say("case radiance_gpu malformed tree", fault)
    with pytest.raises(pt.PtgpuError):
        radiance_frames(ctx, pt.ACCEL_BVH, ctx.to_device(s), n, lc, cam, 16, 16, 1, btn=ctx.to_device(btn),
                        btl=ctx.to_device(btl))
    bvh.close()
