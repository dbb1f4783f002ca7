"""ptgpu_radiance_gpu_noalloc: the RadianceGPU entry that allocates nothing (its scratch is a static
device array of the library) gives the bits of ptgpu_radiance_gpu, call after call, and refuses a tree
that does not fit its scratch before any device work."""
import numpy as np
import pytest

import ptgpu_pc_amd as pt
from ptgpu_pc_amd import scenes
from ptgpu_pc_amd.layout import default_seeds

from test_gpu_parity import gpu_lbvh, same, say

pytestmark = pytest.mark.gpu
F32 = np.float32


def _frames(ctx, accel, d_sh, n, lc, cam, w, h, spp, noalloc, **trees):
    col = ctx.alloc(w * h * 12, F32).zero()
    sd = ctx.to_device(default_seeds(w, h))
    pix = ctx.alloc(w * h * 4, np.uint32)
    d_cam = ctx.to_device(cam)
    for k in range(spp):
        pt.radiance_gpu(ctx, accel, col, sd, d_cam, d_sh, n, lc, width=w, height=h, currentSample=k, pixels=pix,
                        noalloc=noalloc, **trees)
    return col.download(F32), sd.download(np.uint32), pix.download(np.uint32)


def _scene(accel, n, w, h, seed):
    if accel == "none" and n < 100:
        return scenes.cornell(w, h)
    return scenes.soup(n, seed=seed, width=w, height=h)


@pytest.mark.parametrize("accel", ["bvh", "kd", "none"])
def test_noalloc_equals_radiance_gpu_on_two_calls(ctx, accel):
    """Two scenes in a row, the second smaller: the scratch keeps the first call's contents, and the
    second call must not depend on them."""
    sizes = {"bvh": ((20000, 64, 48), (3000, 40, 24)), "kd": ((600, 48, 40), (300, 32, 24)),
             "none": ((2000, 40, 32), (9, 32, 24))}[accel]
    ctx.set_max_depth(5)
    try:
        for call, (n_tri, w, h) in enumerate(sizes):
            shapes, lc, cam = _scene(accel, n_tri, w, h, seed=40 + call)
            pt.update_camera(cam)
            say(f"case noalloc {accel} call {call}: {len(shapes)} shapes, {w}x{h}x3")
            trees, code = {}, pt.ACCEL_NONE
            s = shapes
            if accel == "bvh":
                s, bvh = gpu_lbvh(ctx, shapes)
                btn, btl = bvh.getTrees()
                trees = dict(btn=ctx.to_device(btn), btl=ctx.to_device(btl))
                code = pt.ACCEL_BVH
                bvh.close()
            elif accel == "kd":
                kng, kn = pt.KDTree(ctx, shapes).getTrees()
                trees = dict(kng=ctx.to_device(kng), kngCnt=len(kng), kn=ctx.to_device(kn), knCnt=len(kn))
                code = pt.ACCEL_KD
            d_sh = ctx.to_device(s)
            ref = _frames(ctx, code, d_sh, len(s), lc, cam, w, h, 3, False, **trees)
            got = _frames(ctx, code, d_sh, len(s), lc, cam, w, h, 3, True, **trees)
            for a, b in zip(ref, got):
                assert same(a, b), (accel, call)
            assert ref[0].max() > 0
    finally:
        ctx.set_max_depth(6)


def test_noalloc_refuses_a_tree_beyond_its_scratch(ctx):
    """500,000 shapes need about 156 MB of traversal layout, above the 128 MiB scratch: refused with
    PTGPU_ERR_UNSUPPORTED.  The buffers passed are real and of the stated sizes."""
    shapes, lc, cam = scenes.soup(500000, seed=3, width=8, height=8)
    pt.update_camera(cam)
    say(f"case noalloc refusal: {len(shapes)} shapes")
    s, bvh = gpu_lbvh(ctx, shapes)
    d_btn, d_btl = bvh.device_trees()
    d_sh = ctx.to_device(s)
    col = ctx.alloc(8 * 8 * 12, F32).zero()
    sd = ctx.to_device(default_seeds(8, 8))
    pix = ctx.alloc(8 * 8 * 4, np.uint32)
    d_cam = ctx.to_device(cam)
    with pytest.raises(pt.PtgpuError) as e:
        pt.radiance_gpu(ctx, pt.ACCEL_BVH, col, sd, d_cam, d_sh, len(s), lc, btn=d_btn, btl=d_btl, width=8, height=8,
                        currentSample=0, pixels=pix, noalloc=True)
    assert e.value.status == 4 and b"scratch" in ctx.lib.ptgpu_last_error()
    assert not col.download(F32).any()          # refused before any device work
    pt.radiance_gpu(ctx, pt.ACCEL_BVH, col, sd, d_cam, d_sh, len(s), lc, btn=d_btn, btl=d_btl, width=8, height=8,
                    currentSample=0, pixels=pix)   # the allocating entry has no such limit
    bvh.close()
