"""Tiles and the hybrid CPU/GPU frame (CPU_PARTRENDERING: smallptGPU.cpp:636-884, RadianceBoxGPU
rendering_kernel.cl:1152-1236).

* Every tile of a frame through ptgpu_render_tiles (one launch) gives the frame ptgpu_render gives,
  edge tiles included; every box through ptgpu_radiance_box_gpu gives ptgpu_radiance_gpu's frame.
* Hybrid frames (tiles split between the GPU and host threads by the ratio rcpu) give the GPU-only
  frame's colours and seeds bit for bit, and the oracle's: the host runs the kernels' arithmetic.
  Pixels may differ by one step where the host's powf and the device's differ.
"""
import numpy as np
import pytest

import ptgpu_pc_amd as pt
import pyoracle as O
from ptgpu_pc_amd import scenes
from ptgpu_pc_amd.layout import default_seeds

from test_gpu_parity import gpu_lbvh, image_checks, render_gpu, render_oracle, same, say

pytestmark = pytest.mark.gpu
F32 = np.float32


def _scene(accel, w, h):
    if accel == "none":
        return scenes.cornell(w, h)
    if accel == "kd":
        return scenes.icosphere_in_box(3, seed=3, width=w, height=h)
    return scenes.soup(5000, seed=7, width=w, height=h)


def _trees(ctx, accel, shapes):
    if accel == "bvh":
        s, bvh = gpu_lbvh(ctx, shapes)
        return s, bvh, None
    if accel == "kd":
        return shapes, None, pt.KDTree(ctx, shapes)
    return shapes, None, None


def _buffers(ctx, w, h):
    return (ctx.alloc(w * h * 12, F32).zero(), ctx.to_device(default_seeds(w, h)), ctx.alloc(w * h * 4, np.uint32).zero())


def _grid(w, h, tw, th):
    return np.array([(x, y) for y in range(0, h, th) for x in range(0, w, tw)], dtype=np.int32)


@pytest.mark.parametrize("accel", ["bvh", "kd", "none"])
def test_all_tiles_equal_the_frame(ctx, accel):
    w, h, tw, th = 100, 70, 32, 24
    shapes, lc, cam = _scene(accel, w, h)
    pt.update_camera(cam)
    say(f"case tiles {accel} {w}x{h} in {tw}x{th}")
    s, bvh, kd = _trees(ctx, accel, shapes)
    d_sh, d_cam = ctx.to_device(s), ctx.to_device(cam)
    ref = render_gpu(ctx, s, lc, cam, w, h, 3, 5, None, bvh=bvh, kd=kd, d_shapes=d_sh)[:3]
    col, sd, pix = _buffers(ctx, w, h)
    tiles = _grid(w, h, tw, th)[::-1].copy()        # any order: tiles are independent
    pt.render_tiles(ctx, bvh, kd, tiles, tw, th, colors=col, seeds=sd, camera=d_cam, shapes=d_sh, shapeCnt=len(s),
                    lightCnt=lc, width=w, height=h, currentSample=0, spp=3, max_depth=5, pixels=pix)
    got = col.download(F32), sd.download(np.uint32), pix.download(np.uint32)
    for a, b in zip(ref, got):
        assert same(a, b)


def test_radiance_box_gpu_equals_radiance_gpu(ctx):
    w, h, bw, bh = 64, 40, 24, 16
    shapes, lc, cam = scenes.soup(2000, seed=5, width=w, height=h)
    pt.update_camera(cam)
    say("case RadianceBoxGPU boxes vs RadianceGPU")
    s, bvh = gpu_lbvh(ctx, shapes)
    btn, btl = bvh.getTrees()
    d_sh, d_cam, d_btn, d_btl = ctx.to_device(s), ctx.to_device(cam), ctx.to_device(btn), ctx.to_device(btl)
    ctx.set_max_depth(5)
    try:
        outs = []
        for boxed in (False, True):
            col, sd, pix = _buffers(ctx, w, h)
            for k in range(2):
                if boxed:
                    for x, y in _grid(w, h, bw, bh):
                        pt.radiance_gpu(ctx, pt.ACCEL_BVH, col, sd, d_cam, d_sh, len(s), lc, btn=d_btn, btl=d_btl,
                                        width=w, height=h, currentSample=k, pixels=pix, box=(x, y, bw, bh))
                else:
                    pt.radiance_gpu(ctx, pt.ACCEL_BVH, col, sd, d_cam, d_sh, len(s), lc, btn=d_btn, btl=d_btl,
                                    width=w, height=h, currentSample=k, pixels=pix)
            outs.append((col.download(F32), sd.download(np.uint32), pix.download(np.uint32)))
    finally:
        ctx.set_max_depth(6)
    for a, b in zip(*outs):
        assert same(a, b)
    with pytest.raises(pt.PtgpuError):
        pt.radiance_gpu(ctx, pt.ACCEL_BVH, col, sd, d_cam, d_sh, len(s), lc, btn=d_btn, btl=d_btl, width=w, height=h,
                        currentSample=0, pixels=pix, box=(w, 0, bw, bh))      # a box outside the frame
    bvh.close()


@pytest.mark.parametrize("accel", ["bvh", "kd", "none"])
def test_hybrid_frames_equal_gpu_frames_and_oracle(ctx, accel):
    """Four hybrid frames of one sample each (rcpu from 1.0, then the measured ratio), then two of two
    samples: colours and seeds equal the GPU-only render and the oracle's frame bit for bit."""
    w, h = 96, 72
    shapes, lc, cam = _scene(accel, w, h)
    pt.update_camera(cam)
    s, bvh, kd = _trees(ctx, accel, shapes)
    d_sh, d_cam = ctx.to_device(s), ctx.to_device(cam)
    hy = pt.Hybrid(ctx, 4)
    col, sd, pix = _buffers(ctx, w, h)
    rcpu, seen_cpu = 1.0, 0
    kw = dict(colors=col, seeds=sd, camera=d_cam, shapes=d_sh, shapeCnt=len(s), lightCnt=lc, width=w, height=h,
              max_depth=5, pixels=pix)
    for k in range(4):
        new, st = hy.frame(bvh, kd, 32, 24, rcpu, currentSample=k, spp=1, **kw)
        say(f"case hybrid {accel} frame {k}: rcpu {rcpu:.3f} -> {new:.3f}, tiles GPU {st['gpu_tiles']} "
            f"host {st['cpu_tiles']}, {st['gpu_ms']:.2f} / {st['cpu_ms']:.2f} ms")
        seen_cpu += st["cpu_tiles"]
        assert st["gpu_tiles"] >= 1 and st["gpu_tiles"] + st["cpu_tiles"] == 3 * 3
        if k == 0:
            assert st["cpu_tiles"] == 4            # rcpu 1: GPU and host tiles alternate
        rcpu = new
    for k in (4, 6):
        _, st = hy.frame(bvh, kd, 32, 24, 0.0, currentSample=k, spp=2, **kw)   # rcpu 0: the host takes all but one
        assert st["cpu_tiles"] == 8
        seen_cpu += st["cpu_tiles"]
    assert seen_cpu > 0
    got = col.download(F32), sd.download(np.uint32), pix.download(np.uint32)
    ref = render_gpu(ctx, s, lc, cam, w, h, 8, 5, None, bvh=bvh, kd=kd, d_shapes=d_sh)[:3]
    assert same(got[0], ref[0]) and same(got[1], ref[1])
    assert (np.abs(got[2].view(np.uint8).astype(int) - ref[2].view(np.uint8).astype(int)) <= 1).all()
    acc = O.accel_none()
    if accel == "bvh":
        _, _, obtn, obtl = O.build_lbvh(shapes)
        acc = O.accel_bvh(obtn, obtl)
    elif accel == "kd":
        n5, box, kn = O.KD(shapes).flatten_dfs()
        acc = O.accel_kd_dfs(n5, box, kn)
    image_checks(*got, *render_oracle(s, lc, cam, w, h, 8, 5, acc))
    hy.close()


def test_hybrid_all_gpu_when_the_host_is_slow(ctx):
    w, h = 64, 48
    shapes, lc, cam = scenes.soup(3000, seed=3, width=w, height=h)
    pt.update_camera(cam)
    s, bvh = gpu_lbvh(ctx, shapes)
    d_sh, d_cam = ctx.to_device(s), ctx.to_device(cam)
    hy = pt.Hybrid(ctx, 2)
    col, sd, pix = _buffers(ctx, w, h)
    _, st = hy.frame(bvh, None, 16, 16, 1e9, colors=col, seeds=sd, camera=d_cam, shapes=d_sh, shapeCnt=len(s),
                     lightCnt=lc, width=w, height=h, currentSample=0, spp=2, max_depth=5, pixels=pix)
    assert st["cpu_tiles"] == 0 and st["gpu_tiles"] == 12
    ref = render_gpu(ctx, s, lc, cam, w, h, 2, 5, None, bvh=bvh, d_shapes=d_sh)[:3]
    assert same(col.download(F32), ref[0]) and same(sd.download(np.uint32), ref[1])
    hy.close()
    bvh.close()
