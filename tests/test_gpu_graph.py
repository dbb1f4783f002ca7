"""HIP graphs of a per-frame sequence (ptgpu_graph_*).

A recorded frame (camera upload from pinned memory, seed reset, colour clear, one fused render,
pixel readback into pinned memory) is replayed several times with different cameras written between
replays, with eager work and synchronisations in between; every replay must equal the eager render
of the same frame bit for bit.  Calls that wait for the device must refuse to be recorded.
"""
import numpy as np
import pytest

import ptgpu_pc_amd as pt
from ptgpu_pc_amd import scenes
from ptgpu_pc_amd.layout import CAMERA_DTYPE, default_seeds

from test_gpu_parity import gpu_lbvh, render_gpu, say

pytestmark = pytest.mark.gpu
F32 = np.float32


def test_recorded_frame_replays_equal_eager(ctx):
    w, h, spp = 80, 56, 3
    shapes, lc, _ = scenes.soup(6000, seed=23, width=w, height=h)
    say("case graph replay")
    s, bvh = gpu_lbvh(ctx, shapes)
    d_sh = ctx.to_device(s)
    cams = []
    for yaw in (0.3, 1.1, 2.0, 0.3):
        c = scenes.orbit_camera(yaw, w, h)
        pt.update_camera(c)
        cams.append(c)
    render_gpu(ctx, s, lc, cams[0], w, h, 1, 5, None, bvh=bvh, d_shapes=d_sh)   # prepares the handle
    h_cam = ctx.pinned(CAMERA_DTYPE.itemsize, CAMERA_DTYPE)
    h_pix = ctx.pinned(w * h * 4, np.uint32)
    d_cam = ctx.alloc(CAMERA_DTYPE.itemsize, CAMERA_DTYPE)
    d_seeds0 = ctx.to_device(default_seeds(w, h))
    d_seeds = ctx.alloc(w * h * 8, np.uint32)
    d_col = ctx.alloc(w * h * 12, F32)
    d_pix = ctx.alloc(w * h * 4, np.uint32)

    ctx.graph_begin()
    try:
        with pytest.raises(pt.PtgpuError):
            ctx.synchronize()                       # waits for the device: refused, nothing recorded
        ctx.memcpy_async(d_cam, h_cam, CAMERA_DTYPE.itemsize)
        ctx.memcpy_async(d_seeds, d_seeds0, w * h * 8)
        d_col.zero()                                # a fill kernel
        pt.render(ctx, bvh, None, colors=d_col, seeds=d_seeds, camera=d_cam, shapes=d_sh, shapeCnt=len(s),
                  lightCnt=lc, width=w, height=h, currentSample=0, spp=spp, max_depth=5, pixels=d_pix)
        ctx.memcpy_async(h_pix, d_pix, w * h * 4)
    finally:
        g = ctx.graph_end()
    outs = []
    for k, cam in enumerate(cams):
        h_cam.array[:] = cam
        g.launch()
        ctx.synchronize()
        got = h_pix.array.copy()
        col = d_col.download(F32)
        # eager: the same frame from fresh buffers, with work and waits between the replays
        c_e, s_e, p_e = render_gpu(ctx, s, lc, cam, w, h, spp, 5, None, bvh=bvh, d_shapes=d_sh)[:3]
        assert got.tobytes() == p_e.tobytes(), k
        assert col.tobytes() == c_e.tobytes(), k
        assert d_seeds.download(np.uint32).tobytes() == s_e.tobytes(), k
        outs.append(got)
    assert outs[0].tobytes() == outs[3].tobytes()     # same camera, same frame
    assert outs[0].tobytes() != outs[1].tobytes()     # the camera written between replays is used
    g.close()
    h_cam.close()
    h_pix.close()
    bvh.close()


def test_recording_refuses_unprepared_traversal(ctx):
    shapes, lc, cam = scenes.soup(500, seed=4, width=16, height=16)
    pt.update_camera(cam)
    s, bvh = gpu_lbvh(ctx, shapes)      # built, never traversed: the height read-back is still pending
    d_sh, d_cam = ctx.to_device(s), ctx.to_device(cam)
    d_col, d_seeds = ctx.alloc(16 * 16 * 12, F32), ctx.to_device(default_seeds(16, 16))
    ctx.graph_begin()
    try:
        with pytest.raises(pt.PtgpuError):
            pt.render(ctx, bvh, None, colors=d_col, seeds=d_seeds, camera=d_cam, shapes=d_sh, shapeCnt=len(s),
                      lightCnt=lc, width=16, height=16, spp=1, max_depth=5)
        d_col.zero()                    # something recordable, so the graph is not empty
    finally:
        g = ctx.graph_end()
    g.close()
    # eager use afterwards is unaffected
    pt.render(ctx, bvh, None, colors=d_col, seeds=d_seeds, camera=d_cam, shapes=d_sh, shapeCnt=len(s), lightCnt=lc,
              width=16, height=16, spp=1, max_depth=5)
    ctx.synchronize()
    bvh.close()
