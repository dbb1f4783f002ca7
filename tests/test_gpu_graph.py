"""HIP graphs of a per-frame sequence (ptgpu_graph_*).

A recorded frame (camera upload from pinned memory, seed reset, colour clear, one fused render,
pixel readback into pinned memory) is replayed several times with different cameras written between
replays, with eager work and synchronisations in between; every replay must equal the eager render
of the same frame bit for bit.  Calls that wait for the device must refuse to be recorded.

An animated scene's frame -- new shapes, LBVH rebuild, treelet pass, render -- records too: the render
sizes its stack from the height the handle last read back and checks the built tree's height on the
device, so a replay on a deeper tree renders nothing and reports the error at the next synchronisation.
"""
import numpy as np
import pytest

import ptgpu_pc_amd as pt
from ptgpu_pc_amd import scenes
from ptgpu_pc_amd.layout import CAMERA_DTYPE, SHAPE_DTYPE, default_seeds

from test_gpu_parity import gpu_lbvh, render_gpu, same, say

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


def _frame_buffers(ctx, w, h):
    return dict(d_cam=ctx.alloc(CAMERA_DTYPE.itemsize, CAMERA_DTYPE), d_seeds0=ctx.to_device(default_seeds(w, h)),
                d_seeds=ctx.alloc(w * h * 8, np.uint32), d_col=ctx.alloc(w * h * 12, F32),
                d_pix=ctx.alloc(w * h * 4, np.uint32))


def _record_rebuild_frame(ctx, bvh, d_sh, h_sh, n, lc, w, h, spp, treelet, B, h_cam, h_pix):
    """Records one animated-scene frame: new shapes from pinned memory, LBVH rebuild (+ treelet pass),
    camera upload, seed reset, colour clear, render, pixel readback."""
    ctx.graph_begin()
    try:
        ctx.memcpy_async(d_sh, h_sh, n * SHAPE_DTYPE.itemsize)
        bvh.reset()
        bvh.buildRadixTree()
        bvh.buildBVHTree()
        if treelet:
            bvh.optimize(treelet)
        ctx.memcpy_async(B["d_cam"], h_cam, CAMERA_DTYPE.itemsize)
        ctx.memcpy_async(B["d_seeds"], B["d_seeds0"], w * h * 8)
        B["d_col"].zero()
        pt.render(ctx, bvh, None, colors=B["d_col"], seeds=B["d_seeds"], camera=B["d_cam"], shapes=d_sh, shapeCnt=n,
                  lightCnt=lc, width=w, height=h, currentSample=0, spp=spp, max_depth=5, pixels=B["d_pix"])
        ctx.memcpy_async(h_pix, B["d_pix"], w * h * 4)
    finally:
        g = ctx.graph_end()
    return g


def _eager(ctx, shapes, lc, cam, w, h, spp):
    s, bvh = gpu_lbvh(ctx, shapes)
    out = render_gpu(ctx, s, lc, cam, w, h, spp, 5, None, bvh=bvh)[:3]
    bvh.close()
    return out


@pytest.mark.parametrize("treelet", [0, 5])
def test_recorded_rebuild_frames_replay_equal_eager(ctx, treelet):
    """The per-frame sequence of an animated scene -- new shapes, LBVH rebuild, optional treelet pass,
    render -- recorded once, replayed with other shapes and cameras written between replays: every
    replay equals an eager render of its shapes on a freshly built tree, bit for bit."""
    w, h, spp, n = 80, 56, 2, 6002
    say(f"case graph rebuild replay treelet {treelet}")
    frames = []
    for seed, yaw in ((23, 0.3), (31, 1.1), (37, 2.0), (23, 0.3)):
        sh, lc, _ = scenes.soup(n - 2, seed=seed, width=w, height=h)
        cam = scenes.orbit_camera(yaw, w, h)
        pt.update_camera(cam)
        frames.append((sh, cam))
    d_sh = ctx.to_device(frames[0][0])
    bvh = pt.CLBVH(ctx, d_sh, n)
    bvh.buildRadixTree()
    bvh.buildBVHTree()
    if treelet:
        bvh.optimize(treelet)
    render_gpu(ctx, frames[0][0], lc, frames[0][1], w, h, 1, 5, None, bvh=bvh, d_shapes=d_sh)   # heights known
    B = _frame_buffers(ctx, w, h)
    h_sh = ctx.pinned(n * SHAPE_DTYPE.itemsize, SHAPE_DTYPE)
    h_cam = ctx.pinned(CAMERA_DTYPE.itemsize, CAMERA_DTYPE)
    h_pix = ctx.pinned(w * h * 4, np.uint32)
    g = _record_rebuild_frame(ctx, bvh, d_sh, h_sh, n, lc, w, h, spp, treelet, B, h_cam, h_pix)
    outs = []
    for k, (sh, cam) in enumerate(frames):
        h_sh.array[:] = sh
        h_cam.array[:] = cam
        g.launch()
        ctx.synchronize()
        got = h_pix.array.copy(), B["d_col"].download(F32), B["d_seeds"].download(np.uint32)
        c_e, s_e, p_e = _eager(ctx, sh, lc, cam, w, h, spp)
        say(f"replay {k}: pixels equal {same(got[0], p_e)}")
        assert same(got[0], p_e) and same(got[1], c_e) and same(got[2], s_e), k
        outs.append(got[0])
    assert same(outs[0], outs[3]) and not same(outs[0], outs[1])
    # the handle stays usable eagerly after the replays (its height is read back then)
    c_e, _, p_e = _eager(ctx, frames[3][0], lc, frames[3][1], w, h, spp)
    c_h, _, p_h = render_gpu(ctx, frames[3][0], lc, frames[3][1], w, h, spp, 5, None, bvh=bvh, d_shapes=d_sh)[:3]
    assert same(p_h, p_e) and same(c_h, c_e)
    for x in (g, h_sh, h_cam, h_pix, bvh):
        x.close()


def test_recorded_render_checks_the_built_height(ctx):
    """A frame recorded after a rebuild sizes its traversal stack from the height the handle last read
    back (plus a margin).  Replayed on shapes whose tree is deeper than that stack holds, the render
    checks the built height on the device, renders nothing, and the next ptgpu_synchronize reports it.
    Recorded again once the deep tree's height is known, the frame replays correctly."""
    w, h, spp, n = 80, 56, 2, 6002
    soup, lc, cam = scenes.soup(n - 2, seed=23, width=w, height=h)
    chain, lc2, _ = scenes.morton_chain(n - 2, seed=5, width=w, height=h)
    assert lc2 == lc
    pt.update_camera(cam)
    say("case graph height check")
    d_sh = ctx.to_device(soup)
    bvh = pt.CLBVH(ctx, d_sh, n)
    bvh.buildRadixTree()
    bvh.buildBVHTree()
    render_gpu(ctx, soup, lc, cam, w, h, 1, 5, None, bvh=bvh, d_shapes=d_sh)
    h_soup = bvh.export_traversal()[2]
    probe = pt.CLBVH(ctx, ctx.to_device(chain), n)
    probe.buildRadixTree()
    probe.buildBVHTree()
    h_chain = probe.export_traversal()[2]
    probe.close()
    say(f"heights: soup {h_soup}, chain {h_chain}")
    # the stack recorded for the soup: four-wide, sized for soup height + 4
    need = 3 * ((h_soup + 4 + 2) // 2) + 1
    stack = 32 if need <= 32 else 48 if need <= 48 else 64
    assert h_chain > 2 * ((stack - 1) // 3) - 1, (h_soup, h_chain, stack)

    B = _frame_buffers(ctx, w, h)
    h_sh = ctx.pinned(n * SHAPE_DTYPE.itemsize, SHAPE_DTYPE)
    h_cam = ctx.pinned(CAMERA_DTYPE.itemsize, CAMERA_DTYPE)
    h_pix = ctx.pinned(w * h * 4, np.uint32)
    h_cam.array[:] = cam
    g = _record_rebuild_frame(ctx, bvh, d_sh, h_sh, n, lc, w, h, spp, 0, B, h_cam, h_pix)
    h_sh.array[:] = soup
    g.launch()
    ctx.synchronize()
    assert same(h_pix.array, _eager(ctx, soup, lc, cam, w, h, spp)[2])
    # the deep tree: nothing rendered, the error reported once
    B["d_pix"].upload(np.full(w * h, 0xDEADBEEF, np.uint32))
    h_sh.array[:] = chain
    g.launch()
    with pytest.raises(pt.PtgpuError, match="deeper than the traversal stack"):
        ctx.synchronize()
    ctx.synchronize()
    assert (B["d_pix"].download(np.uint32) == 0xDEADBEEF).all()
    g.close()
    # eager use reads the deep tree's height back and renders it
    want = _eager(ctx, chain, lc, cam, w, h, spp)
    got = render_gpu(ctx, chain, lc, cam, w, h, spp, 5, None, bvh=bvh, d_shapes=d_sh)[:3]
    assert all(same(a, b) for a, b in zip(got, want))
    # recorded again, the frame holds the deep tree
    g = _record_rebuild_frame(ctx, bvh, d_sh, h_sh, n, lc, w, h, spp, 0, B, h_cam, h_pix)
    g.launch()
    ctx.synchronize()
    assert same(h_pix.array, want[2]) and same(B["d_col"].download(F32), want[0])
    for x in (g, h_sh, h_cam, h_pix, bvh):
        x.close()


def test_recorded_batch_rebuild_frames_replay_equal_eager(ctx):
    """A C4-style step -- several scenes' new shapes, one batched LBVH build, one batched render of
    all their frames -- recorded once and replayed with other shapes written between replays: every
    job equals an eager render of its shapes on a freshly built tree.  A replay in which one scene's
    tree is deeper than the recorded stack leaves that scene's frame unrendered (with the other jobs
    of its launch: the batch renders one launch per stack class) and reports the error."""
    spec = [(3000, 48, 32, 2), (6000, 64, 40, 1), (9000, 40, 40, 3)]   # triangles, width, height, spp
    say("case graph batch rebuild replay")
    scenes_k = [[scenes.soup(n, seed=seed + 10 * k, width=w, height=h)[0] for k in range(3)]
                for seed, (n, w, h, _) in zip((3, 5, 7), spec)]
    lc = 2
    cams = []
    for i, (_, w, h, _) in enumerate(spec):
        c = scenes.orbit_camera(0.4 + i, w, h)
        pt.update_camera(c)
        cams.append(c)
    d_sh = [ctx.to_device(v[0]) for v in scenes_k]
    hs = pt.CLBVH.create_batch(ctx, d_sh)
    h_sh = [ctx.pinned(len(v[0]) * SHAPE_DTYPE.itemsize, SHAPE_DTYPE) for v in scenes_k]
    jobs, outs = [], []
    for i, (n, w, h, spp) in enumerate(spec):
        B = _frame_buffers(ctx, w, h)
        B["d_cam"].upload(cams[i])
        jobs.append(dict(bvh=hs[i], shapes=d_sh[i], shapeCnt=n + 2, lightCnt=lc, width=w, height=h, spp=spp,
                         currentSample=0, max_depth=5, camera=B["d_cam"], colors=B["d_col"], seeds=B["d_seeds"],
                         pixels=B["d_pix"], B=B))
        outs.append(ctx.pinned(w * h * 4, np.uint32))

    def step():
        for i, j in enumerate(jobs):
            ctx.memcpy_async(d_sh[i], h_sh[i], (spec[i][0] + 2) * SHAPE_DTYPE.itemsize)
        pt.CLBVH.build_batch(ctx, hs)
        for j in jobs:
            ctx.memcpy_async(j["seeds"], j["B"]["d_seeds0"], j["width"] * j["height"] * 8)
            j["colors"].zero()
        pt.render_batch(ctx, [{k: v for k, v in j.items() if k != "B"} for j in jobs])
        for j, o in zip(jobs, outs):
            ctx.memcpy_async(o, j["pixels"], j["width"] * j["height"] * 4)

    for i in range(3):
        h_sh[i].array[:] = scenes_k[i][0]
    step()                                          # eager: heights read back, light lists built
    ctx.synchronize()
    ctx.graph_begin()
    try:
        step()
    finally:
        g = ctx.graph_end()
    for k in (1, 2, 0):
        for i in range(3):
            h_sh[i].array[:] = scenes_k[i][k]
        g.launch()
        ctx.synchronize()
        for i, (n, w, h, spp) in enumerate(spec):
            want = _eager(ctx, scenes_k[i][k], lc, cams[i], w, h, spp)
            assert same(outs[i].array, want[2]), (k, i)
            assert same(jobs[i]["colors"].download(F32), want[0]), (k, i)
            assert same(jobs[i]["seeds"].download(np.uint32), want[1]), (k, i)
    # one scene replaced by a deep Morton chain of the same size: no job is rendered
    chain, _, _ = scenes.morton_chain(spec[1][0], seed=5, width=spec[1][1], height=spec[1][2])
    h_sh[1].array[:] = chain
    for j in jobs:
        j["pixels"].upload(np.full(j["width"] * j["height"], 0xDEADBEEF, np.uint32))
    g.launch()
    with pytest.raises(pt.PtgpuError, match="deeper than the traversal stack"):
        ctx.synchronize()
    # the deep scene's job is not rendered (nor the other jobs of its stack class's launch)
    assert (jobs[1]["pixels"].download(np.uint32) == 0xDEADBEEF).all()
    g.close()
    for x in (*h_sh, *outs, *hs):
        x.close()
