"""Trees built on one context's stream and rendered on another's.

A render whose CLBVH handle belongs to another context waits for that context's work on the handle
(the build and the traversal layout derived at the first render after it), so a caller can build
frame k+1 on one stream while frame k renders on another.  Checked bit for bit against the same
frames built and rendered on one stream: a handle rendered on a second context straight after its
build, six frames in flight over two handles whose shape arrays are refilled between uses (three
scenes in turn, so a render that read the other slot's tree or a refill that overtook the render
still using it changes the image), and seven frames three in flight (three handles, renders
alternating between two contexts), with and without treelet restructuring.
"""
import numpy as np
import pytest

import ptgpu_pc_amd as pt
from ptgpu_pc_amd import scenes
from ptgpu_pc_amd.layout import default_seeds

from test_gpu_parity import render_gpu, say

pytestmark = pytest.mark.gpu

W, H, SPP, DEPTH = 64, 48, 2, 5
F32 = np.float32


def _scenes():
    out = [scenes.soup(20000, seed=s, width=W, height=H) for s in (31, 32, 33)]
    assert len({len(x[0]) for x in out}) == 1 and len({x[1] for x in out}) == 1
    return out


def _serial(ctx, shapes, lc, cam, treelet):
    bvh = pt.CLBVH(ctx, shapes)
    bvh.buildRadixTree()
    bvh.buildBVHTree()
    if treelet:
        bvh.optimize(treelet)
    d_sh = ctx.to_device(shapes)
    img = render_gpu(ctx, shapes, lc, cam, W, H, SPP, DEPTH, None, bvh=bvh, d_shapes=d_sh)[:3]
    bvh.close()
    return img


def test_render_on_other_context(ctx):
    say("case build on context B, render on context A")
    shapes, lc, cam = scenes.soup(20000, seed=31, width=W, height=H)
    pt.update_camera(cam)
    ref = _serial(ctx, shapes, lc, cam, 0)
    cb = pt.Context(0)
    try:
        d_sh = ctx.to_device(shapes)
        ctx.synchronize()
        bvh = pt.CLBVH(cb, d_sh, len(shapes))
        for rebuild in range(2):
            if rebuild:
                bvh.reset()
            bvh.buildRadixTree()
            bvh.buildBVHTree()
            got = render_gpu(ctx, shapes, lc, cam, W, H, SPP, DEPTH, None, bvh=bvh, d_shapes=d_sh)[:3]
            for x, y in zip(got, ref):
                assert x.tobytes() == y.tobytes()
        bvh.close()
    finally:
        cb.close()


@pytest.mark.parametrize("treelet", [0, 5])
def test_two_frames_in_flight(ctx, treelet):
    say("case two frames in flight, treelet", treelet)
    sc = _scenes()
    n, lc = len(sc[0][0]), sc[0][1]
    cam = sc[0][2]
    pt.update_camera(cam)
    refs = [_serial(ctx, s, lc, cam, treelet) for s, _, _ in sc]
    cb = pt.Context(0)
    try:
        src = [ctx.to_device(s) for s, _, _ in sc]
        arr = [ctx.alloc(src[0].nbytes) for _ in range(2)]
        hs = [pt.CLBVH(cb, arr[i], n) for i in range(2)]
        d_cam = ctx.to_device(cam)
        seeds0 = ctx.to_device(default_seeds(W, H))
        outs = []
        for k in range(6):
            i, j = k & 1, k % 3
            # refill slot i on the render stream (after the render that last read it), then build on
            # the other stream once the refill is done
            arr[i].copy_from(src[j])
            ctx.record(10 + i)
            cb.wait_event(ctx, 10 + i)
            hs[i].reset()
            hs[i].buildRadixTree()
            hs[i].buildBVHTree()
            if treelet:
                hs[i].optimize(treelet)
            col = ctx.alloc(W * H * 12, F32).zero()
            seeds = ctx.alloc(W * H * 8, np.uint32).copy_from(seeds0)
            pix = ctx.alloc(W * H * 4, np.uint32)
            pt.render(ctx, hs[i], None, colors=col, seeds=seeds, camera=d_cam, shapes=arr[i], shapeCnt=n,
                      lightCnt=lc, width=W, height=H, currentSample=0, spp=SPP, max_depth=DEPTH, pixels=pix)
            outs.append((col, seeds, pix))
        ctx.synchronize()
        cb.synchronize()
        for k, (col, seeds, pix) in enumerate(outs):
            got = (col.download(F32), seeds.download(np.uint32), pix.download(np.uint32))
            for x, y in zip(got, refs[k % 3]):
                assert x.tobytes() == y.tobytes(), k
        assert refs[0][0].tobytes() != refs[1][0].tobytes()
        for h in hs:
            h.close()
    finally:
        cb.close()


@pytest.mark.parametrize("treelet", [0, 5])
def test_three_frames_in_flight(ctx, treelet):
    """The bench's three-in-flight schedule: three handles over their own shape arrays, frame k+1's tree
    built on a third context right after frame k's render is queued, renders alternating between two
    contexts with their own frame buffers (so frame k+1 renders while frame k drains).  Shape arrays
    are refilled with three scenes in turn; every frame equals its one-stream render bit for bit."""
    say("case three frames in flight, treelet", treelet)
    sc = _scenes()
    n, lc = len(sc[0][0]), sc[0][1]
    cam = sc[0][2]
    pt.update_camera(cam)
    refs = [_serial(ctx, s, lc, cam, treelet) for s, _, _ in sc]
    cb, r1 = pt.Context(0), pt.Context(0)
    try:
        rc = [ctx, r1]
        src = [ctx.to_device(s) for s, _, _ in sc]
        arr = [cb.alloc(src[0].nbytes) for _ in range(3)]
        hs = [pt.CLBVH(cb, arr[i], n) for i in range(3)]
        d_cam = ctx.to_device(cam)
        seeds0 = ctx.to_device(default_seeds(W, H))

        def build(sl, j, after):
            # the render that last read slot sl ran on context `after`; refill, then rebuild
            cb.wait_event(after, 10 + sl)
            arr[sl].copy_from(src[j])
            hs[sl].reset()
            hs[sl].buildRadixTree()
            hs[sl].buildBVHTree()
            if treelet:
                hs[sl].optimize(treelet)

        for c in rc:
            for sl in range(3):
                c.record(10 + sl)
        outs = []
        build(0, 0, ctx)
        for k in range(7):
            sl, r = k % 3, rc[k & 1]
            col = r.alloc(W * H * 12, F32).zero()
            seeds = r.alloc(W * H * 8, np.uint32).copy_from(seeds0)
            pix = r.alloc(W * H * 4, np.uint32)
            pt.render(r, hs[sl], None, colors=col, seeds=seeds, camera=d_cam, shapes=arr[sl], shapeCnt=n,
                      lightCnt=lc, width=W, height=H, currentSample=0, spp=SPP, max_depth=DEPTH, pixels=pix)
            r.record(10 + sl)
            outs.append((col, seeds, pix))
            build((k + 1) % 3, (k + 1) % 3, rc[(k + 1 - 3) & 1] if k >= 2 else ctx)
        for c in (ctx, r1, cb):
            c.synchronize()
        for k, (col, seeds, pix) in enumerate(outs):
            got = (col.download(F32), seeds.download(np.uint32), pix.download(np.uint32))
            for x, y in zip(got, refs[k % 3]):
                assert x.tobytes() == y.tobytes(), k
        for h in hs:
            h.close()
    finally:
        r1.close()
        cb.close()


def test_height_readback_off(ctx):
    """ptgpu_clbvh_set_height_readback(h, 0): after the first read-back, rebuilds are not waited for;
    frames equal the read-back frames bit for bit, with a treelet pass too.  A rebuild into a tree
    deeper than the bound renders nothing and the next synchronisation reports it; with the read-back
    back on, the deep tree renders."""
    say("case height read-back off")
    sc = _scenes()
    n, lc = len(sc[0][0]), sc[0][1]
    cam = sc[0][2]
    pt.update_camera(cam)
    d_cam = ctx.to_device(cam)
    d_sh = ctx.to_device(sc[0][0])
    bvh = pt.CLBVH(ctx, d_sh, n)
    bvh.set_height_readback(False)

    def frame(shapes_src, treelet):
        d_sh.upload(shapes_src)
        bvh.reset()
        bvh.buildRadixTree()
        bvh.buildBVHTree()
        if treelet:
            bvh.optimize(treelet)
        col = ctx.alloc(W * H * 12, F32).zero()
        seeds = ctx.to_device(default_seeds(W, H))
        pix = ctx.alloc(W * H * 4, np.uint32).upload(np.full(W * H, 0xDEADBEEF, np.uint32))
        pt.render(ctx, bvh, None, colors=col, seeds=seeds, camera=d_cam, shapes=d_sh, shapeCnt=n, lightCnt=lc,
                  width=W, height=H, currentSample=0, spp=SPP, max_depth=DEPTH, pixels=pix)
        return col, seeds, pix

    for treelet in (0, 5):
        refs = [_serial(ctx, s, lc, cam, treelet) for s, _, _ in sc]
        for k in range(4):
            col, seeds, pix = frame(sc[k % 3][0], treelet)
            ctx.synchronize()
            got = (col.download(F32), seeds.download(np.uint32), pix.download(np.uint32))
            for x, y in zip(got, refs[k % 3]):
                assert x.tobytes() == y.tobytes(), (treelet, k)
    chain, lc2, _ = scenes.morton_chain(n - 2, seed=5, width=W, height=H)
    assert lc2 == lc
    _, _, pix = frame(chain, 0)
    with pytest.raises(pt.PtgpuError, match="deeper than the traversal stack"):
        ctx.synchronize()
    assert (pix.download(np.uint32) == 0xDEADBEEF).all()
    bvh.set_height_readback(True)
    col, seeds, pix = frame(chain, 0)
    ctx.synchronize()
    want = _serial(ctx, chain, lc, cam, 0)
    assert pix.download(np.uint32).tobytes() == want[2].tobytes()
    bvh.close()
