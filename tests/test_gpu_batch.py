"""Batched rendering: many frames of different scenes in one launch (ptgpu_render_batch).

Every job's colours, seeds and packed pixels must be bit-identical to rendering that job alone
(ptgpu_render) -- the lanes of the batch walk each job's own tree with each job's own RNG stream.
"""
import numpy as np
import pytest

import ptgpu_pc_amd as pt
from ptgpu_pc_amd import scenes
from ptgpu_pc_amd.layout import default_seeds

from test_gpu_parity import say

pytestmark = pytest.mark.gpu
F32 = np.float32


def _scene(kind, seed, w, h):
    if kind == "soup":
        return scenes.soup(int(seed), seed=seed, width=w, height=h)
    if kind == "ico":
        return scenes.icosphere_in_box(3 + seed % 2, seed=seed, width=w, height=h)
    return scenes.cornell(w, h)


def _jobs(ctx):
    """(scene kind, size seed, w, h, spp, currentSample, tree) for a mixed set of frames."""
    spec = [("soup", 1500, 48, 32, 2, 0, "bvh"), ("soup", 9000, 64, 64, 4, 0, "bvh"), ("ico", 5, 40, 30, 1, 0, "bvh"),
            ("cornell", 0, 32, 32, 3, 0, "bvh"), ("soup", 700, 24, 40, 2, 3, "kd"), ("ico", 6, 56, 24, 2, 0, "kd"),
            ("soup", 9000, 20, 20, 5, 0, "bvh")]
    out = []
    cache = {}
    for kind, seed, w, h, spp, cs, tree in spec:
        shapes, lc, cam = _scene(kind, seed, w, h)
        pt.update_camera(cam)
        key = (kind, seed, tree)
        if key not in cache:   # the last job reuses the second job's tree and scene
            s = shapes.copy()
            if tree == "bvh":
                handle = pt.CLBVH(ctx, s)
                handle.buildRadixTree()
                handle.buildBVHTree()
                kd = None
            else:
                handle, kd = None, pt.KDTree(ctx, s)
            cache[key] = (s, handle, kd, ctx.to_device(s))
        s, bvh, kd, d_sh = cache[key]
        rng = np.random.default_rng(seed + w)
        colors0 = rng.random(w * h * 3).astype(F32) if cs > 0 else np.zeros(w * h * 3, F32)
        out.append(dict(bvh=bvh, kd=kd, shapes=d_sh, shapeCnt=len(s), lightCnt=lc, width=w, height=h, spp=spp,
                        currentSample=cs, max_depth=5, camera=ctx.to_device(cam), colors0=colors0,
                        seeds0=default_seeds(w, h) + np.uint32(seed % 7)))
    return out, cache


def _run(ctx, jobs, batch):
    bufs = []
    for j in jobs:
        w, h = j["width"], j["height"]
        b = dict(j, colors=ctx.to_device(j["colors0"]), seeds=ctx.to_device(j["seeds0"]),
                 pixels=ctx.alloc(w * h * 4, np.uint32).zero())
        for k in ("colors0", "seeds0"):
            b.pop(k)
        bufs.append(b)
    if batch:
        pt.render_batch(ctx, bufs)
    else:
        for b in bufs:
            pt.render(ctx, b["bvh"], b["kd"], **{k: v for k, v in b.items() if k not in ("bvh", "kd")})
    ctx.synchronize()
    return [(b["colors"].download(F32), b["seeds"].download(np.uint32), b["pixels"].download(np.uint32))
            for b in bufs]


def test_batch_equals_each_job_alone(ctx):
    say("case batch vs alone, 7 jobs")
    jobs, cache = _jobs(ctx)
    alone = _run(ctx, jobs, batch=False)
    mixed = _run(ctx, jobs, batch=True)
    for k, (a, b) in enumerate(zip(alone, mixed)):
        for x, y in zip(a, b):
            assert x.tobytes() == y.tobytes(), k
    assert all(a[0].mean() > 0 for a in alone)
    # a sparse grid (several pixels per lane, lanes crossing jobs) and extreme shading thresholds
    for regroup, bpc in ((1, 1), (64, 1)):
        ctx.set_tuning("batch_regroup", regroup)
        ctx.set_tuning("batch_blocks_per_cu", bpc)
        again = _run(ctx, jobs, batch=True)
        for a, b in zip(alone, again):
            for x, y in zip(a, b):
                assert x.tobytes() == y.tobytes()
    ctx.set_tuning("batch_regroup", 32)
    ctx.set_tuning("batch_blocks_per_cu", 3)
    for _, bvh, kd, _ in cache.values():
        (bvh if bvh is not None else kd).close()


def test_batch_rejects_bad_jobs(ctx):
    shapes, lc, cam = scenes.soup(100, seed=3, width=8, height=8)
    s = shapes.copy()
    bvh = pt.CLBVH(ctx, s)
    bvh.buildRadixTree()
    bvh.buildBVHTree()
    d_sh = ctx.to_device(s)
    base = dict(bvh=bvh, shapes=d_sh, shapeCnt=len(s), lightCnt=lc, width=8, height=8, spp=1, camera=ctx.to_device(cam),
                colors=ctx.alloc(8 * 8 * 12, F32), seeds=ctx.to_device(default_seeds(8, 8)))
    with pytest.raises(pt.PtgpuError):
        pt.render_batch(ctx, [dict(base, bvh=None)])             # no tree
    with pytest.raises(pt.PtgpuError):
        pt.render_batch(ctx, [dict(base, width=0)])              # empty frame
    pt.render_batch(ctx, [])                                     # nothing to do
    bvh.close()
