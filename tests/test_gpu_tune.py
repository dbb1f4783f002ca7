"""Autotuner (ptgpu_autotune / ptgpu_autotune_batch).

The tuned parameters change speed only, so what must hold is: the chosen values are in the
candidate sets, the caller's frame buffers are untouched by the trials, the context and handle
carry the result, and renders after tuning are bit-identical to renders before it.
"""
import numpy as np
import pytest

import ptgpu_pc_amd as pt
from ptgpu_pc_amd import scenes
from ptgpu_pc_amd.layout import default_seeds

from test_gpu_parity import gpu_lbvh, render_gpu, say

pytestmark = pytest.mark.gpu
F32 = np.float32
DEFAULTS = {"regroup": 16, "blocks_per_cu": 2, "wide": 2, "quad": 16, "batch_regroup": 32, "batch_blocks_per_cu": 3}


def _restore_defaults(ctx):
    for k, v in DEFAULTS.items():
        ctx.set_tuning(k, v)


def test_autotune_frame(ctx):
    w, h = 96, 64
    shapes, lc, cam = scenes.soup(20000, seed=31, width=w, height=h)
    pt.update_camera(cam)
    say("case autotune frame")
    s, bvh = gpu_lbvh(ctx, shapes)
    d_sh = ctx.to_device(s)
    img0 = render_gpu(ctx, s, lc, cam, w, h, 3, 5, None, bvh=bvh, d_shapes=d_sh)[:3]
    rng = np.random.default_rng(3)
    col0 = rng.random(w * h * 3).astype(F32)
    seed0 = default_seeds(w, h) + np.uint32(9)
    colors, seeds = ctx.to_device(col0), ctx.to_device(seed0)
    pixels = ctx.alloc(w * h * 4, np.uint32).zero()
    try:
        t = pt.autotune(ctx, bvh, frames_per_build=4, trials=1, colors=colors, seeds=seeds, camera=ctx.to_device(cam),
                        shapes=d_sh, shapeCnt=len(s), lightCnt=lc, width=w, height=h, currentSample=2, spp=2,
                        max_depth=5, pixels=pixels)
        say(f"  {t}")
        assert t["blocks_per_cu"] in (1, 2, 3, 4) and t["regroup"] in (8, 16, 32) and t["wide"] in (0, 1, 2)
        assert t["digit_bits"] in (4, 8) and t["treelet"] in (0, 5, 7)
        assert t["render_ms"] > 0 and t["build_ms"] > 0
        # the caller's buffers are untouched
        assert colors.download(F32).tobytes() == col0.tobytes()
        assert seeds.download(np.uint32).tobytes() == seed0.tobytes()
        assert not pixels.download(np.uint32).any()
        # the tuned context and tree render the same image
        img1 = render_gpu(ctx, s, lc, cam, w, h, 3, 5, None, bvh=bvh, d_shapes=d_sh)[:3]
        for a, b in zip(img0, img1):
            assert a.tobytes() == b.tobytes()
        # and a rebuild with the tuned digit width reproduces the reference-layout sort
        bvh.reset()
        bvh.buildRadixTree()
        bvh.buildBVHTree()
        s2, ref = gpu_lbvh(ctx, shapes)
        assert np.array_equal(bvh.sorted_indices(), ref.sorted_indices())
        ref.close()
    finally:
        _restore_defaults(ctx)
        bvh.close()


def test_autotune_overlapped_builds(ctx):
    """frames_per_build 0 (each rebuild beside the previous frame's render): the render alone is
    minimised, over treelet sizes whose rebuild + optimise is at most 1/8 of the render, and over one
    or two workgroups per CU (the next frame's render needs room beside this one)."""
    w, h = 128, 96
    shapes, lc, cam = scenes.soup(20000, seed=32, width=w, height=h)
    pt.update_camera(cam)
    say("case autotune overlapped builds")
    s, bvh = gpu_lbvh(ctx, shapes)
    d_sh = ctx.to_device(s)
    img0 = render_gpu(ctx, s, lc, cam, w, h, 4, 5, None, bvh=bvh, d_shapes=d_sh)[:3]
    try:
        t = pt.autotune(ctx, bvh, frames_per_build=0, trials=1, colors=ctx.alloc(w * h * 12, F32).zero(),
                        seeds=ctx.to_device(default_seeds(w, h)), camera=ctx.to_device(cam), shapes=d_sh,
                        shapeCnt=len(s), lightCnt=lc, width=w, height=h, currentSample=0, spp=4, max_depth=5,
                        pixels=ctx.alloc(w * h * 4, np.uint32).zero())
        say(f"  {t}")
        assert t["treelet"] in (0, 5, 7)
        assert t["treelet"] == 0 or t["build_ms"] <= 0.125 * t["render_ms"] * (1 + 1e-6)
        assert t["blocks_per_cu"] in (1, 2)
        img1 = render_gpu(ctx, s, lc, cam, w, h, 4, 5, None, bvh=bvh, d_shapes=d_sh)[:3]
        for a, b in zip(img0, img1):
            assert a.tobytes() == b.tobytes()
    finally:
        _restore_defaults(ctx)
        bvh.close()


def test_autotune_batch(ctx):
    say("case autotune batch")
    jobs, keep = [], []
    for k, (n, w, h) in enumerate([(3000, 48, 32), (15000, 64, 64), (800, 40, 24)]):
        shapes, lc, cam = scenes.soup(n, seed=50 + k, width=w, height=h)
        pt.update_camera(cam)
        s, bvh = gpu_lbvh(ctx, shapes)
        keep.append(bvh)
        jobs.append(dict(bvh=bvh, shapes=ctx.to_device(s), shapeCnt=len(s), lightCnt=lc, width=w, height=h, spp=2,
                         max_depth=5, camera=ctx.to_device(cam), colors=ctx.alloc(w * h * 12, F32).zero(),
                         seeds=ctx.to_device(default_seeds(w, h)), pixels=ctx.alloc(w * h * 4, np.uint32).zero()))
    try:
        pt.render_batch(ctx, jobs)
        ctx.synchronize()
        before = [(j["colors"].download(F32), j["seeds"].download(np.uint32), j["pixels"].download(np.uint32))
                  for j in jobs]
        t = pt.autotune_batch(ctx, jobs, trials=1)
        say(f"  {t}")
        assert t["batch_blocks_per_cu"] in (2, 3, 4) and t["batch_regroup"] in (16, 32, 64)
        for j, (c, s_, p) in zip(jobs, before):   # untouched by the trials
            assert j["colors"].download(F32).tobytes() == c.tobytes()
            assert j["seeds"].download(np.uint32).tobytes() == s_.tobytes()
        # same frames again from the same start, with the tuned launch shape
        for j in jobs:
            j["colors"].zero()
            j["seeds"].upload(default_seeds(j["width"], j["height"]))
        pt.render_batch(ctx, jobs)
        ctx.synchronize()
        for j, (c, s_, p) in zip(jobs, before):
            assert j["colors"].download(F32).tobytes() == c.tobytes()
            assert j["pixels"].download(np.uint32).tobytes() == p.tobytes()
    finally:
        _restore_defaults(ctx)
        for b in keep:
            b.close()
