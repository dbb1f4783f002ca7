"""Served rendering (ptgpu_server_*): continuous batching with admission by HBM footprint.

Each request's result -- served in a mix with other requests, split over ticks, on a scene that may
have been built for an earlier request -- must be bit-identical to rendering that request alone with
a fresh LBVH and one fused ptgpu_render.
"""
import numpy as np
import pytest

import ptgpu_pc_amd as pt
from ptgpu_pc_amd import scenes
from ptgpu_pc_amd.layout import default_seeds

from test_gpu_parity import same, say

pytestmark = pytest.mark.gpu
F32 = np.float32


def _scene(key, w=8, h=8):
    kind, seed = key
    if kind == "soup":
        return scenes.soup(seed, seed=seed, width=w, height=h)
    if kind == "ico":
        return scenes.icosphere_in_box(4, seed=seed, width=w, height=h)
    return scenes.cornell(w, h)


def _camera(key, w, h):
    """The scene generator's camera for a w x h frame (shapes do not depend on the frame size)."""
    cam = _scene(key, w, h)[2]
    pt.update_camera(cam)
    return cam


def _alone(ctx, shapes, lc, cam, w, h, spp, seeds):
    s = shapes.copy()
    bvh = pt.CLBVH(ctx, s)
    bvh.buildRadixTree()
    bvh.buildBVHTree()
    d_col = ctx.alloc(w * h * 12, F32).zero()
    d_seed = ctx.to_device(seeds)
    d_pix = ctx.alloc(w * h * 4, np.uint32)
    pt.render(ctx, bvh, None, colors=d_col, seeds=d_seed, camera=ctx.to_device(cam), shapes=ctx.to_device(s),
              shapeCnt=len(s), lightCnt=lc, width=w, height=h, currentSample=0, spp=spp, max_depth=5, pixels=d_pix)
    out = d_col.download(F32), d_pix.download(np.uint32), d_seed.download(np.uint32)
    bvh.close()
    return out


def _serve(ctx, reqs, budget, spp_per_tick):
    srv = pt.Server(ctx, budget, spp_per_tick)
    ids, results, ticks = {}, {}, 0
    for k, (key, w, h, spp, seeds) in enumerate(reqs):
        shapes, lc, _ = _scene(key)
        key_id = hash(key) & 0xFFFFFFFF
        ids[srv.submit(key_id, shapes, lc, _camera(key, w, h), w, h, spp, 5, seeds)] = k
    while len(results) < len(reqs):
        active, queued = srv.tick()
        ticks += 1
        for rid in srv.finished():
            results[ids[rid]] = srv.result(rid, colors=True, pixels=True, seeds=True)
        assert ticks < 10000
    st = srv.stats()
    srv.close()
    return results, st, ticks


def _requests(rng, keys, n):
    out = []
    for _ in range(n):
        key = keys[rng.integers(len(keys))]
        w, h = [(24, 24), (48, 32), (64, 64), (16, 40)][rng.integers(4)]
        spp = int([1, 3, 5][rng.integers(3)])
        seeds = default_seeds(w, h) + np.uint32(rng.integers(0, 1000)) if rng.random() < 0.5 else None
        out.append((key, w, h, spp, seeds))
    return out


def test_served_equals_alone(ctx):
    say("case serve mix vs alone")
    rng = np.random.default_rng(21)
    keys = [("soup", 3000), ("soup", 20000), ("ico", 7), ("cornell", 0)]
    reqs = _requests(rng, keys, 12)
    results, st, ticks = _serve(ctx, reqs, budget=4 << 30, spp_per_tick=2)
    say(f"  ticks {ticks}, builds {st['builds']}, rays {st['closest'] + st['anyhit']}")
    assert st["builds"] == len({r[0] for r in reqs})       # one build per scene: the rest hit the cache
    for k, (key, w, h, spp, seeds) in enumerate(reqs):
        shapes, lc, _ = _scene(key)
        sd = seeds if seeds is not None else default_seeds(w, h)
        a_col, a_pix, a_seed = _alone(ctx, shapes, lc, _camera(key, w, h), w, h, spp, sd)
        c, p, s = results[k]
        assert same(c, a_col) and same(p, a_pix) and same(s, a_seed), k


def test_served_small_budget_queues_and_evicts(ctx):
    say("case serve small budget")
    rng = np.random.default_rng(5)
    keys = [("soup", 20000), ("soup", 20001), ("ico", 8)]
    reqs = _requests(rng, keys, 8)
    # room for one scene and a few frames at a time: requests queue, idle scenes are evicted
    budget = 20002 * 600 + 6 * (64 * 64 * 24 + 1024)
    results, st, ticks = _serve(ctx, reqs, budget=budget, spp_per_tick=1)
    switches = sum(1 for i in range(len(reqs)) if i == 0 or reqs[i][0] != reqs[i - 1][0])
    say(f"  ticks {ticks}, builds {st['builds']} (scene switches {switches})")
    assert st["builds"] >= len({r[0] for r in reqs})
    for k in (0, len(reqs) - 1):
        key, w, h, spp, seeds = reqs[k]
        shapes, lc, _ = _scene(key)
        sd = seeds if seeds is not None else default_seeds(w, h)
        a_col, _, _ = _alone(ctx, shapes, lc, _camera(key, w, h), w, h, spp, sd)
        assert same(results[k][0], a_col)


def test_server_rejects_oversized(ctx):
    srv = pt.Server(ctx, 2 << 20, 1)
    shapes, lc, cam = scenes.soup(20000, seed=1, width=8, height=8)
    pt.update_camera(cam)
    with pytest.raises(pt.PtgpuError):
        srv.submit(1, shapes, lc, cam, 8, 8, 1)
    srv.close()


def _frame_bytes(w, h):
    """The server's per-request frame footprint: colors, seeds and pixels (each 256-byte aligned) + camera."""
    px = w * h
    r = lambda b: (b + 255) & ~255  # noqa: E731
    return r(px * 12) + r(px * 8) + r(px * 4) + 256


def _run_until_done(srv, rid, limit=10000):
    for _ in range(limit):
        srv.tick()
        if rid in srv.finished():
            return
    raise AssertionError(f"request {rid} did not finish")


def test_served_own_idle_scene_survives_eviction(ctx):
    """Admission evicts idle scenes to fit a new frame.  When the only idle scene is the request's own
    (its earlier requests taken) and the new frame is larger than every pooled buffer, the scene must
    not be freed under the request: the server either keeps it or brings it in anew."""
    say("case serve own idle scene under eviction")
    key = ("soup", 20000)
    shapes, lc, _ = _scene(key)
    # the scene's built footprint, measured on a server with room to spare
    srv = pt.Server(ctx, 4 << 30, 1)
    rid = srv.submit(1, shapes, lc, _camera(key, 8, 8), 8, 8, 1)
    _run_until_done(srv, rid)
    real = srv.stats()["hbm_used"] - _frame_bytes(8, 8)
    srv.result(rid)
    srv.close()
    estimate = len(shapes) * 600
    assert real > estimate, "the case needs a built scene larger than the admission estimate"
    # room for the scene at its estimate and the large frame, not for the built scene and the large frame
    budget = estimate + _frame_bytes(64, 64)
    srv = pt.Server(ctx, budget, 1)
    a = srv.submit(1, shapes, lc, _camera(key, 8, 8), 8, 8, 1)
    _run_until_done(srv, a)
    srv.result(a)                              # the 8x8 frame buffer goes to the pool; the scene is idle
    b = srv.submit(1, shapes, lc, _camera(key, 64, 64), 64, 64, 2)
    _run_until_done(srv, b)
    col, pix, _ = srv.result(b, colors=True, pixels=True)
    st = srv.stats()
    srv.close()
    assert st["builds"] == 2                   # the scene was dropped and built anew, not used after free
    a_col, a_pix, _ = _alone(ctx, shapes, lc, _camera(key, 64, 64), 64, 64, 2, default_seeds(64, 64))
    assert same(col, a_col) and same(pix, a_pix)
