"""The five workload configurations at their full sizes, on the GPU, against the oracle.

C1  Cornell spheres, no acceleration, 256x256 x 4 spp, depth 5: the whole frame against the oracle.
C3  Displaced icosphere (subdivision 7, 327,680 triangles) in a 12-triangle box: the LBVH bit-exact,
    treelet-7 restructuring invariants, closest / any hits of the LBVH, the mean-split KD tree and the
    SAH KD tree (depth 24) against brute force, and the 1024x768 x 16 spp depth-6 frame (every tree
    the same bits; the oracle on a pixel subset).
C4  64 scenes of 10^3..10^6 triangles: the batched build equals building each scene alone (three
    scenes against the oracle), one batched render of all 64 frames equals rendering each alone, and
    three frames against the oracle on pixel subsets.
C5  A 32-request slice of the served Poisson trace (256^2..1024^2, 1/4/16 spp): every served frame
    equals its request rendered alone, and two against the oracle on pixel subsets.

The oracle walks shadow rays through its BVH (shadow_accel) on the large scenes: the same answers as
the reference's brute-force IntersectP, in seconds instead of minutes.
"""
import numpy as np
import pytest

import ptgpu_pc_amd as pt
import pyoracle as O
from ptgpu_pc_amd import scenes
from ptgpu_pc_amd.layout import default_seeds

from test_gpu_parity import (assert_closest_parity, clamp01, gpu_lbvh, gpu_trace, image_checks, random_rays,
                             render_gpu, same, say)

pytestmark = pytest.mark.gpu
F32 = np.float32


def subset_vs_oracle(c_g, s_g, shapes, lc, cam, w, h, spp, depth, pix, acc, seeds0=None, tol=1e-3):
    """The oracle's frame on the listed pixel ids only; colours compared at the flipped row the
    kernel writes (colors[(H-1-y) W + x]), seeds at the pixel id."""
    colors = np.zeros(w * h * 3, dtype=F32)
    seeds = (default_seeds(w, h) if seeds0 is None else seeds0).copy()
    O.radiance_frame(colors, seeds, cam, shapes, lc, acc, w, h, 0, spp, None, depth, sem=0, pix_list=pix, threads=16)
    x, y = pix % w, pix // w
    ci = (h - 1 - y) * w + x
    cg, co = c_g.reshape(-1, 3)[ci], colors.reshape(-1, 3)[ci]
    rmse = float(np.sqrt(np.mean((clamp01(cg) - clamp01(co)) ** 2)))
    seq = float(np.mean(s_g.reshape(-1, 2)[pix] == seeds.reshape(-1, 2)[pix]))
    say(f"  subset of {len(pix)} pixels: rmse {rmse:.3e}, seeds equal {seq:.4f}")
    assert rmse <= tol, rmse
    assert same(cg, co) and seq == 1.0      # bit-equal: see image_checks


def sah_of_export(refs, boxes, ci=1.2, ct=1.0):
    """SAH cost of an exported traversal tree (Ci = 1.2 per internal node, Ct = 1.0 per leaf, surface
    areas in float64; clheader.h:14-15), vectorised over the nodes."""
    b = boxes.astype(np.float64)

    def area(x):
        dx, dy, dz = x[..., 1] - x[..., 0], x[..., 3] - x[..., 2], x[..., 5] - x[..., 4]
        return 2.0 * (dx * dy + dy * dz + dz * dx)
    u = np.stack([np.minimum(b[:, 0, 0], b[:, 1, 0]), np.maximum(b[:, 0, 1], b[:, 1, 1]),
                  np.minimum(b[:, 0, 2], b[:, 1, 2]), np.maximum(b[:, 0, 3], b[:, 1, 3]),
                  np.minimum(b[:, 0, 4], b[:, 1, 4]), np.maximum(b[:, 0, 5], b[:, 1, 5])], axis=1)
    leaf = np.asarray(refs)[:, :2] < 0
    return float(ci * area(u).sum() + ct * (area(b) * leaf).sum())


def pick_pixels(w, h, k, seed):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(w * h, k, replace=False)).astype(np.int32)


# ---------------------------------------------------------------------------------------------
# C1
# ---------------------------------------------------------------------------------------------
def test_c1_cornell_full_frame_vs_oracle(ctx):
    w = h = 256
    shapes, lc, cam = scenes.cornell(w, h)
    pt.update_camera(cam)
    say("case C1 cornell 256x256x4 none, full frame")
    c_g, s_g, p_g, _, _ = render_gpu(ctx, shapes, lc, cam, w, h, 4, 5, None)
    colors = np.zeros(w * h * 3, dtype=F32)
    seeds = default_seeds(w, h)
    pixels = np.zeros(w * h, dtype=np.uint32)
    O.radiance_frame(colors, seeds, cam, shapes, lc, O.accel_none(), w, h, 0, 4, pixels, 5, sem=0, threads=16)
    image_checks(c_g, s_g, p_g, colors, seeds, pixels)
    # the reference's per-sample launches (RadianceGPU, no acceleration) give the fused frame's bits
    col = ctx.alloc(w * h * 12, F32).zero()
    sd = ctx.to_device(default_seeds(w, h))
    pix = ctx.alloc(w * h * 4, np.uint32)
    d_cam, d_sh = ctx.to_device(cam), ctx.to_device(shapes)
    ctx.set_max_depth(5)
    try:
        for k in range(4):
            pt.radiance_gpu(ctx, pt.ACCEL_NONE, col, sd, d_cam, d_sh, len(shapes), lc, width=w, height=h,
                            currentSample=k, pixels=pix)
    finally:
        ctx.set_max_depth(6)
    assert same(col.download(F32), c_g) and same(pix.download(np.uint32), p_g)


# ---------------------------------------------------------------------------------------------
# C3
# ---------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c3():
    shapes, lc, cam = scenes.icosphere_in_box(7, seed=3, width=1024, height=768)
    pt.update_camera(cam)
    assert len(shapes) == 327692
    ref = O.build_lbvh(shapes)
    return shapes, lc, cam, ref


def leaves_below_root(refs):
    seen, stack = [], [0]
    while stack:
        v = stack.pop()
        for c in refs[v][:2]:
            if c < 0:
                seen.append(~int(c))
            else:
                stack.append(int(c))
    return seen


def test_c3_lbvh_bitexact_and_treelet7(ctx, c3):
    shapes, lc, cam, (ref_s, ref_sorted, ref_btn, ref_btl) = c3
    say("case C3 LBVH bit-exact, 327,692 shapes")
    s, bvh = gpu_lbvh(ctx, shapes)
    btn, btl = bvh.getTrees()
    assert same(s, ref_s)
    np.testing.assert_array_equal(bvh.sorted_indices(), ref_sorted)
    assert same(btn, ref_btn) and same(btl, ref_btl)
    say("case C3 treelet 7 invariants")
    d_sh = ctx.to_device(s)
    rays = random_rays(4000, 31, lo=-1.4, hi=1.4)
    t0, id0 = gpu_trace(ctx, bvh, None, d_sh, len(s), rays)
    fac = np.random.default_rng(1).choice(np.array([0.9, 1.5], F32), size=len(rays))
    maxt = np.where(id0 >= 0, t0 * fac, F32(4.0)).astype(F32)   # before and past the closest hit
    h0 = gpu_trace(ctx, bvh, None, d_sh, len(s), rays, maxt)
    refs0, boxes0, height0 = bvh.export_traversal()
    sah0 = sah_of_export(refs0, boxes0)
    bvh.optimize(7)
    refs1, boxes1, height1 = bvh.export_traversal()
    sah1 = sah_of_export(refs1, boxes1)
    say(f"  SAH {sah0:.6g} -> {sah1:.6g}, height {height0} -> {height1}")
    assert sah1 <= sah0 * (1 + 1e-6) and sah1 < sah0
    assert sorted(leaves_below_root(refs1)) == list(range(len(s)))
    t1, id1 = gpu_trace(ctx, bvh, None, d_sh, len(s), rays)
    assert same(t1, t0) and same(id1, id0)
    assert same(gpu_trace(ctx, bvh, None, d_sh, len(s), rays, maxt), h0)
    bvh.close()


def test_c3_closest_and_anyhit_every_tree_vs_bruteforce(ctx, c3):
    shapes, lc, cam, (s, ref_sorted, ref_btn, ref_btl) = c3
    rays = random_rays(2000, 32, lo=-1.4, hi=1.4)
    say("case C3 closest / any hit vs brute force (2,000 rays)")
    t_o, id_o = O.trace_closest(s, O.accel_none(), rays, threads=16)
    assert (id_o >= 0).mean() > 0.9      # rays start inside the closed box
    fac = np.random.default_rng(2).choice(np.array([0.75, 1.5], F32), size=len(rays))
    maxt = np.where(id_o >= 0, t_o * fac, F32(4.0)).astype(F32)   # before and past the closest hit
    h_o = O.trace_any_bvh(s, O.accel_bvh(ref_btn, ref_btl, shadow_accel=True), rays, maxt, threads=16)
    assert 0.3 < h_o.mean() < 0.7
    d_sh = ctx.to_device(s)
    bvh = pt.CLBVH(ctx, d_sh, len(s))
    bvh.buildRadixTree()
    bvh.buildBVHTree()
    kd = pt.KDTree(ctx, s)
    kd_sah = pt.KDTree(ctx, s, sah=True, max_depth=24, max_leaf=4)
    for name, b, k in (("lbvh", bvh, None), ("kd_mean_split", None, kd), ("kd_sah24", None, kd_sah)):
        say(" ", name)
        t_g, id_g = gpu_trace(ctx, b, k, d_sh, len(s), rays)
        assert_closest_parity(t_g, id_g, t_o, id_o)
        assert (gpu_trace(ctx, b, k, d_sh, len(s), rays, maxt) == h_o).mean() >= 0.9999
    bvh.close()


def test_c3_full_frame_every_tree(ctx, c3):
    """1024x768 x 16 spp, depth 6: LBVH + treelet 7, the mean-split KD tree and the SAH KD tree render
    the same bits; properties over the whole frame; the oracle on 64 pixels."""
    shapes, lc, cam, (ref_s, ref_sorted, ref_btn, ref_btl) = c3
    w, h = 1024, 768
    say("case C3 full frame 1024x768x16 depth 6, LBVH + treelet 7")
    s, bvh = gpu_lbvh(ctx, shapes)
    bvh.optimize(7)
    d_sh = ctx.to_device(s)
    c_g, s_g, p_g, _, _ = render_gpu(ctx, s, lc, cam, w, h, 16, 6, None, bvh=bvh, d_shapes=d_sh)
    assert np.isfinite(c_g).all() and (c_g >= 0).all()
    lit = (c_g.reshape(-1, 3) > 0).any(axis=1).mean()
    say(f"  lit fraction {lit:.3f}, mean {c_g.mean():.4f}")
    assert lit > 0.9                                       # a closed box with a light: nearly every path returns light
    assert (s_g != default_seeds(w, h)).all()              # every pixel drew its samples
    for name, kd in (("kd_mean_split", pt.KDTree(ctx, s)), ("kd_sah24", pt.KDTree(ctx, s, sah=True, max_depth=24,
                                                                                      max_leaf=4))):
        say(" ", name)
        c_k, s_k, p_k, _, _ = render_gpu(ctx, s, lc, cam, w, h, 16, 6, None, kd=kd, d_shapes=d_sh)
        assert same(c_k, c_g) and same(s_k, s_g) and same(p_k, p_g)
    subset_vs_oracle(c_g, s_g, s, lc, cam, w, h, 16, 6, pick_pixels(w, h, 64, 3),
                     O.accel_bvh(ref_btn, ref_btl, shadow_accel=True))
    bvh.close()


# ---------------------------------------------------------------------------------------------
# C4
# ---------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c4():
    W = scenes.batch_set(64)
    assert len(W) == 64
    return W


def test_c4_batched_build_equals_alone_and_oracle(ctx, c4):
    say(f"case C4 batched build of 64 scenes, {sum(len(w['shapes']) for w in c4):,} shapes")
    devs = [ctx.to_device(w["shapes"]) for w in c4]
    hs = pt.CLBVH.create_batch(ctx, devs)
    smallest = sorted(range(64), key=lambda k: len(c4[k]["shapes"]))[:2]
    largest = max(range(64), key=lambda k: len(c4[k]["shapes"]))
    for k, (w, d, hb) in enumerate(zip(c4, devs, hs)):
        if k % 16 == 0:
            say(f"  scene {k} alone")
        s_alone, alone = gpu_lbvh(ctx, w["shapes"])
        bt_b, bl_b = hb.getTrees()
        bt_a, bl_a = alone.getTrees()
        assert same(d.download(np.uint8), s_alone.view(np.uint8)), k
        assert (hb.sorted_indices() == alone.sorted_indices()).all(), k
        assert same(bt_b, bt_a) and same(bl_b, bl_a), k
        if k in smallest or k == largest:
            say(f"  scene {k}: {len(w['shapes'])} shapes vs oracle")
            ref_s, ref_sorted, ref_btn, ref_btl = O.build_lbvh(w["shapes"])
            assert same(s_alone, ref_s)
            assert (hb.sorted_indices() == ref_sorted).all()
            assert same(bt_b, ref_btn) and same(bl_b, ref_btl)
        alone.close()
    for hb in hs:
        hb.close()


def test_c4_render_batch_equals_alone_and_oracle(ctx, c4):
    say("case C4 one batched render of 64 frames (256^2..1024^2, 4 spp, depth 5) vs alone")
    devs = [ctx.to_device(w["shapes"]) for w in c4]
    hs = pt.CLBVH.create_batch(ctx, devs)
    jobs = []
    for w, d, hb in zip(c4, devs, hs):
        jobs.append(dict(bvh=hb, kd=None, colors=ctx.alloc(w["w"] * w["h"] * 12, F32).zero(),
                         seeds=ctx.to_device(default_seeds(w["w"], w["h"])), camera=ctx.to_device(w["cam"]),
                         shapes=d, shapeCnt=len(w["shapes"]), lightCnt=w["lc"], width=w["w"], height=w["h"],
                         currentSample=0, spp=4, max_depth=5,
                         pixels=ctx.alloc(w["w"] * w["h"] * 4, np.uint32).zero()))
    pt.render_batch(ctx, jobs)
    ctx.synchronize()
    out = [(j["colors"].download(F32), j["seeds"].download(np.uint32), j["pixels"].download(np.uint32)) for j in jobs]
    for k, (w, d, hb) in enumerate(zip(c4, devs, hs)):
        a = render_gpu(ctx, w["shapes"], w["lc"], w["cam"], w["w"], w["h"], 4, 5, None, bvh=hb, d_shapes=d)[:3]
        for x, y in zip(out[k], a):
            assert same(x, y), k
    shapes_now = [d.download(np.uint8).view(pt.SHAPE_DTYPE) for d in devs]   # index / Morton / bounds filled
    for k in (0, 31, 63):
        w = c4[k]
        say(f"  frame {k}: {len(w['shapes'])} shapes, {w['w']}x{w['h']} vs oracle")
        _, _, rbtn, rbtl = O.build_lbvh(w["shapes"])
        subset_vs_oracle(out[k][0], out[k][1], shapes_now[k], w["lc"], w["cam"], w["w"], w["h"], 4, 5,
                         pick_pixels(w["w"], w["h"], 256, k), O.accel_bvh(rbtn, rbtl, shadow_accel=True))
    for hb in hs:
        hb.close()


# ---------------------------------------------------------------------------------------------
# C5
# ---------------------------------------------------------------------------------------------
def test_c5_served_slice_equals_alone_and_oracle(ctx):
    pool, reqs = scenes.serve_trace(32, 500.0, 16)
    say(f"case C5 served slice: {len(reqs)} requests over {len({r['scene'] for r in reqs})} scenes")
    srv = pt.Server(ctx, 24 << 30, 4)
    cams, ids = [], {}
    for i, r in enumerate(reqs):
        cam = scenes.orbit_camera(r["yaw"], r["w"], r["h"])
        pt.update_camera(cam)
        cams.append(cam)
        ss, shapes, lc, _ = pool[r["scene"]]
        ids[srv.submit(ss, shapes, lc, cam, r["w"], r["h"], r["spp"], 5)] = i
    served = {}
    for tick in range(100000):
        if tick % 50 == 0:
            say(f"  tick {tick}: {len(served)} of {len(reqs)} served")
        srv.tick()
        for rid in srv.finished():
            served[ids[rid]] = srv.result(rid, colors=True, pixels=True, seeds=True)
        if len(served) == len(reqs):
            break
    st = srv.stats()
    srv.close()
    assert len(served) == len(reqs)
    say(f"  builds {st['builds']}, queries {st['closest'] + st['anyhit']:,}")
    handles = {}
    say("  rendering each request alone")
    for i, r in enumerate(reqs):
        ss, shapes, lc, _ = pool[r["scene"]]
        if r["scene"] not in handles:
            handles[r["scene"]] = gpu_lbvh(ctx, shapes)
        s, bvh = handles[r["scene"]]
        a_col, a_seed, a_pix = render_gpu(ctx, s, lc, cams[i], r["w"], r["h"], r["spp"], 5, None, bvh=bvh)[:3]
        col, pix, seeds = served[i]                     # ptgpu_server_result: colors, pixels, seeds
        assert same(col, a_col) and same(pix, a_pix) and same(seeds, a_seed), i
    checked = 0
    for i, r in enumerate(reqs):
        ss, shapes, lc, _ = pool[r["scene"]]
        if checked == 2 or len(shapes) > 200000:
            continue
        s = handles[r["scene"]][0]
        say(f"  request {i}: {len(s)} shapes, {r['w']}^2 x {r['spp']} spp vs oracle")
        _, _, rbtn, rbtl = O.build_lbvh(shapes)
        subset_vs_oracle(served[i][0], served[i][2], s, lc, cams[i], r["w"], r["h"], r["spp"], 5,
                         pick_pixels(r["w"], r["h"], 128, i), O.accel_bvh(rbtn, rbtl, shadow_accel=True))
        checked += 1
    assert checked == 2
    for s, bvh in handles.values():
        bvh.close()
