"""GPU parity tests: libptgpu.so (gfx950) against the CPU oracle on the same seeded inputs.

Bars (from the path's parity contract):
  Morton codes, sorted order, tree topology and bounds: bit-exact.
  Closest hit: >= 99.99% of rays with the same shape id, t within 1e-5 relative.
  Images: RMSE <= 1e-3 on clamped linear colours at matched seeds and spp.
"""
import sys
import time

import numpy as np
import pytest

import ptgpu_pc_amd as pt
import pyoracle as O
from ptgpu_pc_amd import scenes
from ptgpu_pc_amd.layout import RAY_DTYPE, SHAPE_DTYPE, SPHERE, default_seeds

pytestmark = pytest.mark.gpu
F32 = np.float32


_T0 = time.perf_counter()


def say(*a):
    print(f"[{time.perf_counter() - _T0:7.1f}s]", *a, file=sys.stderr, flush=True)


def same(a, b):
    """Bitwise equality of two arrays, as a bool: a failing assert then reports array summaries instead
    of diffing megabytes of bytes."""
    return a.tobytes() == b.tobytes()


def gpu_lbvh(ctx, shapes, digit_bits=8):
    s = shapes.copy()
    bvh = pt.CLBVH(ctx, s, digit_bits=digit_bits)
    bvh.buildRadixTree()
    bvh.buildBVHTree()
    return s, bvh


def small_cases():
    out = []
    for n in (1, 2, 3, 5, 9, 64, 1000, 4095, 4096, 4097):
        sh, _, _ = scenes.soup(n, seed=100 + n)
        out.append((f"soup{n}", sh))
    sph = np.zeros(7, dtype=SHAPE_DTYPE)
    sph["type"] = SPHERE
    sph["rad"] = 0.5
    sph["p"] = [(0, 0, 0), (1, 0, 0), (1, 0, 0), (-3, -2, -1), (-3, -2, -1), (0, 5, 0), (0, 5, 0)]
    out.append(("spheres_dupes", sph))
    out.append(("cornell", scenes.cornell()[0]))
    return out


@pytest.mark.parametrize("case", range(12))
def test_morton_sort_topology_bounds_bitexact_small(ctx, case):
    name, shapes = small_cases()[case]
    say("case", name, len(shapes))
    ref_s, ref_sorted, ref_btn, ref_btl = O.build_lbvh(shapes, literal_sort=len(shapes) <= 4096)
    for bits in (8, 4):
        s, bvh = gpu_lbvh(ctx, shapes, bits)
        btn, btl = bvh.getTrees()
        np.testing.assert_array_equal(s.view(np.uint8), ref_s.view(np.uint8))   # index, morton_code, b
        np.testing.assert_array_equal(bvh.sorted_indices(), ref_sorted)
        np.testing.assert_array_equal(btn.view(np.uint8), ref_btn.view(np.uint8))
        np.testing.assert_array_equal(btl.view(np.uint8), ref_btl.view(np.uint8))
        bvh.close()


@pytest.mark.parametrize("which", ["c2_soup_65538", "icosphere_l6_box", "soup_300k"])
def test_lbvh_bitexact_config_sizes(ctx, which):
    say("case", which)
    if which == "c2_soup_65538":
        shapes = scenes.soup(65536, seed=1)[0]
    elif which == "icosphere_l6_box":
        shapes = scenes.icosphere_in_box(6, seed=3)[0]
    else:
        shapes = scenes.soup(300000, seed=5)[0]
    ref_s, ref_sorted, ref_btn, ref_btl = O.build_lbvh(shapes)
    s, bvh = gpu_lbvh(ctx, shapes)
    btn, btl = bvh.getTrees()
    assert (s["morton_code"] == ref_s["morton_code"]).all()
    np.testing.assert_array_equal(bvh.sorted_indices(), ref_sorted)
    np.testing.assert_array_equal(btn.view(np.uint8), ref_btn.view(np.uint8))
    np.testing.assert_array_equal(btl.view(np.uint8), ref_btl.view(np.uint8))
    # size-independent properties: keys descending along leaf order, parents contain children
    keys = s["morton_code"][bvh.sorted_indices()]
    assert (np.diff(keys.astype(np.int64)) <= 0).all()


@pytest.mark.parametrize("n", [1, 2, 3, 5000])
def test_radix_tree_only_export(ctx, n):
    """After buildRadixTree alone the arrays hold the reference's state at that point (CLBVH.cpp:83-88
    fills, then BVH.cl:127-208's topology): internal nodes zeroed but for the Karras fields, leaves
    0xFF but for nParent.  Derived from the oracle's full build by clearing what the refit adds."""
    say("case radix-only export", n)
    shapes = scenes.soup(max(n - 2, 1), seed=8)[0][:n].copy()
    ref_s, ref_sorted, ref_btn, ref_btl = O.build_lbvh(shapes, literal_sort=n <= 4096)
    s = shapes.copy()
    bvh = pt.CLBVH(ctx, s)
    bvh.buildRadixTree()
    btn, btl = bvh.getTrees()
    exp_btn = ref_btn.copy()
    if len(exp_btn):
        exp_btn["bound"] = 0
    exp_btl = ref_btl.copy()
    raw = exp_btl.view(np.uint8).reshape(len(exp_btl), 60)
    raw[:, 8:12] = 0xFF     # nShape
    raw[:, 28:52] = 0xFF    # bound
    if n > 1:
        np.testing.assert_array_equal(btn.view(np.uint8), exp_btn.view(np.uint8))
    np.testing.assert_array_equal(btl.view(np.uint8), exp_btl.view(np.uint8))
    bvh.buildBVHTree()
    btn, btl = bvh.getTrees()
    if n > 1:
        np.testing.assert_array_equal(btn.view(np.uint8), ref_btn.view(np.uint8))
    np.testing.assert_array_equal(btl.view(np.uint8), ref_btl.view(np.uint8))
    bvh.close()


def check_traversal_tree(bvh, s):
    """Invariants of the traversal tree: a binary tree over all leaves, each child box the exact
    union (select-merge) of its leaf range, children partition the parent's range."""
    refs, boxes, height = bvh.export_traversal()
    n = bvh.n
    srt = bvh.sorted_indices()
    lb = s["b"][srt]
    seen = np.zeros(n, np.int32)
    seen_int = np.zeros(max(n - 1, 1), np.int32)
    seen_int[0] = 1
    for i in range(n - 1):
        l, r, lo, hi = refs[i]
        llo, lhi = (~l, ~l) if l < 0 else (refs[l][2], refs[l][3])
        rlo, rhi = (~r, ~r) if r < 0 else (refs[r][2], refs[r][3])
        assert llo == lo and rhi == hi and lhi + 1 == rlo, (i, refs[i])
        for side, (a, b) in enumerate(((llo, lhi), (rlo, rhi))):
            seg = lb[a:b + 1]
            exp = [seg[:, 0].min(), seg[:, 1].max(), seg[:, 2].min(), seg[:, 3].max(), seg[:, 4].min(), seg[:, 5].max()]
            assert boxes[i, side].tolist() == [float(x) for x in exp], (i, side)
        for c in (l, r):
            if c < 0:
                seen[~c] += 1
            else:
                seen_int[c] += 1
    assert (seen == 1).all() and (seen_int == 1).all()
    assert 0 < height <= 64
    return height


@pytest.mark.parametrize("topology", ["morton", "reference"])
@pytest.mark.parametrize("n", [2, 3, 1000, 1025, 4097, 70000])
def test_traversal_tree_invariants(ctx, topology, n):
    shapes = scenes.soup(n, seed=40 + n)[0] if n > 2 else scenes.soup(n, seed=3)[0][:n]
    say("case traversal", topology, n)
    s = shapes.copy()
    bvh = pt.CLBVH(ctx, s, traversal=topology)
    bvh.buildRadixTree()
    bvh.buildBVHTree()
    h = check_traversal_tree(bvh, s)
    say("height", h)


def test_traversal_tree_with_duplicate_codes(ctx):
    # many shapes share a centroid (equal Morton codes): ties break on the leaf index
    sph = np.zeros(3000, dtype=SHAPE_DTYPE)
    sph["type"] = SPHERE
    sph["rad"] = 0.1
    rng = np.random.default_rng(2)
    sph["p"] = np.round(rng.random((3000, 3)) * 4) / 4
    say("case duplicates")
    s = sph.copy()
    bvh = pt.CLBVH(ctx, s)
    bvh.buildRadixTree()
    bvh.buildBVHTree()
    check_traversal_tree(bvh, s)


def test_rebuild_is_deterministic(ctx):
    shapes = scenes.soup(100000, seed=9)[0]
    say("case rebuild x3")
    s, bvh = gpu_lbvh(ctx, shapes)
    btn0, btl0 = bvh.getTrees()
    for _ in range(2):
        bvh.reset()
        bvh.buildRadixTree()
        bvh.buildBVHTree()
        btn, btl = bvh.getTrees()
        assert btn.tobytes() == btn0.tobytes() and btl.tobytes() == btl0.tobytes()


# ---------------------------------------------------------------------------------------------
def random_rays(n, seed, lo=-0.2, hi=1.2):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, dtype=RAY_DTYPE)
    r["o"] = (rng.random((n, 3)) * (hi - lo) + lo).astype(F32)
    d = rng.normal(size=(n, 3)).astype(F32)
    r["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(F32)
    return r


def gpu_trace(ctx, bvh, kd, shapes_dev, n_shapes, rays, maxt=None):
    d_rays = ctx.to_device(rays)
    if maxt is None:
        t = ctx.alloc(len(rays) * 4, F32)
        ids = ctx.alloc(len(rays) * 4, np.int32)
        pt.trace_closest(ctx, bvh, kd, shapes_dev, n_shapes, d_rays, len(rays), t, ids)
        return t.download(F32), ids.download(np.int32)
    d_m = ctx.to_device(np.asarray(maxt, F32))
    hit = ctx.alloc(len(rays) * 4, np.int32)
    pt.trace_any(ctx, bvh, kd, shapes_dev, n_shapes, d_rays, d_m, len(rays), hit)
    return hit.download(np.int32)


def assert_closest_parity(t_g, id_g, t_o, id_o):
    match = (id_g == id_o).mean()
    assert match >= 0.9999, f"id match {match}"
    both = (id_g == id_o) & (id_o >= 0)
    rel = np.abs(t_g[both] - t_o[both]) / np.maximum(np.abs(t_o[both]), 1e-30)
    assert (rel <= 1e-5).all(), rel.max()


@pytest.mark.parametrize("accel", ["bvh", "kd", "kd_sah", "none"])
def test_closest_and_anyhit_vs_bruteforce(ctx, accel):
    n = 20000 if accel != "none" else 3000
    shapes = scenes.soup(n, seed=77)[0]
    s, _ = O.clbvh_construct(shapes)
    rays = random_rays(4000, 3)
    say("case", accel, n)
    t_o, id_o = O.trace_closest(s, O.accel_none(), rays)
    d_sh = ctx.to_device(s)
    bvh = kd = None
    if accel == "bvh":
        bvh = pt.CLBVH(ctx, d_sh, len(s))
        bvh.buildRadixTree()
        bvh.buildBVHTree()
    elif accel == "kd":
        kd = pt.KDTree(ctx, s)
    elif accel == "kd_sah":
        kd = pt.KDTree(ctx, s, sah=True, max_depth=24, max_leaf=4)
    t_g, id_g = gpu_trace(ctx, bvh, kd, d_sh, len(s), rays)
    assert (id_o >= 0).sum() > 100
    assert_closest_parity(t_g, id_g, t_o, id_o)
    maxt = np.where(id_o >= 0, t_o * F32(1.5), F32(0.5)).astype(F32)
    h_o = O.trace_any(s, rays, maxt)
    h_g = gpu_trace(ctx, bvh, kd, d_sh, len(s), rays, maxt)
    assert (h_o == h_g).mean() >= 0.9999


def test_closest_hit_c2_camera_rays(ctx):
    shapes, lc, cam = scenes.soup(65536, seed=1, width=64, height=64)
    pt.update_camera(cam)
    s, bvh = gpu_lbvh(ctx, shapes)
    # camera rays of a 64x64 frame through the C2 soup
    rng = np.random.default_rng(0)
    ys, xs = np.mgrid[0:64, 0:64]
    kx = ((xs.ravel() + rng.random(4096) - 0.5) / 64 - 0.5).astype(F32)
    ky = ((ys.ravel() + rng.random(4096) - 0.5) / 64 - 0.5).astype(F32)
    d = cam["x"][0] * kx[:, None] + cam["y"][0] * ky[:, None] + cam["dir"][0]
    rays = np.zeros(4096, dtype=RAY_DTYPE)
    rays["o"] = cam["orig"][0] + F32(0.1) * d
    rays["d"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    say("case c2 camera rays")
    t_o, id_o = O.trace_closest(s, O.accel_none(), rays)
    t_g, id_g = gpu_trace(ctx, bvh, None, ctx.to_device(s), len(s), rays)
    assert (id_o >= 0).mean() > 0.4
    assert_closest_parity(t_g, id_g, t_o, id_o)


# ---------------------------------------------------------------------------------------------
def clamp01(c):
    return np.clip(c, 0.0, 1.0)


def render_gpu(ctx, shapes, lc, cam, w, h, spp, depth, accel, bvh=None, kd=None, d_shapes=None, first=0,
               colors=None, seeds=None):
    d_sh = d_shapes if d_shapes is not None else ctx.to_device(shapes)
    d_col = colors if colors is not None else ctx.alloc(w * h * 12, F32).zero()
    d_seed = seeds if seeds is not None else ctx.to_device(default_seeds(w, h))
    d_pix = ctx.alloc(w * h * 4, np.uint32)
    d_cam = ctx.to_device(cam)
    pt.render(ctx, bvh, kd, colors=d_col, seeds=d_seed, camera=d_cam, shapes=d_sh, shapeCnt=len(shapes),
              lightCnt=lc, width=w, height=h, currentSample=first, spp=spp, max_depth=depth, pixels=d_pix)
    return d_col.download(F32), d_seed.download(np.uint32), d_pix.download(np.uint32), d_col, d_seed


def render_oracle(shapes, lc, cam, w, h, spp, depth, acc):
    colors = np.zeros(w * h * 3, dtype=F32)
    seeds = default_seeds(w, h)
    pixels = np.zeros(w * h, dtype=np.uint32)
    O.radiance_frame(colors, seeds, cam, shapes, lc, acc, w, h, 0, spp, pixels, depth, sem=0)
    return colors, seeds, pixels


def image_checks(c_g, s_g, p_g, c_o, s_o, p_o, tol=1e-3):
    rmse = float(np.sqrt(np.mean((clamp01(c_g) - clamp01(c_o)) ** 2)))
    say(f"rmse {rmse:.3e}  colours bit-equal {np.mean(c_g == c_o):.5f}  seeds equal {np.mean(s_g == s_o):.5f}  "
        f"pixels equal {np.mean(p_g == p_o):.5f}")
    assert rmse <= tol, rmse
    # Beyond the RMSE bar: the kernel and the oracle's kernel semantics evaluate every FP32 expression
    # in the same order with IEEE division / sqrt and one shared sin / cos, so colours and seeds are
    # bit-equal; only the 8-bit pack's powf comes from two math libraries (within one step).
    assert same(c_g, c_o), f"colours bit-equal on {np.mean(c_g == c_o):.6f} only"
    assert same(s_g, s_o), f"seeds equal on {np.mean(s_g == s_o):.6f} only"
    dp = np.abs(p_g.view(np.uint8).astype(int) - p_o.view(np.uint8).astype(int))
    assert (dp <= 1).all()


def test_render_cornell_none_vs_oracle(ctx):
    shapes, lc, cam = scenes.cornell(64, 64)
    pt.update_camera(cam)
    say("case cornell 64x64x4 none")
    c_g, s_g, p_g, _, _ = render_gpu(ctx, shapes, lc, cam, 64, 64, 4, 5, None)
    c_o, s_o, p_o = render_oracle(shapes, lc, cam, 64, 64, 4, 5, O.accel_none())
    image_checks(c_g, s_g, p_g, c_o, s_o, p_o)
    assert c_g.mean() > 0.05


ENGINES = ["auto", "megakernel", "wavefront"]


@pytest.mark.parametrize("engine", ENGINES)
def test_render_soup_bvh_vs_oracle(ctx, engine):
    shapes, lc, cam = scenes.soup(3000, seed=4, width=48, height=48)
    pt.update_camera(cam)
    say("case soup3002 48x48x4 bvh", engine)
    s, bvh = gpu_lbvh(ctx, shapes)
    d_sh = ctx.to_device(s)
    # the handle's traversal layout was built from `s`; render with the same shapes
    ctx.set_engine(engine)
    try:
        c_g, s_g, p_g, _, _ = render_gpu(ctx, s, lc, cam, 48, 48, 4, 5, None, bvh=bvh, d_shapes=d_sh)
    finally:
        ctx.set_engine("auto")
    _, _, btn, btl = O.build_lbvh(shapes)
    c_o, s_o, p_o = render_oracle(s, lc, cam, 48, 48, 4, 5, O.accel_bvh(btn, btl))
    image_checks(c_g, s_g, p_g, c_o, s_o, p_o)


@pytest.mark.parametrize("engine", ENGINES)
def test_render_icosphere_kd_vs_oracle(ctx, engine):
    shapes, lc, cam = scenes.icosphere_in_box(3, seed=3, width=40, height=30)
    pt.update_camera(cam)
    say("case icosphere l3 40x30x4 kd", engine)
    kd = pt.KDTree(ctx, shapes)
    ctx.set_engine(engine)
    try:
        c_g, s_g, p_g, _, _ = render_gpu(ctx, shapes, lc, cam, 40, 30, 4, 6, None, kd=kd)
    finally:
        ctx.set_engine("auto")
    okd = O.KD(shapes)
    n5, box, kn = okd.flatten_dfs()
    c_o, s_o, p_o = render_oracle(shapes, lc, cam, 40, 30, 4, 6, O.accel_kd_dfs(n5, box, kn))
    image_checks(c_g, s_g, p_g, c_o, s_o, p_o)


@pytest.mark.parametrize("scene", ["soup_bvh", "cornell_bvh", "icosphere_kd", "soup_kd_sah"])
def test_wavefront_equals_megakernel(ctx, scene):
    """The two schedules run the same device functions in the same order: same bits."""
    say("case engines", scene)
    if scene == "soup_bvh":
        shapes, lc, cam = scenes.soup(20000, seed=12, width=96, height=64)
    elif scene == "cornell_bvh":
        shapes, lc, cam = scenes.cornell(64, 48)
    else:
        shapes, lc, cam = scenes.icosphere_in_box(4, seed=3, width=64, height=48)
    if scene == "soup_kd_sah":
        shapes, lc, cam = scenes.soup(20000, seed=13, width=64, height=64)
    pt.update_camera(cam)
    w, h = int(cam["width"][0]), int(cam["height"][0])
    s = shapes.copy()
    bvh = kd = None
    if scene.endswith("bvh"):
        bvh = pt.CLBVH(ctx, s)
        bvh.buildRadixTree()
        bvh.buildBVHTree()
    elif scene == "icosphere_kd":
        kd = pt.KDTree(ctx, s)
    else:
        kd = pt.KDTree(ctx, s, sah=True, max_depth=24, max_leaf=2)
    d_sh = ctx.to_device(s)
    out = {}
    for eng in ("megakernel", "wavefront", "lanes"):
        ctx.set_engine(eng)
        out[eng] = render_gpu(ctx, s, lc, cam, w, h, 6, 5, None, bvh=bvh, kd=kd, d_shapes=d_sh)[:3]
    # per-lane states with few lanes (several pixels per lane) and extreme regroup thresholds
    # (quad: traversing lanes at or below which four lanes share each query, binary16 nodes only)
    for regroup, bpc, wide, quad in ((1, 1, 1, 16), (64, 1, 0, 16), (7, 2, 1, 16), (8, 8, 0, 16), (3, 1, 2, 16),
                                     (64, 3, 2, 0), (64, 1, 2, 16), (1, 1, 2, 4), (16, 8, 2, 16)):
        ctx.set_tuning("regroup", regroup)
        ctx.set_tuning("blocks_per_cu", bpc)
        ctx.set_tuning("wide", wide)
        ctx.set_tuning("quad", quad)
        out[f"lanes{regroup}/{bpc}/{wide}/{quad}"] = render_gpu(ctx, s, lc, cam, w, h, 6, 5, None, bvh=bvh, kd=kd,
                                                              d_shapes=d_sh)[:3]
    ctx.set_tuning("regroup", 16)
    ctx.set_tuning("blocks_per_cu", 2)
    ctx.set_tuning("wide", 2)
    ctx.set_tuning("quad", 16)
    ctx.set_engine("auto")
    for eng, res in out.items():
        for a, b in zip(out["megakernel"], res):
            assert a.tobytes() == b.tobytes(), eng


def test_fused_spp_equals_separate_launches(ctx):
    shapes, lc, cam = scenes.soup(5000, seed=6, width=32, height=32)
    pt.update_camera(cam)
    s, bvh = gpu_lbvh(ctx, shapes)
    d_sh = ctx.to_device(s)
    say("case fused spp")
    c1, s1, p1, _, _ = render_gpu(ctx, s, lc, cam, 32, 32, 4, 5, None, bvh=bvh, d_shapes=d_sh)
    col = ctx.alloc(32 * 32 * 12, F32).zero()
    sd = ctx.to_device(default_seeds(32, 32))
    for k in range(4):
        c2, s2, p2, col, sd = render_gpu(ctx, s, lc, cam, 32, 32, 1, 5, None, bvh=bvh, d_shapes=d_sh, first=k,
                                         colors=col, seeds=sd)
    assert c1.tobytes() == c2.tobytes() and s1.tobytes() == s2.tobytes() and p1.tobytes() == p2.tobytes()


def test_radiance_gpu_reference_layout(ctx):
    """RadianceGPU argument list with btn/btl (BVH), kng/kn (KD) and brute force give the same image
    as the prepared-layout path."""
    ctx.set_max_depth(5)
    shapes, lc, cam = scenes.soup(2000, seed=8, width=32, height=32)
    pt.update_camera(cam)
    s, bvh = gpu_lbvh(ctx, shapes)
    d_sh = ctx.to_device(s)
    say("case radiance_gpu ref layout")
    c_ref, s_ref, p_ref, _, _ = render_gpu(ctx, s, lc, cam, 32, 32, 1, 5, None, bvh=bvh, d_shapes=d_sh)
    d_btn, d_btl = bvh.device_trees()
    col = ctx.alloc(32 * 32 * 12, F32).zero()
    sd = ctx.to_device(default_seeds(32, 32))
    pix = ctx.alloc(32 * 32 * 4, np.uint32)
    d_cam = ctx.to_device(cam)
    pt.radiance_gpu(ctx, pt.ACCEL_BVH, col, sd, d_cam, d_sh, len(s), lc, btn=d_btn, btl=d_btl, width=32, height=32,
                    currentSample=0, pixels=pix)
    assert col.download(F32).tobytes() == c_ref.tobytes()
    assert pix.download(np.uint32).tobytes() == p_ref.tobytes()
    # brute force: same closest hits -> same image
    col2 = ctx.alloc(32 * 32 * 12, F32).zero()
    sd2 = ctx.to_device(default_seeds(32, 32))
    pt.radiance_gpu(ctx, pt.ACCEL_NONE, col2, sd2, d_cam, d_sh, len(s), lc, width=32, height=32, currentSample=0,
                    pixels=pix)
    assert np.mean(col2.download(F32) == c_ref) > 0.999
    # KD heap (int16 export of a small tree) through the reference layout
    small = scenes.soup(300, seed=3, width=16, height=16)
    sh3, lc3, cam3 = small
    pt.update_camera(cam3)
    kd = pt.KDTree(ctx, sh3)
    kng, kn = kd.getTrees()
    d3 = ctx.to_device(sh3)
    c_kd, _, _, _, _ = render_gpu(ctx, sh3, lc3, cam3, 16, 16, 1, 5, None, kd=kd, d_shapes=d3)
    col3 = ctx.alloc(16 * 16 * 12, F32).zero()
    sd3 = ctx.to_device(default_seeds(16, 16))
    pix3 = ctx.alloc(16 * 16 * 4, np.uint32)
    pt.radiance_gpu(ctx, pt.ACCEL_KD, col3, sd3, ctx.to_device(cam3), d3, len(sh3), lc3, kng=ctx.to_device(kng),
                    kngCnt=len(kng), kn=ctx.to_device(kn), knCnt=len(kn), width=16, height=16, currentSample=0,
                    pixels=pix3)
    assert np.mean(col3.download(F32) == c_kd) > 0.999
    ctx.set_max_depth(6)


def test_render_c2_full_frame_properties(ctx):
    """C2 at full size (65,538 shapes, 512x512, 8 spp): oracle on a pixel subset, properties overall."""
    shapes, lc, cam = scenes.soup(65536, seed=1, width=512, height=512)
    pt.update_camera(cam)
    s, bvh = gpu_lbvh(ctx, shapes)
    d_sh = ctx.to_device(s)
    say("case c2 full frame")
    c_g, s_g, p_g, _, _ = render_gpu(ctx, s, lc, cam, 512, 512, 8, 5, None, bvh=bvh, d_shapes=d_sh)
    assert np.isfinite(c_g).all() and (c_g >= 0).all()
    rng = np.random.default_rng(1)
    pix = np.sort(rng.choice(512 * 512, 256, replace=False)).astype(np.int32)
    _, _, btn, btl = O.build_lbvh(shapes)
    colors = np.zeros(512 * 512 * 3, dtype=F32)
    seeds = default_seeds(512, 512)
    O.radiance_frame(colors, seeds, cam, s, lc, O.accel_bvh(btn, btl), 512, 512, 0, 8, None, 5, sem=0, pix_list=pix)
    x, y = pix % 512, pix // 512
    ci = (511 - y) * 512 + x
    cg = c_g.reshape(-1, 3)[ci]
    co = colors.reshape(-1, 3)[ci]
    rmse = float(np.sqrt(np.mean((clamp01(cg) - clamp01(co)) ** 2)))
    say(f"subset rmse {rmse:.3e}, seeds {np.mean(s_g.reshape(-1, 2)[pix] == seeds.reshape(-1, 2)[pix]):.4f}")
    assert rmse <= 1e-3
    assert same(cg, co) and same(s_g.reshape(-1, 2)[pix], seeds.reshape(-1, 2)[pix])
