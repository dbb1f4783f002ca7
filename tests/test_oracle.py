"""CPU tests that pin the oracle (oracle/oracle.cpp) before it is trusted as the parity checker.

Each check is independent of the oracle's own formula: hand-computed cases, closed forms, a
second implementation in numpy/Python, or the reference's own host path compiled from the
reference sources (oracle/_ref, when present).
"""
import ctypes as C
import math
import struct

import numpy as np
import pytest

import pyoracle as O
from ptgpu_pc_amd import scenes
from ptgpu_pc_amd.layout import (BVHNODE_DTYPE, DIFF, RAY_DTYPE, SHAPE_DTYPE, SPHERE, TRIANGLE, default_seeds,
                                 make_camera)

F32 = np.float32
FLT_MIN = np.finfo(np.float32).tiny


# ---------------------------------------------------------------------------------------------
# GetRandom (simplernd.h:35-49)
# ---------------------------------------------------------------------------------------------
def py_get_random(s0, s1):
    s0 = (36969 * (s0 & 0xFFFF) + (s0 >> 16)) & 0xFFFFFFFF
    s1 = (18000 * (s1 & 0xFFFF) + (s1 >> 16)) & 0xFFFFFFFF
    ires = ((s0 << 16) + s1) & 0xFFFFFFFF
    f = struct.unpack("<f", struct.pack("<I", (ires & 0x007FFFFF) | 0x40000000))[0]
    return F32((F32(f) - F32(2.0)) / F32(2.0)), s0, s1


def test_get_random_known_answers():
    # seeds (2, 2): s0 = 36969*2 = 73938, s1 = 18000*2 = 36000,
    # ires = (73938 << 16) + 36000 = 0x20D2_8CA0 -> mantissa 0x528CA0
    out, s0, s1 = O.get_random(2, 2, 1)
    assert (s0, s1) == (73938, 36000)
    mant = 0x528CA0
    expect = (1.0 + mant / 2 ** 23) * 2.0   # 2.xxx
    assert out[0] == F32((F32(expect) - 2) / 2)
    assert 0.0 <= out[0] < 1.0


@pytest.mark.parametrize("seed", [(2, 2), (7, 123456), (0xFFFFFFFF, 0xFFFFFFFF), (65535, 65536), (123, 4000000)])
def test_get_random_matches_python(seed):
    got, a, b = O.get_random(seed[0], seed[1], 200)
    s0, s1 = seed
    for k in range(200):
        v, s0, s1 = py_get_random(s0, s1)
        assert got[k] == v
    assert (a, b) == (s0, s1)


def test_get_random_matches_reference_host():
    fn = O.ref_fn("ref_get_random", 1, 6)
    if fn is None:
        pytest.skip("reference host path not built (no reference tree)")
    fn.restype = C.c_float
    for seed in [(2, 2), (99, 12345), (0xFFFFFFFF, 3)]:
        a, b = C.c_uint(seed[0]), C.c_uint(seed[1])
        got, _, _ = O.get_random(seed[0], seed[1], 64)
        for k in range(64):
            assert fn(C.byref(a), C.byref(b)) == got[k]


# ---------------------------------------------------------------------------------------------
# Morton codes (CLBVH.cpp:90-164)
# ---------------------------------------------------------------------------------------------
def spheres_at(points, r=0.1):
    s = np.zeros(len(points), dtype=SHAPE_DTYPE)
    s["type"] = SPHERE
    s["rad"] = r
    s["p"] = np.asarray(points, dtype=F32)
    return s


def interleave_py(a, b, c):
    code = 0
    for bit in range(10):
        code |= ((a >> bit) & 1) << (3 * bit + 2)
        code |= ((b >> bit) & 1) << (3 * bit + 1)
        code |= ((c >> bit) & 1) << (3 * bit)
    return code


def test_morton_wrap_at_max_and_degenerate_axes():
    # x in {0, 0.5, 1}; y and z constant -> 0/0 -> code bits 0; x = 1 -> 1024 -> wraps to 0
    s, scene = O.clbvh_construct(spheres_at([(0, .5, .5), (.5, .5, .5), (1, .5, .5)]))
    assert list(s["morton_code"]) == [0, 1 << 29, 0]
    assert list(s["index"]) == [0, 1, 2]
    assert scene.tolist() == [0.0, 1.0, 0.5, 0.5, 0.5, 0.5]


def test_morton_flt_min_max_initialisation():
    # all x negative: max_x stays FLT_MIN (CLBVH.cpp:92), so x = -1 quantises to 512, not 1024
    s, scene = O.clbvh_construct(spheres_at([(-2, 0, 0), (-1, 1, 1)]))
    assert scene[1] == F32(FLT_MIN)
    a1 = int(((F32(-1) - F32(-2)) / (F32(FLT_MIN) - F32(-2))) * F32(1024))
    assert a1 == 512
    assert s["morton_code"][1] == interleave_py(512, 0, 0)   # y, z of shape 1 sit at the max -> 1024 -> 0
    assert s["morton_code"][0] == 0


def test_morton_triangle_centroid_is_aabb_midpoint():
    s = np.zeros(3, dtype=SHAPE_DTYPE)
    s["type"] = TRIANGLE
    s[0]["p1"], s[0]["p2"], s[0]["p3"] = (0, 0, 0), (2, 0, 0), (0, 4, 0)   # AABB mid (1, 2, 0)
    s[1]["p1"], s[1]["p2"], s[1]["p3"] = (0, 0, 0), (0, 0, 0), (0, 0, 0)
    s[2]["p1"], s[2]["p2"], s[2]["p3"] = (4, 8, 1), (4, 8, 1), (4, 8, 1)
    out, scene = O.clbvh_construct(s)
    assert scene.tolist() == [0.0, 4.0, 0.0, 8.0, 0.0, 1.0]
    # (1-0)/4*1024 = 256, (2-0)/8*1024 = 256, z 0
    assert out["morton_code"][0] == interleave_py(256, 256, 0)
    assert out["b"][0].tolist() == [0, 2, 0, 4, 0, 0]


def numpy_morton(shapes):
    """Independent vectorised restatement of CLBVH.cpp:90-164 (numpy float32 is IEEE)."""
    tri = shapes["type"] == TRIANGLE
    v = np.stack([shapes["p1"], shapes["p2"], shapes["p3"]], 1)
    lo = np.minimum(np.minimum(v[:, 0], v[:, 1]), v[:, 2])
    hi = np.maximum(np.maximum(v[:, 0], v[:, 1]), v[:, 2])
    p = np.where(tri[:, None], (lo + hi) / F32(2), shapes["p"]).astype(F32)
    mn = np.minimum(p.min(0), F32(np.finfo(F32).max))
    mx = np.maximum(p.max(0), F32(FLT_MIN))
    with np.errstate(invalid="ignore", divide="ignore"):
        q = ((p - mn) / (mx - mn)) * F32(1024)
    q = np.where(np.isnan(q), 0, q).astype(np.int64) & 1023

    def expand(x):
        x = x.astype(np.uint64)
        out = np.zeros_like(x)
        for bit in range(10):
            out |= ((x >> np.uint64(bit)) & np.uint64(1)) << np.uint64(3 * bit)
        return out
    return ((expand(q[:, 0]) << np.uint64(2)) | (expand(q[:, 1]) << np.uint64(1)) | expand(q[:, 2])).astype(np.uint32)


def test_morton_matches_numpy_restatement_on_soup():
    shapes, _, _ = scenes.soup(5000, seed=11)
    out, _ = O.clbvh_construct(shapes)
    np.testing.assert_array_equal(out["morton_code"], numpy_morton(shapes))
    assert (out["morton_code"] < (1 << 30)).all()


def test_interleave_bit_positions():
    # CLBVH.cpp:158-163: x bit (9-j) -> bit (10-j)*3-1
    s, _ = O.clbvh_construct(spheres_at([(0, 0, 0), (1023.0 / 1024, 0, 0), (1, 1, 1)]))
    assert s["morton_code"][1] == interleave_py(1023, 0, 0)


# ---------------------------------------------------------------------------------------------
# sort (CLBVH.cpp:31-47)
# ---------------------------------------------------------------------------------------------
def test_bubblesort_is_stable_descending():
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 50, size=1500).astype(np.uint32)   # many duplicates
    b = O.bubblesort(keys)
    s = O.stable_sort_desc(keys)
    ref = np.lexsort((np.arange(len(keys)), -keys.astype(np.int64)))
    np.testing.assert_array_equal(b, ref)
    np.testing.assert_array_equal(s, ref)


def test_bubblesort_tiny():
    assert O.bubblesort(np.array([3, 1, 3, 2], np.uint32)).tolist() == [0, 2, 3, 1]
    assert O.bubblesort(np.array([7], np.uint32)).tolist() == [0]


# ---------------------------------------------------------------------------------------------
# Karras radix tree (BVH.cl:141-208)
# ---------------------------------------------------------------------------------------------
def check_topology_closed_form(n):
    btn, btl = O.radix_tree(n)
    left, right, lo, hi, pint, pleaf = O.topology_closed_form(n)
    m = n - 1
    assert (btn["nLeft"] == left[:m]).all()
    assert (btn["nRight"] == right[:m]).all()
    assert (btn["min"] == lo[:m]).all()
    assert (btn["max"] == hi[:m]).all()
    assert (btn["nParent"][1:] == pint[1:m]).all()
    assert btn["nParent"][0] == 0            # root keeps the memset value (CLBVH.cpp:84)
    assert (btl["nParent"] == pleaf).all()
    mixed = (btn["nLeft"] == btn["min"]) | (btn["nRight"] == btn["max"])
    assert (btn["leaf"] == np.where(mixed, 2, 0)).all()
    return btn, btl


def test_radix_tree_equals_closed_form_small_n():
    for n in range(2, 1200):
        check_topology_closed_form(n)


@pytest.mark.parametrize("n", [4097, 65537, 65538, 262146, 327692])
def test_radix_tree_equals_closed_form_config_sizes(n):
    check_topology_closed_form(n)


def test_radix_tree_fill_patterns_and_mixed_nodes():
    for n in (2, 3, 9, 64, 65, 100):
        btn, btl = O.radix_tree(n)
        raw_l = btl.view(np.uint8).reshape(n, 60)
        # leaves: only nParent written, everything else 0xFF (CLBVH.cpp:88)
        keep = np.ones(60, bool)
        keep[24:28] = False
        assert (raw_l[:, keep] == 0xFF).all()
        raw_n = btn.view(np.uint8).reshape(n - 1, 60)
        assert (raw_n[:, 28:60] == 0).all()      # bound/cost/area untouched by the topology kernel
        # exactly one node with one leaf child and one internal child iff n is odd
        lleaf = btn["nLeft"] == btn["min"]
        rleaf = btn["nRight"] == btn["max"]
        one = int(np.count_nonzero(lleaf ^ rleaf))
        assert one == (1 if n % 2 == 1 and n > 2 else 0), (n, one)


# ---------------------------------------------------------------------------------------------
# refit (BVH.cl:214-258, intended semantics)
# ---------------------------------------------------------------------------------------------
def test_refit_equals_bruteforce_union():
    shapes, _, _ = scenes.soup(3000, seed=21)
    s, srt, btn, btl = O.build_lbvh(shapes)
    b = s["b"][srt]                      # leaf bounds in leaf order
    for i in range(len(btn)):
        lo, hi = btn["min"][i], btn["max"][i]
        seg = b[lo:hi + 1]
        exp = [seg[:, 0].min(), seg[:, 1].max(), seg[:, 2].min(), seg[:, 3].max(), seg[:, 4].min(), seg[:, 5].max()]
        assert btn["bound"][i].tolist() == [float(x) for x in exp]
    assert (btl["nShape"] == srt).all()
    assert (btl["bound"] == b).all()


# ---------------------------------------------------------------------------------------------
# KD tree (KDTree.cpp)
# ---------------------------------------------------------------------------------------------
def test_kd_single_shape():
    kd = O.KD(spheres_at([(1, 2, 3)]))
    heap, kn = kd.export_heap()
    assert kd.stats()["max_depth"] == 0 and len(heap) == 2
    assert heap[1]["leaf"] == 1 and heap[1]["min"] == 0 and heap[1]["max"] == 1 and kn.tolist() == [0]
    assert heap[1]["nLeft"] == 2 and heap[1]["nRight"] == 3 and heap[1]["nParent"] == 0


def test_kd_mean_skips_first_shape_and_ties_go_right():
    # two spheres at x=0 and x=10: the mean skips s[0] -> 10/2 = 5; 0 <= 5 -> right, 10 -> left
    kd = O.KD(spheres_at([(0, 0, 0), (10, 0, 0)], r=0.5))
    heap, kn = kd.export_heap()
    assert heap[1]["leaf"] == 0
    assert heap[2]["leaf"] == 1 and kn[heap[2]["min"]:heap[2]["max"]].tolist() == [1]
    assert heap[3]["leaf"] == 1 and kn[heap[3]["min"]:heap[3]["max"]].tolist() == [0]


def test_kd_equal_midpoints_make_one_leaf():
    kd = O.KD(spheres_at([(1, 1, 1)] * 4))
    heap, kn = kd.export_heap()
    # mean = 3 * (p * 0.25) = 0.75 p < p -> every shape goes left -> one leaf holding all four
    assert heap[1]["leaf"] == 1 and kn.tolist() == [0, 1, 2, 3]


def test_kd_triangle_midpoint_is_p3_over_3():
    s = np.zeros(3, dtype=SHAPE_DTYPE)
    s["type"] = TRIANGLE
    # p3 decides: p3.x/3 = 0, 3, 30 ; p1/p2 are decoys that would reverse the split if averaged
    s["p1"] = [(100, 0, 0), (100, 0, 0), (-100, 0, 0)]
    s["p2"] = [(100, 0, 0), (100, 0, 0), (-100, 0, 0)]
    s["p3"] = [(0, 0, 0), (9, 0, 0), (90, 0, 0)]
    kd = O.KD(s)
    node5, box, kn = kd.flatten_dfs()
    # mean over s[1:], each * (1/3): 3/3 + 30/3 = 11 -> p3/3 <= 11: shapes 0, 1 right; shape 2 left
    assert node5[0].tolist()[0] == 0
    left = node5[0][1]
    assert node5[left][0] == 1 and kn[node5[left][3]:node5[left][3] + node5[left][4]].tolist() == [2]


def test_kd_library_builder_matches_oracle():
    import ptgpu_pc_amd as pt
    for seed, n in [(1, 1), (2, 2), (3, 17), (4, 500), (5, 4000)]:
        shapes, _, _ = scenes.soup(n, seed=seed)
        ok = O.KD(shapes)
        lk = pt.KDTree(None, shapes)
        a5, abox, akn = ok.flatten_dfs()
        b5, bbox, bkn = lk.export_dfs()
        np.testing.assert_array_equal(a5, b5)
        np.testing.assert_array_equal(abox.view(np.uint32), bbox.view(np.uint32))
        np.testing.assert_array_equal(akn, bkn)
        st = ok.stats()
        if st["max_depth"] <= 13:
            h1, k1 = ok.export_heap()
            h2, k2 = lk.getTrees()
            np.testing.assert_array_equal(h1.view(np.uint8), h2.view(np.uint8))
            np.testing.assert_array_equal(k1, k2)


def test_kd_heap_export_refuses_int16_overflow():
    import ptgpu_pc_amd as pt
    shapes, _, _ = scenes.soup(40000, seed=9)
    lk = pt.KDTree(None, shapes)
    with pytest.raises(pt.PtgpuError):
        lk.getTrees()


# ---------------------------------------------------------------------------------------------
# intersections (rendering_kernel.cl:26-192)
# ---------------------------------------------------------------------------------------------
def ray(o, d):
    r = np.zeros(1, dtype=RAY_DTYPE)
    r["o"][0] = o
    r["d"][0] = d
    return r


def tri_shape(p1, p2, p3):
    s = np.zeros(1, dtype=SHAPE_DTYPE)
    s["type"] = TRIANGLE
    s["p1"], s["p2"], s["p3"] = p1, p2, p3
    return s


def triangle_t(s, r, sem=0):
    buf = np.ascontiguousarray(s).view(np.uint8)[4:40].copy()
    return O.lib().or_triangle_intersect(buf.ctypes.data, r.ctypes.data, sem)


def test_triangle_closed_forms():
    s = tri_shape((0, 0, 0), (1, 0, 0), (0, 1, 0))
    assert triangle_t(s, ray((0.25, 0.25, 2), (0, 0, -1))) == F32(2.0)          # hit at t = 2
    assert triangle_t(s, ray((0.25, 0.25, -2), (0, 0, 1))) == F32(2.0)          # two-sided
    assert triangle_t(s, ray((0.25, 0.25, 2), (1, 0, 0))) == 0.0                # parallel: det == 0
    assert triangle_t(s, ray((0.9, 0.9, 2), (0, 0, -1))) == 0.0                 # u + v > 1
    assert triangle_t(s, ray((0.25, 0.25, 0.005), (0, 0, -1))) == 0.0           # t <= EPSILON
    assert triangle_t(s, ray((0.25, 0.25, 0.02), (0, 0, -1))) == F32(0.02)


def test_sphere_closed_forms():
    sph = np.zeros(1, dtype=SHAPE_DTYPE)
    sph["type"] = SPHERE
    sph["rad"] = 1.0
    sph["p"] = (0, 0, 0)
    buf = np.ascontiguousarray(sph).view(np.uint8)[4:20].copy()
    f = O.lib().or_sphere_intersect
    assert f(buf.ctypes.data, ray((0, 0, 5), (0, 0, -1)).ctypes.data) == F32(4.0)
    assert f(buf.ctypes.data, ray((0, 0, 0), (0, 0, -1)).ctypes.data) == F32(1.0)   # inside: b + det
    assert f(buf.ctypes.data, ray((0, 3, 5), (0, 0, -1)).ctypes.data) == 0.0


def test_intersections_match_reference_host_bitwise():
    ft = O.ref_fn("ref_triangle_intersect", 1, 6)
    fs = O.ref_fn("ref_sphere_intersect", 1, 6)
    if ft is None:
        pytest.skip("reference host path not built")
    ft.restype = fs.restype = C.c_float
    rng = np.random.default_rng(3)
    shapes, _, _ = scenes.soup(400, seed=4)
    rays = np.zeros(400, dtype=RAY_DTYPE)
    rays["o"] = rng.random((400, 3), dtype=F32) * F32(1.2) - F32(0.1)
    d = rng.normal(size=(400, 3)).astype(F32)
    rays["d"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    for i in range(400):
        buf = np.ascontiguousarray(shapes[i:i + 1]).view(np.uint8)[4:40].copy()
        for j in range(0, 400, 7):
            r = np.ascontiguousarray(rays[j:j + 1])
            a = O.lib().or_triangle_intersect(buf.ctypes.data, r.ctypes.data, 1)
            b = ft(buf.ctypes.data, r.ctypes.data)
            assert struct.pack("<f", a) == struct.pack("<f", b)
    sph = scenes.cornell()[0]
    for i in range(9):
        buf = np.ascontiguousarray(sph[i:i + 1]).view(np.uint8)[4:20].copy()
        for j in range(0, 400, 3):
            r = np.ascontiguousarray(rays[j:j + 1]).copy()
            r["o"] = r["o"] * F32(100)
            a = O.lib().or_sphere_intersect(buf.ctypes.data, r.ctypes.data)
            b = fs(buf.ctypes.data, r.ctypes.data)
            assert struct.pack("<f", a) == struct.pack("<f", b)


# ---------------------------------------------------------------------------------------------
# traversal restatements agree with brute force
# ---------------------------------------------------------------------------------------------
def random_rays(n, seed, lo=-0.2, hi=1.2):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, dtype=RAY_DTYPE)
    r["o"] = (rng.random((n, 3)) * (hi - lo) + lo).astype(F32)
    d = rng.normal(size=(n, 3)).astype(F32)
    r["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(F32)
    return r


def test_oracle_bvh_and_kd_traversal_equal_bruteforce():
    shapes, _, _ = scenes.soup(4000, seed=8)
    s, srt, btn, btl = O.build_lbvh(shapes)
    rays = random_rays(3000, 1)
    tb, ib = O.trace_closest(s, O.accel_none(), rays)
    tv, iv = O.trace_closest(s, O.accel_bvh(btn, btl), rays)
    np.testing.assert_array_equal(ib, iv)
    np.testing.assert_array_equal(tb, tv)
    kd = O.KD(shapes)
    n5, box, kn = kd.flatten_dfs()
    tk, ik = O.trace_closest(s, O.accel_kd_dfs(n5, box, kn), rays)
    np.testing.assert_array_equal(ib, ik)
    assert (ib >= 0).sum() > 50    # the sparse soup is hit by ~5% of random rays


def test_oracle_quirk_mode_matches_reference_bvh_intersect():
    fn = O.ref_fn("ref_intersect", 1, 6)
    if fn is None:
        pytest.skip("reference host path not built")
    shapes, _, _ = scenes.soup(2000, seed=12)     # even N: no mixed nodes, stack need 11 < 18
    s, srt, btn, btl = O.build_lbvh(shapes)
    rays = random_rays(500, 2)
    acc = O.accel_bvh(btn, btl, quirks=1)
    for i in range(500):
        r = np.ascontiguousarray(rays[i:i + 1])
        t1, id1 = C.c_float(), C.c_uint(0)
        h1 = O.lib().or_intersect(s.ctypes.data, len(s), C.byref(acc), r.ctypes.data, C.byref(t1), C.byref(id1), 1)
        t2, id2 = C.c_float(), C.c_uint(0)
        h2 = fn(s.ctypes.data, len(s), btn.ctypes.data, btl.ctypes.data, 0, 0, r.ctypes.data, C.byref(t2),
                C.byref(id2))
        assert h1 == h2
        if h1:
            assert (t1.value, id1.value) == (t2.value, id2.value)


# ---------------------------------------------------------------------------------------------
# radiance: the oracle in host-twin semantics equals the reference's own host path bit for bit
# ---------------------------------------------------------------------------------------------
def ref_render(accel, depth, shapes, light_cnt, cam, w, h, spp, btn=None, btl=None):
    fn = O.ref_fn("ref_render", accel, depth)
    colors = np.zeros(w * h * 3, dtype=F32)
    seeds = default_seeds(w, h)
    a = btn.ctypes.data if btn is not None else None
    b = btl.ctypes.data if btl is not None else None
    fn(colors.ctypes.data, seeds.ctypes.data, cam.ctypes.data, shapes.ctypes.data, len(shapes), light_cnt, a, b,
       0, 0, w, h, 0, spp, None, 0, 8)
    return colors, seeds


def oracle_render(acc, depth, shapes, light_cnt, cam, w, h, spp, sem):
    colors = np.zeros(w * h * 3, dtype=F32)
    seeds = default_seeds(w, h)
    pixels = np.zeros(w * h, dtype=np.uint32)
    O.radiance_frame(colors, seeds, cam, shapes, light_cnt, acc, w, h, 0, spp, pixels, depth, sem=sem)
    return colors, seeds, pixels


def test_radiance_host_twin_equals_reference_cornell():
    if O.ref_lib() is None:
        pytest.skip("reference host path not built")
    shapes, lc, cam = scenes.cornell(32, 32)
    O.update_camera(cam)
    c1, s1 = ref_render(0, 5, shapes, lc, cam, 32, 32, 3)
    c2, s2, _ = oracle_render(O.accel_none(), 5, shapes, lc, cam, 32, 32, 3, sem=1)
    np.testing.assert_array_equal(c1.view(np.uint32), c2.view(np.uint32))
    np.testing.assert_array_equal(s1, s2)


def test_radiance_host_twin_equals_reference_bvh_triangles():
    if O.ref_lib() is None:
        pytest.skip("reference host path not built")
    shapes, lc, cam = scenes.soup(1998, seed=2, width=24, height=24)   # N = 2000 (even)
    O.update_camera(cam)
    s, srt, btn, btl = O.build_lbvh(shapes)
    c1, s1 = ref_render(1, 5, s, lc, cam, 24, 24, 2, btn, btl)
    c2, s2, _ = oracle_render(O.accel_bvh(btn, btl, quirks=1), 5, s, lc, cam, 24, 24, 2, sem=1)
    np.testing.assert_array_equal(c1.view(np.uint32), c2.view(np.uint32))
    np.testing.assert_array_equal(s1, s2)


def test_radiance_analytic_cases():
    # a camera that sees nothing returns 0 everywhere
    shapes = spheres_at([(0, 0, -100)], r=1.0)
    shapes["e"] = (12, 12, 12)
    shapes["area"] = scenes.sphere_area(shapes["rad"])
    cam = make_camera((0, 0, 0), yaw=0.0, width=8, height=8)
    cam["yaw"] = math.pi   # looking +z, away from the sphere
    O.update_camera(cam)
    c, _, _ = oracle_render(O.accel_none(), 5, shapes, 1, cam, 8, 8, 2, sem=0)
    assert (c == 0).all()
    # first hit on an emitter: e * |dp| with specularBounce = 1 (rendering_kernel.cl:600-610)
    cam["yaw"] = 0.0
    O.update_camera(cam)
    shapes2 = spheres_at([(0, 0, -100)], r=50.0)
    shapes2["e"] = (2, 2, 2)
    shapes2["area"] = scenes.sphere_area(shapes2["rad"])
    c, _, _ = oracle_render(O.accel_none(), 5, shapes2, 1, cam, 8, 8, 1, sem=0)
    c = c.reshape(8, 8, 3)
    # |dp| = cos of the angle between ray and surface normal: 1 on the axis, < 1 off it
    assert (c >= 0).all() and (c <= 2.0).all() and (c[:, :, 0] == c[:, :, 1]).all()
    assert c[3:5, 3:5].min() > 1.85 and c.max() > 1.99 and c[0, 0, 0] < c[3, 3, 0]


def test_update_camera_library_matches_oracle():
    import ptgpu_pc_amd as pt
    for yaw, pitch, w, h in [(0.0, 0.0, 256, 256), (1.3, 0.2, 1024, 768), (4.0, -0.5, 512, 512)]:
        a = make_camera((1, 2, 3), yaw=yaw, pitch=pitch, width=w, height=h)
        b = a.copy()
        pt.update_camera(a)
        O.update_camera(b)
        assert a.tobytes() == b.tobytes()


# ---------------------------------------------------------------------------------------------
# The per-pixel wrapper: UpdateCamera, GenerateCameraRay, running mean, row flip, RGBA8 pack
# ---------------------------------------------------------------------------------------------
FOV = F32(F32(3.14159265358979323846) / F32(180.0)) * F32(45.0)   # displayfunc.cpp:742, camera.h:40


def _both_update_camera(yaw, pitch, w, h):
    import ptgpu_pc_amd as pt
    a = make_camera((1, 2, 3), yaw=yaw, pitch=pitch, width=w, height=h)
    b = a.copy()
    pt.update_camera(a)
    O.update_camera(b)
    return a, b


@pytest.mark.parametrize("yaw,pitch,w,h", [(0.0, 0.0, 256, 256), (math.pi / 2, 0.0, 1024, 768),
                                           (0.0, 0.3, 640, 480)])
def test_update_camera_closed_form(yaw, pitch, w, h):
    """displayfunc.cpp:732-756: dir = -(sin yaw cos pitch, sin pitch, cos yaw cos pitch);
    x = norm(dir x up) * W fov / H; y = norm(x x dir) * fov.  Closed forms per case:
      yaw 0, pitch 0:  dir (0, 0, -1), x (W fov/H, 0, 0), y (0, fov, 0)
      yaw 90 deg:      dir (-1, 0, 0), x (0, 0, -W fov/H), y (0, fov, 0)
      pitch p:         dir (0, -sin p, -cos p), x (W fov/H, 0, 0), y (0, cos p, -sin p) * fov"""
    k = float(FOV) * w / h
    if pitch == 0.0 and yaw == 0.0:
        want = [(0, 0, -1), (k, 0, 0), (0, float(FOV), 0)]
    elif pitch == 0.0:
        want = [(-1, 0, 0), (0, 0, -k), (0, float(FOV), 0)]
    else:
        sp, cp = math.sin(pitch), math.cos(pitch)
        want = [(0, -sp, -cp), (k, 0, 0), (0, cp * float(FOV), -sp * float(FOV))]
    for cam in _both_update_camera(yaw, pitch, w, h):
        for name, v in zip(("dir", "x", "y"), want):
            np.testing.assert_allclose(cam[name][0].astype(np.float64), v, rtol=0, atol=2e-6 * max(1.0, k), err_msg=name)


def test_generate_camera_ray_hand_computed():
    """rendering_kernel.cl:1042-1070 by hand in float32: r1 then r2 drawn from the pixel's seeds,
    kcx = (x + r1)/W - .5, kcy = (y + r2)/H - .5, d = X kcx + Y kcy + dir, o = orig + 0.1 d, then d
    normalised; the pixel is x = gid % W, y = gid / W."""
    w, h = 40, 30
    cam = make_camera((0.5, 0.25, 3.0), yaw=0.4, pitch=-0.15, width=w, height=h)
    O.update_camera(cam)
    rng = np.random.default_rng(11)
    seeds = rng.integers(2, 2 ** 32 - 1, size=2 * w * h, dtype=np.uint64).astype(np.uint32)
    pix = np.array([0, w - 1, w * (h - 1), w * h - 1, 17 * w + 3, 5 * w + 31], dtype=np.int32)
    got_seeds = seeds.copy()
    rays = O.camera_rays(cam, got_seeds, w, h, pix)
    X, Y, D, P = (cam[n][0].astype(F32) for n in ("x", "y", "dir", "orig"))
    for q, gid in enumerate(pix):
        s0, s1 = int(seeds[2 * gid]), int(seeds[2 * gid + 1])
        u1, s0, s1 = py_get_random(s0, s1)
        u2, s0, s1 = py_get_random(s0, s1)
        r1, r2 = F32(u1 - F32(0.5)), F32(u2 - F32(0.5))
        x, y = F32(gid % w), F32(gid // w)
        kcx = F32(F32(F32(x + r1) * F32(F32(1) / F32(w))) - F32(0.5))
        kcy = F32(F32(F32(y + r2) * F32(F32(1) / F32(h))) - F32(0.5))
        d = np.array([F32(F32(X[i] * kcx) + F32(Y[i] * kcy)) + D[i] for i in range(3)], dtype=F32)
        o = np.array([F32(F32(0.1) * d[i]) + P[i] for i in range(3)], dtype=F32)
        dd = F32(F32(d[0] * d[0]) + F32(d[1] * d[1])) + F32(d[2] * d[2])
        inv = F32(F32(1) / F32(np.sqrt(dd)))
        dn = np.array([inv * d[i] for i in range(3)], dtype=F32)
        assert rays["o"][q].tobytes() == o.tobytes(), q
        assert rays["d"][q].tobytes() == dn.tobytes(), q
        assert (int(got_seeds[2 * gid]), int(got_seeds[2 * gid + 1])) == (s0, s1)


def _emitter_scene(center, rad, e, floor=False):
    shapes = spheres_at([center], r=rad)
    shapes["e"] = e
    shapes["area"] = scenes.sphere_area(shapes["rad"])
    if floor:   # a large diffuse sphere below: radiance then varies from sample to sample
        f = spheres_at([(0.0, -1e4 - 2.0, -10.0)], r=1e4)
        f["c"] = (0.75, 0.75, 0.75)
        f["refl"] = DIFF
        f["area"] = scenes.sphere_area(f["rad"])
        shapes = _concat(shapes, f)
    return shapes


def _concat(a, b):
    from ptgpu_pc_amd.layout import concat_shapes
    return concat_shapes([a, b])


def test_row_flip_and_pixel_orientation():
    """RadianceGPU writes colors[(H-1-y) W + x] and pixels[y W + x] (rendering_kernel.cl:1127,1141):
    an emitter up and to the right of the view axis (+x, +y world; camera looking down -z) is seen at
    pixel ids with x > W/2 and y > H/2, so it lands in the upper-right block of `pixels` (rows by y)
    and in the first rows of `colors` (rows by H-1-y)."""
    w = h = 32
    shapes = _emitter_scene((2.0, 2.0, -10.0), 1.5, (4.0, 4.0, 4.0))
    cam = make_camera((0, 0, 0), yaw=0.0, width=w, height=h)
    O.update_camera(cam)
    c, _, p = oracle_render(O.accel_none(), 5, shapes, 1, cam, w, h, 2, sem=0)
    lit_c = (c.reshape(h, w, 3) != 0).any(axis=2)
    lit_p = p.reshape(h, w) != 0xFF000000
    assert lit_c.sum() > 20 and lit_c.sum() == lit_p.sum()
    assert not lit_c[h // 2:, :].any() and not lit_c[:, : w // 2].any()     # colors: rows H-1-y < H/2
    assert not lit_p[: h // 2, :].any() and not lit_p[:, : w // 2].any()     # pixels: rows y >= H/2
    np.testing.assert_array_equal(lit_c[::-1], lit_p)


def test_pack_known_values_and_channel_order():
    """toInt(v) = (int)(pow(clamp(v, 0, 1), 1/2.2) * 255 + .5) (vec.h:60) and the pack
    r | g << 8 | b << 16 | 0xff000000 (rendering_kernel.cl:1141-1145).  From inside an emissive sphere
    every camera ray meets the surface head-on (|dp| = 1 to within an ulp), so the colour is e:
    toInt(0.5) = 186, toInt(1) = 255, toInt(2) = 255 (clamped); a frame that sees nothing is 0 -> 0."""
    w = h = 8
    shapes = _emitter_scene((0.0, 0.0, 0.0), 10.0, (0.5, 1.0, 2.0))
    cam = make_camera((0, 0, 0), yaw=0.3, width=w, height=h)
    O.update_camera(cam)
    c, _, p = oracle_render(O.accel_none(), 5, shapes, 1, cam, w, h, 1, sem=0)
    np.testing.assert_allclose(c.reshape(-1, 3), np.tile([0.5, 1.0, 2.0], (w * h, 1)), rtol=1e-6)
    assert (p == np.uint32(186 | (255 << 8) | (255 << 16) | 0xFF000000)).all()
    cam["yaw"] = F32(math.pi)
    away = _emitter_scene((0.0, 0.0, -50.0), 1.0, (0.5, 1.0, 2.0))
    O.update_camera(cam)
    c, _, p = oracle_render(O.accel_none(), 5, away, 1, cam, w, h, 1, sem=0)
    assert (c == 0).all() and (p == np.uint32(0xFF000000)).all()


def test_running_mean_over_samples():
    """Sample k > 0 folds in as colors = (colors k + r) / (k + 1) (rendering_kernel.cl:1127-1135): after
    K launches a pixel holds the arithmetic mean of its K sample radiances.  Each r_k is rendered on
    its own (currentSample 0 assigns) from the seeds sample k starts from."""
    w, h = 24, 16
    shapes = _emitter_scene((2.0, 2.0, -10.0), 1.5, (4.0, 4.0, 4.0), floor=True)
    cam = make_camera((0, 0, 0), yaw=0.0, pitch=-0.2, width=w, height=h)
    O.update_camera(cam)
    seeds = default_seeds(w, h)
    samples = []
    for _ in range(4):
        col = np.zeros(w * h * 3, F32)
        O.radiance_frame(col, seeds, cam, shapes, 1, O.accel_none(), w, h, 0, 1, None, 5, sem=0)
        samples.append(col.astype(np.float64))
    col = np.zeros(w * h * 3, F32)
    s2 = default_seeds(w, h)
    for k in range(4):   # four launches, currentSample 0..3
        O.radiance_frame(col, s2, cam, shapes, 1, O.accel_none(), w, h, k, 1, None, 5, sem=0)
        mean = np.mean(samples[:k + 1], axis=0)
        np.testing.assert_allclose(col, mean, rtol=2e-6, atol=1e-7)
    np.testing.assert_array_equal(s2, seeds)
    assert np.std(np.array(samples), axis=0).max() > 0.01   # the samples differ


def test_shadow_rays_through_bvh_equal_bruteforce():
    """The oracle's BVH-walked IntersectP (shadow_accel) answers the reference's brute-force scan on
    every ray, and a frame rendered with it has the same bits."""
    shapes, lc, cam = scenes.soup(20000, seed=21, width=20, height=20)
    s, srt, btn, btl = O.build_lbvh(shapes)
    rays = random_rays(4000, 9, lo=0.2, hi=0.8)
    t, ids = O.trace_closest(s, O.accel_none(), rays)
    rng = np.random.default_rng(2)
    maxt = np.where(ids >= 0, t * F32(rng.choice([0.5, 1.0, 1.5], size=len(t))), F32(2.0)).astype(F32)
    h_b = O.trace_any(s, rays, maxt)
    h_v = O.trace_any_bvh(s, O.accel_bvh(btn, btl, shadow_accel=True), rays, maxt)
    np.testing.assert_array_equal(h_b, h_v)
    assert 100 < h_b.sum() < len(h_b) - 100
    O.update_camera(cam)
    a = oracle_render(O.accel_bvh(btn, btl), 5, s, lc, cam, 20, 20, 2, sem=0)
    b = oracle_render(O.accel_bvh(btn, btl, shadow_accel=True), 5, s, lc, cam, 20, 20, 2, sem=0)
    for x, y in zip(a, b):
        assert x.tobytes() == y.tobytes()


def test_sincos_kernel_semantics_against_the_mathematics():
    """The kernel semantics' sin/cos (Cody-Waite reduction and Taylor series, all in binary32) against
    float64 sin/cos rounded to float32: within one binary32 ulp everywhere on the sampling angles
    2 pi u (u in [0, 1)) and beyond, and equal to the correctly rounded value for most arguments.
    A dropped series term or a wrong reduction constant costs far more than one ulp."""
    rng = np.random.default_rng(4)
    u = rng.random(6000).astype(F32)
    xs = np.concatenate([F32(2.0) * F32(math.pi) * u,                       # the angles the kernel draws
                         rng.uniform(-20.0, 20.0, 3000).astype(F32),
                         np.array([0.0, -0.0, math.pi / 4, math.pi / 2, math.pi, 1.5 * math.pi, 2 * math.pi,
                                   1e-7, 3.0, 6.2831], dtype=F32)])
    exact_s, exact_c = 0, 0
    for x in xs:
        s, c = O.sincos(x, sem=0)
        rs, rc = np.float32(math.sin(float(x))), np.float32(math.cos(float(x)))
        for got, want in ((s, rs), (c, rc)):
            ulp = np.spacing(np.abs(want)) if want != 0 else np.float32(1e-45)
            # pi/2 split in three binary32 parts leaves a residue of 1.7e-15 per multiple removed (plus
            # the rounding of k P3): an absolute error below 3e-15 |x|, seen only near zeros of sin and cos
            assert abs(float(got) - float(want)) <= float(ulp) + 3e-15 * abs(float(x)), (x, got, want)
        exact_s += s == rs
        exact_c += c == rc
    assert exact_s >= 0.75 * len(xs) and exact_c >= 0.75 * len(xs)
    assert O.sincos(0.0) == (0.0, 1.0)
