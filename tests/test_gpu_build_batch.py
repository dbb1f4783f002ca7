"""Batched LBVH builds (ptgpu_clbvh_create_batch / ptgpu_clbvh_build_batch).

Every scene of a batch must come out bit-identical to building it alone (CLBVH + buildRadixTree +
buildBVHTree), which the single-scene parity tests pin to the oracle: Morton codes and AABBs
written into the shapes, sorted order, the exported btn/btl trees, and the traversal layout with
its height.  Sizes cover the stage boundaries: 1 and 2 shapes, the 1024-leaf refit chunk, the
4096-key sort tile, and several level-refit passes.
"""
import numpy as np
import pytest

import ptgpu_pc_amd as pt
import pyoracle as O
from ptgpu_pc_amd import scenes
from ptgpu_pc_amd.layout import SHAPE_DTYPE

from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def _mixed_shapes():
    out = [scenes.soup(1, seed=2)[0][:1], scenes.soup(1, seed=3)[0][:2], scenes.cornell()[0]]
    for n in (1022, 1023, 1024, 4094, 4095, 4096, 30000):
        out.append(scenes.soup(n, seed=n)[0])
    out.append(scenes.icosphere_in_box(5, seed=4)[0])
    rng = np.random.default_rng(17)
    for _ in range(5):   # log-uniform sizes, as in the many-scene workload
        n = int(np.exp(rng.uniform(np.log(1e3), np.log(3e5))))
        out.append(scenes.soup(n, seed=int(rng.integers(1 << 20)))[0])
    return out


def _alone(ctx, shapes, bits, topo):
    d = ctx.to_device(shapes)
    h = pt.CLBVH(ctx, d, digit_bits=bits, traversal=topo)
    h.buildRadixTree()
    h.buildBVHTree()
    return d, h


def _state(h, d):
    btn, btl = h.getTrees()
    refs, boxes, height = h.export_traversal()
    return dict(shapes=d.download(SHAPE_DTYPE).view(np.uint8), sorted=h.sorted_indices(), btn=btn.view(np.uint8),
                btl=btl.view(np.uint8), refs=refs, boxes=boxes, height=height)


def _assert_same(a, b, k):
    for key in a:
        assert np.array_equal(np.asarray(a[key]), np.asarray(b[key])), (k, key)


@pytest.mark.parametrize("bits,topo", [(8, "morton"), (4, "morton"), (8, "reference")])
def test_batch_build_equals_alone(ctx, bits, topo):
    shapes = _mixed_shapes()
    say(f"case batch build, {len(shapes)} scenes, {sum(len(s) for s in shapes)} shapes, {bits}-bit digits, {topo}")
    alone = []
    for s in shapes:
        d, h = _alone(ctx, s, bits, topo)
        alone.append(_state(h, d))
        h.close()
    d_batch = [ctx.to_device(s) for s in shapes]
    if bits == 8 and topo == "morton":
        hs = pt.CLBVH.create_batch(ctx, d_batch)
    else:
        hs = [pt.CLBVH(ctx, d, digit_bits=bits, traversal=topo) for d in d_batch]
        pt.CLBVH.build_batch(ctx, hs)
    for k, (h, d) in enumerate(zip(hs, d_batch)):
        _assert_same(alone[k], _state(h, d), k)
    # the oracle on two of the scenes, directly
    for k in (2, 7):
        ref_s, ref_sorted, ref_btn, ref_btl = O.build_lbvh(shapes[k])
        btn, btl = hs[k].getTrees()
        assert np.array_equal(hs[k].sorted_indices(), ref_sorted)
        assert btn.view(np.uint8).tobytes() == ref_btn.view(np.uint8).tobytes()
        assert btl.view(np.uint8).tobytes() == ref_btl.view(np.uint8).tobytes()
    for h in hs:
        h.close()


def test_batch_rebuild_after_moving_shapes(ctx):
    """build_batch on existing handles after their shapes moved equals fresh single builds."""
    shapes = [scenes.soup(n, seed=n)[0] for n in (500, 5000, 50000)]
    say("case batch rebuild")
    d = [ctx.to_device(s) for s in shapes]
    hs = pt.CLBVH.create_batch(ctx, d)
    moved = []
    for k, s in enumerate(shapes):
        m = s.copy()
        m["p1"] = m["p1"] * np.float32(1.5) + np.float32(0.25 * k)
        moved.append(m)
        d[k].upload(m)
    pt.CLBVH.build_batch(ctx, hs)
    for k, m in enumerate(moved):
        d1, h1 = _alone(ctx, m, 8, "morton")
        _assert_same(_state(h1, d1), _state(hs[k], d[k]), k)
        h1.close()
    for h in hs:
        h.close()


def test_batch_build_rejects_bad_batches(ctx):
    shapes = scenes.soup(200, seed=9)[0]
    a, b = ctx.to_device(shapes), ctx.to_device(shapes)
    ha = pt.CLBVH(ctx, a)
    hb = pt.CLBVH(ctx, b, digit_bits=4)
    with pytest.raises(pt.PtgpuError):
        pt.CLBVH.build_batch(ctx, [ha, ha])     # the same handle twice
    with pytest.raises(pt.PtgpuError):
        pt.CLBVH.build_batch(ctx, [ha, hb])     # mixed digit widths
    pt.CLBVH.build_batch(ctx, [])               # nothing to do
    assert pt.CLBVH.create_batch(ctx, []) == []
    ha.close()
    hb.close()
