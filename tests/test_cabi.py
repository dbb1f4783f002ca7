"""CPU checks of the C ABI: the library loads, exports every symbol include/ptgpu.h declares, reports
the reference layout sizes, and rejects bad arguments without touching a device."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import ptgpu_pc_amd as pt

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "ptgpu.h").read_text()
    return sorted(set(re.findall(r"\b(ptgpu_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = pt.load_library()
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_layout_sizes_match_reference():
    lib = pt.load_library()
    v = [C.c_int() for _ in range(6)]
    lib.ptgpu_layout_sizes(*[C.byref(x) for x in v])
    assert [x.value for x in v] == [104, 76, 60, 36, 24, 24]


def test_null_arguments_are_rejected():
    lib = pt.load_library()
    assert lib.ptgpu_clbvh_create(None, None, 10, None) == 1
    assert lib.ptgpu_render(None, None, None, None) == 1
    assert lib.ptgpu_kdtree_build(None, None, 5, None) == 1
    assert b"null" in lib.ptgpu_last_error()
    assert lib.ptgpu_ctx_wait_event(None, None, 0) == 1
    h = C.c_void_p()
    shapes = pt.empty_shapes(4)
    assert lib.ptgpu_kdtree_build(None, shapes.ctypes.data, 0, C.byref(h)) == 1


def test_update_camera_rejects_zero_height():
    cam = pt.make_camera((0, 0, 0), width=0, height=0)
    with pytest.raises(pt.PtgpuError):
        pt.update_camera(cam)


def test_shapes_must_use_the_104_byte_layout():
    bad = np.zeros(3, dtype=[("type", "<i4"), ("x", "<f4")])
    with pytest.raises(ValueError):
        pt.KDTree(None, bad)
