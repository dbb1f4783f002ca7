"""Scene ingest (ptgpu_scene_read) and PPM output (ptgpu_write_ppm): host code, no GPU needed.

The library's reader must agree bit for bit with the oracle restatement of ReadTxt on the
reference's own scene (scenes/obj-model.txt, fixture copied by tests/golden/make_reference_scene.py),
and both are pinned by hand: 326 shapes (2+2+2+12+2+2+60+240 model triangles, 2 triangles,
2 spheres), 2 lights, exact plane vertices and areas, sphere materials SPEC / REFR.
"""
from pathlib import Path

import numpy as np
import pytest

import ptgpu_pc_amd as pt
import scene_txt as S
from ptgpu_pc_amd.layout import CAMERA_DTYPE, SHAPE_DTYPE

SCENE_DIR = Path(__file__).resolve().parent / "golden" / "reference_scene"
SCENE = SCENE_DIR / "scenes" / "obj-model.txt"
F32 = np.float32


def test_reference_scene_library_equals_oracle():
    shapes, lc, cam = pt.read_scene(SCENE, SCENE_DIR, 640, 480)
    o_shapes, o_lc, o_cam = S.read_txt(SCENE, SCENE_DIR, SHAPE_DTYPE, CAMERA_DTYPE, 640, 480)
    assert len(shapes) == 326 and lc == 2
    assert shapes.view(np.uint8).tobytes() == o_shapes.view(np.uint8).tobytes()
    assert lc == o_lc
    assert cam.view(np.uint8).tobytes() == o_cam.view(np.uint8).tobytes()


def test_reference_scene_hand_checks():
    shapes, lc, cam = pt.read_scene(SCENE, SCENE_DIR, 320, 240)
    assert cam["orig"][0].tolist() == [0.0, 45.0, F32(205.6)]
    assert cam["target"][0].tolist() == [0.0, F32(44.957388), F32(204.6)]
    assert cam["width"][0] == 320 and cam["height"][0] == 240 and cam["yaw"][0] == 0 and cam["pitch"][0] == 0
    # first object: plane-x.obj scaled (-1, 70, 60) and moved to (90, 69, 0); its face "f 2 4 3"
    # takes vertices 2 = (-1,-1,-1), 4 = (-1,1,-1), 3 = (-1,1,1)
    assert shapes["type"][0] == S.TRIANGLE
    assert shapes["p1"][0].tolist() == [91.0, -1.0, -60.0]
    assert shapes["p2"][0].tolist() == [91.0, 139.0, -60.0]
    assert shapes["p3"][0].tolist() == [91.0, 139.0, 60.0]
    assert shapes["area"][0] == 8400.0            # half of the 140 x 120 rectangle
    assert shapes["c"][0].tolist() == [0.25, 0.25, 0.75]
    # per-object shape counts in file order
    counts = [2, 2, 2, 12, 2, 2, 60, 240, 1, 1, 1, 1]
    assert sum(counts) == len(shapes)
    starts = np.cumsum([0] + counts)
    # the two emissive triangles (20, 20, 20), then the SPEC and the REFR sphere of radius 5
    for k in (starts[8], starts[9]):
        assert shapes["type"][k] == S.TRIANGLE and shapes["e"][k].tolist() == [20.0, 20.0, 20.0]
    s1, s2 = starts[10], starts[11]
    assert shapes["type"][s1] == S.SPHERE and shapes["refl"][s1] == S.SPEC
    assert shapes["type"][s2] == S.SPHERE and shapes["refl"][s2] == S.REFR
    assert shapes["p"][s1].tolist() == [-24.0, 5.0, 44.0] and shapes["rad"][s2] == 5.0
    assert shapes["area"][s1] == F32(F32(F32(4) * F32(np.pi)) * F32(5)) * F32(5)
    assert (shapes["e"][:starts[8]] == 0).all()      # the models do not emit
    assert int((np.abs(shapes["e"]).sum(axis=1) > 0).sum()) == lc == 2


def test_obj_number_parsing():
    """The bundled OBJ reader's digit-by-digit double arithmetic, then one rounding to float."""
    assert S.obj_float("1.000000") == F32(1.0) and S.obj_float("-1.000000") == F32(-1.0)
    assert S.obj_float("0.5") == F32(0.5) and S.obj_float("+2.25") == F32(2.25)
    assert S.obj_float("12") == F32(12.0)
    # 0.3 accumulates as 0 + 3 * 0.1 = 0.30000000000000004 in double; float32 of it is 0.3f
    assert S.obj_float("0.3") == F32(0.30000000000000004)


def test_scene_errors(tmp_path):
    bad = tmp_path / "bad.txt"
    bad.write_text("camera 0 0 0  0 0 -1\nmaterials 0\nobjects 1\ncone 1 2 3\n")
    with pytest.raises(pt.PtgpuError):
        pt.read_scene(bad)
    missing = tmp_path / "missing.txt"
    missing.write_text("camera 0 0 0  0 0 -1\nmaterials 0\nobjects 1\nmodel models/none.obj 0 0 0 1 1 1 0 0 0 1 1 1 0\n")
    with pytest.raises(pt.PtgpuError):
        pt.read_scene(missing, tmp_path)
    with pytest.raises(pt.PtgpuError):
        pt.read_scene(tmp_path / "absent.txt")


def test_ppm_writer(tmp_path):
    """Header, rows from the top of the image (last buffer row) down, 'r g b ' from the low bytes."""
    w, h = 3, 2
    px = np.array([0xff030201, 0xff060504, 0xff090807,      # buffer row 0 (bottom of the image)
                   0xff0c0b0a, 0xff0f0e0d, 0xff807f10], dtype=np.uint32)
    out = tmp_path / "image.ppm"
    pt.write_ppm(out, px, w, h)
    want = "P3\n3 2\n255\n" + "10 11 12 13 14 15 16 127 128 " + "1 2 3 4 5 6 7 8 9 "
    assert out.read_text() == want
    with pytest.raises(ValueError):
        pt.write_ppm(out, px[:5], w, h)
