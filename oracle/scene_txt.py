"""Oracle (test infrastructure only): the reference's text-scene reader restated in Python.

Follows ReadTxt (displayfunc.cpp:144-505) and the number parsing of the OBJ reader the reference
bundles (tinyobj_loader_c.h:267-402, tryParseDouble / parseFloat).  Only the tests import this file;
the library's own reader (ptgpu_scene_read, csrc/scene.cpp) is what they check against it.

Float semantics: scene-file numbers go through the C library's strtof (what fscanf("%f") does);
OBJ numbers through the bundled reader's double arithmetic (integer digits m = 10 m + d, fractional
digits m += d * 0.1^k with 0.1^k a product of k factors 0.1), rounded to float; vertex scale and
translation, triangle areas (|e1 x e2| / 2, displayfunc.cpp:430-439) and sphere areas (4 pi r r,
displayfunc.cpp:303) in float32, one rounding per operation in the reference's order.
"""
from __future__ import annotations

import ctypes
import ctypes.util
from pathlib import Path

import numpy as np

F32 = np.float32
_libc = ctypes.CDLL(ctypes.util.find_library("c"))
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]

SPHERE, TRIANGLE = 0, 1
DIFF, SPEC, REFR = 0, 1, 2


def scene_float(tok: str) -> np.float32:
    end = ctypes.c_char_p()
    return F32(_libc.strtof(tok.encode(), ctypes.byref(end)))


def obj_float(tok: str) -> np.float32:
    """tryParseDouble (tinyobj_loader_c.h:267-391) for the forms sign? digits ('.' digits)?, then (float)."""
    m, i, sign = 0.0, 0, 1.0
    if tok and tok[0] in "+-":
        sign = -1.0 if tok[0] == "-" else 1.0
        i = 1
    read = 0
    while i < len(tok) and tok[i].isdigit():
        m = m * 10
        m = m + int(tok[i])
        i += 1
        read += 1
    if read == 0:
        return F32(0.0)
    if i < len(tok) and tok[i] == ".":
        i += 1
        read = 1
        while i < len(tok) and tok[i].isdigit():
            frac = 1.0
            for _ in range(read):
                frac = frac * 0.1
            m = m + int(tok[i]) * frac
            read += 1
            i += 1
    return F32(sign * (m * 1.0) * 1.0)


def read_obj(path: Path):
    """Vertices (float32 triples) and triangle faces (0-based) of an OBJ file."""
    verts, faces = [], []
    for line in path.read_text().splitlines():
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "v":
            verts.append([obj_float(t) for t in tok[1:4]])
        elif tok[0] == "f":
            assert len(tok) == 4, "triangles only"
            faces.append([int(t.split("/")[0]) - 1 for t in tok[1:4]])
    return verts, faces


def tri_area(p1, p2, p3) -> np.float32:
    e1 = [p2[k] - p1[k] for k in range(3)]
    e2 = [p3[k] - p1[k] for k in range(3)]
    n = [e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]]
    return F32(np.sqrt(F32(F32(n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]))) * F32(0.5)


def read_txt(path: Path, base: Path, shape_dtype, camera_dtype, width: int = 640, height: int = 480):
    """(shapes, lightCnt, camera) as ReadTxt builds them (the camera's basis not yet computed)."""
    tok = Path(path).read_text().split()
    i = 0

    def take(n=1):
        nonlocal i
        out = tok[i:i + n]
        i += n
        return out

    if tok[0] == "gamma":
        take(10)   # gamma g sky1 r g b sky2 r g b
    assert take()[0] == "camera"
    cam = np.zeros(1, dtype=camera_dtype)
    cam["orig"][0] = [scene_float(t) for t in take(3)]
    cam["target"][0] = [scene_float(t) for t in take(3)]
    cam["width"][0], cam["height"][0] = width, height
    assert take()[0] == "materials"
    for _ in range(int(take()[0])):
        kind = take()[0]
        take(4 if kind == "matte" else 7)
    assert take()[0] == "objects"
    rows, lights = [], 0

    def refl(mat):
        return SPEC if mat == 1 else REFR if mat == 2 else DIFF

    def emissive(e):
        return any(float(x) != 0.0 for x in e)

    for _ in range(int(take()[0])):
        kind = take()[0]
        if kind == "sphere":
            v = [scene_float(t) for t in take(10)]
            mat = int(take()[0])
            r = v[0]
            rows.append(dict(type=SPHERE, rad=r, p=v[1:4], e=v[4:7], c=v[7:10], refl=refl(mat),
                             area=F32(F32(F32(F32(4) * F32(3.14159265358979323846)) * r) * r)))
            lights += emissive(v[4:7])
        elif kind == "triangle":
            v = [scene_float(t) for t in take(15)]
            mat = int(take()[0])
            p1, p2, p3 = v[0:3], v[3:6], v[6:9]
            rows.append(dict(type=TRIANGLE, p1=p1, p2=p2, p3=p3, e=v[9:12], c=v[12:15], refl=refl(mat),
                             area=tri_area(p1, p2, p3)))
            lights += emissive(v[9:12])
        else:
            assert kind == "model", kind
            mpath = take()[0]
            v = [scene_float(t) for t in take(12)]
            mat = int(take()[0])
            pos, scl, e, c = v[0:3], v[3:6], v[6:9], v[9:12]
            verts, faces = read_obj(Path(base) / mpath)
            for f in faces:
                p = [[F32(verts[j][k] * scl[k]) + pos[k] for k in range(3)] for j in f]
                rows.append(dict(type=TRIANGLE, p1=p[0], p2=p[1], p3=p[2], e=e, c=c, refl=refl(mat),
                                 area=tri_area(*p)))
                lights += emissive(e)
    shapes = np.zeros(len(rows), dtype=shape_dtype)
    for k, r in enumerate(rows):
        shapes["type"][k] = r["type"]
        if r["type"] == SPHERE:
            shapes["rad"][k] = r["rad"]
            shapes["p"][k] = r["p"]
        else:
            shapes["p1"][k], shapes["p2"][k], shapes["p3"][k] = r["p1"], r["p2"], r["p3"]
        shapes["e"][k], shapes["c"][k] = r["e"], r["c"]
        shapes["refl"][k] = r["refl"]
        shapes["area"][k] = r["area"]
    return shapes, lights, cam
