"""numpy views of the reference's binary layouts (geom.h, camera.h, BVHNodeGPU.h, KDNodeGPU.h).

The offsets are those of the C structures in include/ptgpu.h; the library reports its compiled
sizes through ptgpu_layout_sizes() and the loader checks them against these dtypes.
"""
from __future__ import annotations

import numpy as np

SPHERE, TRIANGLE = 0, 1          # geom.h:72
DIFF, SPEC, REFR = 0, 1, 2       # geom.h:55-57
FLOAT_PI = np.float32(3.14159265358979323846)   # geom.h:30

SHAPE_DTYPE = np.dtype({
    "names": ["type", "p1", "p2", "p3", "rad", "p", "e", "c", "refl", "index", "morton_code", "b", "area"],
    "formats": ["<i4", ("<f4", 3), ("<f4", 3), ("<f4", 3), "<f4", ("<f4", 3), ("<f4", 3), ("<f4", 3),
                "<i4", "<i4", "<u4", ("<f4", 6), "<f4"],
    "offsets": [0, 4, 16, 28, 4, 8, 40, 52, 64, 68, 72, 76, 100],
    "itemsize": 104,
})

CAMERA_DTYPE = np.dtype({
    "names": ["orig", "target", "pitch", "yaw", "width", "height", "dir", "x", "y"],
    "formats": [("<f4", 3), ("<f4", 3), "<f4", "<f4", "<f4", "<f4", ("<f4", 3), ("<f4", 3), ("<f4", 3)],
    "offsets": [0, 12, 24, 28, 32, 36, 40, 52, 64],
    "itemsize": 76,
})

BVHNODE_DTYPE = np.dtype({
    "names": ["min", "max", "nShape", "leaf", "nLeft", "nRight", "nParent", "bound", "cost", "area"],
    "formats": ["<i4", "<i4", "<i4", "<i4", "<i4", "<i4", "<i4", ("<f4", 6), "<f4", "<f4"],
    "offsets": [0, 4, 8, 12, 16, 20, 24, 28, 52, 56],
    "itemsize": 60,
})

KDNODE_DTYPE = np.dtype({
    "names": ["leaf", "nLeft", "nRight", "nParent", "bound", "min", "max"],
    "formats": ["<i2", "<i2", "<i2", "<i2", ("<f4", 6), "<i2", "<i2"],
    "offsets": [0, 2, 4, 6, 8, 32, 34],
    "itemsize": 36,
})

RAY_DTYPE = np.dtype({"names": ["o", "d"], "formats": [("<f4", 3), ("<f4", 3)], "offsets": [0, 12],
                      "itemsize": 24})

BOUND_DTYPE = np.dtype(("<f4", 6))   # min_x max_x min_y max_y min_z max_z


def empty_shapes(n: int) -> np.ndarray:
    return np.zeros(n, dtype=SHAPE_DTYPE)


def concat_shapes(parts) -> np.ndarray:
    """Concatenate SHAPE_DTYPE arrays byte-wise (np.concatenate repacks dtypes whose fields
    overlap, as the Triangle/Sphere union does, into a 120-byte layout)."""
    parts = [np.ascontiguousarray(p) for p in parts]
    for p in parts:
        if p.dtype != SHAPE_DTYPE:
            raise ValueError("not a SHAPE_DTYPE array")
    out = np.zeros(sum(len(p) for p in parts), dtype=SHAPE_DTYPE)
    raw = out.view(np.uint8).reshape(len(out), SHAPE_DTYPE.itemsize)
    k = 0
    for p in parts:
        raw[k:k + len(p)] = p.view(np.uint8).reshape(len(p), SHAPE_DTYPE.itemsize)
        k += len(p)
    return out


def make_camera(orig, target=(0.0, 0.0, 0.0), yaw: float = 0.0, pitch: float = 0.0, width: int = 256,
                height: int = 256) -> np.ndarray:
    cam = np.zeros(1, dtype=CAMERA_DTYPE)
    cam["orig"][0] = orig
    cam["target"][0] = target
    cam["yaw"][0] = yaw
    cam["pitch"][0] = pitch
    cam["width"][0] = width
    cam["height"][0] = height
    return cam


def default_seeds(width: int, height: int) -> np.ndarray:
    """Deterministic replacement for the reference's rand() seeds (smallptGPU.cpp:192-195):
    seeds[2*gid] = seeds[2*gid+1] = max(gid, 2) with gid = y*W + x."""
    gid = np.arange(width * height, dtype=np.uint32)
    s = np.maximum(gid, np.uint32(2))
    return np.repeat(s, 2).astype(np.uint32)
