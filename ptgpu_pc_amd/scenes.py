"""Seeded synthetic scenes for the benchmark configurations (numpy, deterministic).

Randomness comes from numpy's PCG64 generator (np.random.default_rng(seed)); every generator
returns (shapes, light_count, camera) with `area` filled as the reference's scene reader does
(sphere 4*pi*r*r, triangle |e1 x e2| * 0.5 in FP32, displayfunc.cpp:282,318) and lightCnt counted
explicitly (the reference's AddWallLight path sets neither, Q21).

  cornell(w, h)                    9-sphere Cornell box of scene.h:36-56 (WALL_RAD 1e4f), no accel
  soup(n, seed)                    n triangles, centroids U[0,1)^3, vertices +-0.01, plus an
                                   emissive quad at y=1.25 stored last
  icosphere_in_box(level, seed)    displaced icosphere (r = 1 + 0.05 * value noise) inside a
                                   12-triangle box whose two ceiling triangles emit (stored last)
  mixed_scene(seed, n)             soup or truncated displaced icosphere + emissive quad, for the
                                   batched / served workloads
"""
from __future__ import annotations

import hashlib
import math
from functools import lru_cache

import numpy as np

from .layout import DIFF, REFR, SHAPE_DTYPE, SPEC, SPHERE, TRIANGLE, concat_shapes, make_camera

F32 = np.float32
PI = F32(3.14159265358979323846)


def sphere_area(r: np.ndarray) -> np.ndarray:
    r = np.asarray(r, dtype=F32)
    return (F32(4) * PI * r) * r


def triangle_area(p1: np.ndarray, p2: np.ndarray, p3: np.ndarray) -> np.ndarray:
    e1 = (p2 - p1).astype(F32)
    e2 = (p3 - p1).astype(F32)
    cx = e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1]
    cy = e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2]
    cz = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]
    return (np.sqrt(cx * cx + cy * cy + cz * cz) * F32(0.5)).astype(F32)


def shapes_digest(shapes: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(shapes).tobytes()).hexdigest()


def triangles_to_shapes(tri: np.ndarray, e=None, c=(0.75, 0.75, 0.75), refl=DIFF) -> np.ndarray:
    """tri: (n, 3, 3) float32 vertices."""
    n = tri.shape[0]
    s = np.zeros(n, dtype=SHAPE_DTYPE)
    s["type"] = TRIANGLE
    s["p1"] = tri[:, 0]
    s["p2"] = tri[:, 1]
    s["p3"] = tri[:, 2]
    s["e"] = (0.0, 0.0, 0.0) if e is None else e
    s["c"] = c
    s["refl"] = refl
    s["area"] = triangle_area(tri[:, 0], tri[:, 1], tri[:, 2])
    return s


def light_count(shapes: np.ndarray) -> int:
    e = shapes["e"]
    return int(np.count_nonzero((e[:, 0] != 0) | (e[:, 1] != 0) | (e[:, 2] != 0)))


# ---------------------------------------------------------------------------------------------
def cornell(width: int = 256, height: int = 256):
    """CornellSpheres + t[] of scene.h:29-56; camera of smallptGPU.cpp:1419-1420."""
    W = F32(1e4)
    spheres = [
        (W, (W + 1.0, 40.8, 81.6), (0, 0, 0), (.75, .25, .25), DIFF),
        (W, (-W + 99.0, 40.8, 81.6), (0, 0, 0), (.25, .25, .75), DIFF),
        (W, (50.0, 40.8, W), (0, 0, 0), (.75, .75, .75), DIFF),
        (W, (50.0, 40.8, -W + 270.0), (0, 0, 0), (0, 0, 0), DIFF),
        (W, (50.0, W, 81.6), (0, 0, 0), (.75, .75, .75), DIFF),
        (W, (50.0, -W + 81.6, 81.6), (0, 0, 0), (.75, .75, .75), DIFF),
        (16.5, (27.0, 16.5, 47.0), (0, 0, 0), (.9, .9, .9), SPEC),
        (16.5, (73.0, 16.5, 78.0), (0, 0, 0), (.9, .9, .9), REFR),
        (7.0, (50.0, 81.6 - 15.0, 81.6), (12, 12, 12), (0, 0, 0), DIFF),
    ]
    s = np.zeros(len(spheres), dtype=SHAPE_DTYPE)
    for i, (r, p, e, c, refl) in enumerate(spheres):
        s[i]["type"] = SPHERE
        s[i]["rad"] = F32(r)
        # the reference evaluates e.g. WALL_RAD + 1.f and 81.6f - 15.f in float
        s[i]["p"] = np.array(p, dtype=np.float64).astype(F32)
        s[i]["e"] = e
        s[i]["c"] = c
        s[i]["refl"] = refl
    s["p"][0, 0] = W + F32(1.0)
    s["p"][1, 0] = -W + F32(99.0)
    s["p"][3, 2] = -W + F32(270.0)
    s["p"][5, 1] = -W + F32(81.6)
    s["p"][8, 1] = F32(81.6) - F32(15.0)
    s["area"] = sphere_area(s["rad"])
    cam = make_camera((50.0, 45.0, 205.6), (50.0, 45.0 - 0.042612, 204.6), 0.0, 0.0, width, height)
    return s, 1, cam


def soup(n: int = 65536, seed: int = 1, width: int = 512, height: int = 512):
    rng = np.random.default_rng(seed)
    cen = rng.random((n, 3), dtype=F32)
    off = rng.uniform(-0.01, 0.01, size=(n, 3, 3)).astype(F32)
    tri = (cen[:, None, :] + off).astype(F32)
    body = triangles_to_shapes(tri)
    quad = np.array([[(0.25, 1.25, 0.25), (0.75, 1.25, 0.25), (0.75, 1.25, 0.75)],
                     [(0.25, 1.25, 0.25), (0.75, 1.25, 0.75), (0.25, 1.25, 0.75)]], dtype=F32)
    light = triangles_to_shapes(quad, e=(12.0, 12.0, 12.0), c=(0.0, 0.0, 0.0))
    shapes = concat_shapes([body, light])
    cam = make_camera((0.5, 0.5, 2.5), (0.5, 0.5, 0.0), 0.0, 0.0, width, height)
    return shapes, 2, cam



def morton_chain(n: int = 6000, seed: int = 5, width: int = 80, height: int = 56):
    """n small triangles on 31 points whose quantised centroids are 0 and each power of two on one axis
    (10 per axis), plus the soup's light quad: Morton codes 0 and 2^0 .. 2^29, so the Morton-keyed LBVH
    is a chain about 30 levels deep with small balanced subtrees at its links (height 39 at n = 6000,
    seed 5, against 17 for a uniform soup of the same size)."""
    rng = np.random.default_rng(seed)
    pts = [(0.0, 0.0, 0.0)]
    for i in range(10):
        v = (2.0 ** i + 0.5) / 1024.0
        pts += [(v, 0.0, 0.0), (0.0, v, 0.0), (0.0, 0.0, v)]
    pts = np.array(pts) * np.array([1.0, 1.25, 1.0])   # the light's centroids set the y extent to 1.25
    cen = pts[rng.integers(0, len(pts), size=(n,))].astype(F32)
    cen[:8] = (1.0, 1.25, 1.0)                          # and these the x and z extents to 1
    off = rng.uniform(-1e-5, 1e-5, size=(n, 3, 3)).astype(F32)
    body = triangles_to_shapes((cen[:, None, :] + off).astype(F32))
    quad = np.array([[(0.25, 1.25, 0.25), (0.75, 1.25, 0.25), (0.75, 1.25, 0.75)],
                     [(0.25, 1.25, 0.25), (0.75, 1.25, 0.75), (0.25, 1.25, 0.75)]], dtype=F32)
    light = triangles_to_shapes(quad, e=(12.0, 12.0, 12.0), c=(0.0, 0.0, 0.0))
    cam = make_camera((0.5, 0.5, 2.5), (0.5, 0.5, 0.0), 0.0, 0.0, width, height)
    return concat_shapes([body, light]), 2, cam


# ---------------------------------------------------------------------------------------------
@lru_cache(maxsize=4)
def _icosphere(level: int):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                  (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
                  (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], dtype=np.int64)
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (len(v) + 1) + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk // (len(v) + 1), uk % (len(v) + 1)
        m = v[a] + v[b]
        m /= np.linalg.norm(m, axis=1, keepdims=True)
        mid = (len(v) + inv).reshape(3, -1).T   # midpoint ids for edges (01, 12, 20)
        v = np.concatenate([v, m])
        f = np.concatenate([
            np.stack([f[:, 0], mid[:, 0], mid[:, 2]], 1),
            np.stack([f[:, 1], mid[:, 1], mid[:, 0]], 1),
            np.stack([f[:, 2], mid[:, 2], mid[:, 1]], 1),
            np.stack([mid[:, 0], mid[:, 1], mid[:, 2]], 1),
        ])
    return v, f


def value_noise(p: np.ndarray, seed: int, freq: float = 4.0, grid: int = 32) -> np.ndarray:
    """Seeded 3D value noise in [-1, 1): random lattice values, smoothstep trilinear blend."""
    rng = np.random.default_rng(seed)
    lat = rng.uniform(-1.0, 1.0, size=(grid, grid, grid))
    q = p * freq + grid / 2.0
    i0 = np.floor(q).astype(np.int64)
    fr = q - i0
    w = fr * fr * (3.0 - 2.0 * fr)
    out = np.zeros(len(p))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                ww = (w[:, 0] if dx else 1 - w[:, 0]) * (w[:, 1] if dy else 1 - w[:, 1]) * \
                     (w[:, 2] if dz else 1 - w[:, 2])
                out += ww * lat[(i0[:, 0] + dx) % grid, (i0[:, 1] + dy) % grid, (i0[:, 2] + dz) % grid]
    return out


def displaced_icosphere(level: int, seed: int, center=(0.0, 0.0, 0.0), radius: float = 1.0) -> np.ndarray:
    v, f = _icosphere(level)
    r = radius * (1.0 + 0.05 * value_noise(v, seed))
    pv = v * r[:, None] + np.asarray(center)
    return pv[f].astype(F32)


def _box(lo: float, hi: float) -> np.ndarray:
    c = lambda x, y, z: (x, y, z)  # noqa: E731
    a, b = lo, hi
    quads = [
        (c(a, a, a), c(b, a, a), c(b, a, b), c(a, a, b)),   # floor
        (c(a, a, a), c(a, b, a), c(a, b, b), c(a, a, b)),   # left
        (c(b, a, a), c(b, b, a), c(b, b, b), c(b, a, b)),   # right
        (c(a, a, a), c(b, a, a), c(b, b, a), c(a, b, a)),   # back
        (c(a, a, b), c(b, a, b), c(b, b, b), c(a, b, b)),   # front
        (c(a, b, a), c(b, b, a), c(b, b, b), c(a, b, b)),   # ceiling (last: emissive)
    ]
    tris = []
    for q in quads:
        tris.append((q[0], q[1], q[2]))
        tris.append((q[0], q[2], q[3]))
    return np.array(tris, dtype=F32)


def icosphere_in_box(level: int = 7, seed: int = 3, width: int = 1024, height: int = 768):
    sph = triangles_to_shapes(displaced_icosphere(level, seed), c=(0.8, 0.6, 0.4))
    box = _box(-3.0, 3.0)
    walls = triangles_to_shapes(box[:10], c=(0.75, 0.75, 0.75))
    walls["c"][2:4] = (0.75, 0.25, 0.25)
    walls["c"][4:6] = (0.25, 0.25, 0.75)
    ceil = triangles_to_shapes(box[10:], e=(4.0, 4.0, 4.0), c=(0.0, 0.0, 0.0))
    shapes = concat_shapes([sph, walls, ceil])
    cam = make_camera((0.0, 0.0, 2.9), (0.0, 0.0, 0.0), 0.0, 0.0, width, height)
    return shapes, 2, cam


def orbit_camera(yaw: float, width: int, height: int) -> np.ndarray:
    """Camera of the batched / served workloads: on a radius-2 orbit around the unit cube's centre."""
    orig = (0.5 + 2.0 * math.sin(yaw), 0.5, 0.5 + 2.0 * math.cos(yaw))
    return make_camera(orig, (0.5, 0.5, 0.5), yaw, 0.0, width, height)


def mixed_scene(seed: int, n: int, width: int = 256, height: int = 256):
    """One scene of the batched / served workloads: n triangles of soup or displaced icosphere
    (truncated to n faces) in the unit cube, an emissive quad (stored last), camera on an orbit."""
    rng = np.random.default_rng(seed)
    kind = int(rng.integers(0, 2))
    yaw = float(rng.uniform(0.0, 2.0 * math.pi))
    if kind == 0:
        cen = rng.random((n, 3), dtype=F32)
        off = rng.uniform(-0.01, 0.01, size=(n, 3, 3)).astype(F32)
        tri = (cen[:, None, :] + off).astype(F32)
    else:
        level = 0
        while 20 * 4 ** level < n:
            level += 1
        tri = displaced_icosphere(level, seed, center=(0.5, 0.5, 0.5), radius=0.4)[:n]
    body = triangles_to_shapes(tri)
    quad = np.array([[(0.25, 1.25, 0.25), (0.75, 1.25, 0.25), (0.75, 1.25, 0.75)],
                     [(0.25, 1.25, 0.25), (0.75, 1.25, 0.75), (0.25, 1.25, 0.75)]], dtype=F32)
    light = triangles_to_shapes(quad, e=(12.0, 12.0, 12.0), c=(0.0, 0.0, 0.0))
    shapes = concat_shapes([body, light])
    return shapes, 2, orbit_camera(yaw, width, height), kind


def batch_set(n_scenes: int = 64, seed: int = 4000):
    """The many-scene workload: independent scenes with triangle counts log-uniform in [1e3, 1e6],
    soup or displaced icosphere by scene seed, a camera at a random orbit angle (UpdateCamera
    applied), 256^2 / 512^2 / 1024^2 by seed; rendered at 4 spp, depth 5."""
    from . import lib
    out = []
    for k in range(n_scenes):
        ss = seed + k
        r = np.random.default_rng([ss, 1])
        n = int(np.exp(r.uniform(np.log(1e3), np.log(1e6))))
        res = int(r.choice([256, 512, 1024]))
        shapes, lc, cam, kind = mixed_scene(ss, n, res, res)
        lib.update_camera(cam)
        out.append({"shapes": shapes, "lc": lc, "cam": cam, "w": res, "h": res, "kind": kind, "seed": ss})
    return out


def serve_trace(n_req: int, rate: float, pool: int, seed: int = 2024):
    """The served workload: Poisson arrivals at `rate`/s; each request picks one of `pool` scenes
    (triangle count log-uniform in [1e3, 2e6] and mesh type from the scene seed), a camera at a
    random orbit angle, a resolution in {256^2, 512^2, 1024^2} and spp in {1, 4, 16}; depth 5.
    Returns (pool_scenes [(scene seed, shapes, lightCnt, kind)], requests)."""
    rng = np.random.default_rng(seed)
    pool_scenes = []
    for k in range(pool):
        ss = 1000 + k
        n = int(np.exp(np.random.default_rng([ss, 1]).uniform(np.log(1e3), np.log(2e6))))
        shapes, lc, _, kind = mixed_scene(ss, n, 64, 64)
        pool_scenes.append((ss, shapes, lc, kind))
    t, reqs = 0.0, []
    for i in range(n_req):
        t += rng.exponential(1.0 / rate)
        sc = int(rng.integers(pool))
        res = int(rng.choice([256, 512, 1024]))
        spp = int(rng.choice([1, 4, 16]))
        yaw = float(rng.uniform(0, 2 * np.pi))
        reqs.append({"arrival": t, "scene": sc, "w": res, "h": res, "spp": spp, "yaw": yaw})
    return pool_scenes, reqs
