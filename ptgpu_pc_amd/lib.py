"""ctypes host mirror of the reference's interface for the ray-acceleration and radiance path.

`CLBVH` keeps the method names and call order of the reference class (CLBVH.h:39-47):
    bvh = CLBVH(ctx, shapes)          # ctor: AABBs + Morton codes, written back into `shapes`
    bvh.buildRadixTree()              # Karras topology
    bvh.buildBVHTree()                # sort + refit
    bvh.optimize(7)                   # treelet restructuring
    btn, btl = bvh.getTrees()
`KDTree` mirrors KDTree::build / getTrees (KDTree.h:88-89).  `radiance_gpu` takes the RadianceGPU
argument list in its original order (rendering_kernel.cl:1072-1099).

Everything runs through libptgpu.so; there is no CPU fallback.  If the library is missing the
import of this module still works, but every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from pathlib import Path

import numpy as np

from .layout import (BVHNODE_DTYPE, CAMERA_DTYPE, KDNODE_DTYPE, RAY_DTYPE, SHAPE_DTYPE)

LIB_PATH = Path(__file__).resolve().parent / "libptgpu.so"

ACCEL_NONE, ACCEL_BVH, ACCEL_KD = 0, 1, 2
STATUS_NAMES = {0: "OK", 1: "INVALID", 2: "HIP", 3: "OOM", 4: "UNSUPPORTED", 5: "NO_DEVICE"}

_LIB = None
vp = C.c_void_p
i32 = C.c_int
u32 = C.c_uint
f32 = C.c_float


class PtgpuError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"ptgpu error {STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class RenderArgs(C.Structure):
    _fields_ = [("colors", vp), ("seeds", vp), ("camera", vp), ("shapes", vp), ("shapeCnt", u32),
                ("lightCnt", u32), ("width", i32), ("height", i32), ("currentSample", i32), ("spp", i32),
                ("max_depth", i32), ("pixels", vp), ("ray_counter", vp), ("trav_stats", vp)]


class Request(C.Structure):
    _fields_ = [("scene_key", C.c_ulonglong), ("h_shapes", vp), ("shapeCnt", u32), ("lightCnt", u32),
                ("camera", C.c_ubyte * 76), ("width", i32), ("height", i32), ("spp", i32), ("max_depth", i32),
                ("h_seeds", vp)]


class RenderJob(C.Structure):
    _fields_ = [("bvh", vp), ("kd", vp), ("args", RenderArgs)]


class Tuning(C.Structure):
    _fields_ = [("blocks_per_cu", i32), ("regroup", i32), ("wide", i32), ("batch_blocks_per_cu", i32),
                ("batch_regroup", i32), ("digit_bits", i32), ("treelet", i32), ("render_ms", f32), ("build_ms", f32)]

    def as_dict(self) -> dict:
        return {k: (round(getattr(self, k), 4) if isinstance(getattr(self, k), float) else getattr(self, k))
                for k, _ in self._fields_}


class RadianceArgs(C.Structure):
    _fields_ = [("colors", vp), ("seeds", vp), ("camera", vp), ("shapes", vp), ("shapeCnt", u32),
                ("lightCnt", u32), ("btn", vp), ("btl", vp), ("kng", vp), ("kngCnt", i32), ("kn", vp),
                ("knCnt", i32), ("width", i32), ("height", i32), ("currentSample", i32), ("pixels", vp)]


_SIGS = {
    "ptgpu_layout_sizes": [C.POINTER(i32)] * 6,
    "ptgpu_last_error": [],
    "ptgpu_device_count": [C.POINTER(i32)],
    "ptgpu_ctx_create": [i32, vp, C.POINTER(vp)],
    "ptgpu_ctx_destroy": [vp],
    "ptgpu_ctx_set_max_depth": [vp, i32],
    "ptgpu_ctx_set_engine": [vp, i32],
    "ptgpu_ctx_set_tuning": [vp, C.c_char_p, i32],
    "ptgpu_synchronize": [vp],
    "ptgpu_malloc": [vp, C.c_size_t, C.POINTER(vp)],
    "ptgpu_free": [vp, vp],
    "ptgpu_memcpy_h2d": [vp, vp, vp, C.c_size_t],
    "ptgpu_memcpy_d2h": [vp, vp, vp, C.c_size_t],
    "ptgpu_memset_d": [vp, vp, i32, C.c_size_t],
    "ptgpu_memcpy_d2d": [vp, vp, vp, C.c_size_t],
    "ptgpu_host_free": [vp],
    "ptgpu_event_record": [vp, i32],
    "ptgpu_ctx_wait_event": [vp, vp, i32],
    "ptgpu_copy_calibration": [vp, vp, vp, C.c_size_t],
    "ptgpu_event_elapsed_ms": [vp, i32, i32, C.POINTER(f32)],
    "ptgpu_clbvh_create": [vp, vp, i32, C.POINTER(vp)],
    "ptgpu_clbvh_create_device": [vp, vp, i32, C.POINTER(vp)],
    "ptgpu_clbvh_reset": [vp, i32],
    "ptgpu_clbvh_build_radix_tree": [vp],
    "ptgpu_clbvh_build_bvh_tree": [vp],
    "ptgpu_clbvh_optimize": [vp, i32],
    "ptgpu_clbvh_set_digit_bits": [vp, i32],
    "ptgpu_clbvh_set_traversal_topology": [vp, i32],
    "ptgpu_clbvh_set_height_readback": [vp, i32],
    "ptgpu_clbvh_export_traversal": [vp, vp, vp, C.POINTER(i32)],
    "ptgpu_clbvh_get_trees": [vp, C.POINTER(vp), C.POINTER(vp)],
    "ptgpu_clbvh_device_trees": [vp, C.POINTER(vp), C.POINTER(vp)],
    "ptgpu_clbvh_device_sorted": [vp, C.POINTER(vp)],
    "ptgpu_clbvh_stage_times": [vp, C.POINTER(f32)],
    "ptgpu_clbvh_stage_times6": [vp, C.POINTER(f32)],
    "ptgpu_clbvh_destroy": [vp],
    "ptgpu_clbvh_build_batch": [vp, vp, i32],
    "ptgpu_clbvh_create_batch": [vp, vp, vp, i32, vp],
    "ptgpu_kdtree_build": [vp, vp, i32, C.POINTER(vp)],
    "ptgpu_kdtree_build_sah": [vp, vp, i32, i32, i32, C.POINTER(vp)],
    "ptgpu_kdtree_get_trees": [vp, C.POINTER(vp), C.POINTER(i32), C.POINTER(vp), C.POINTER(i32)],
    "ptgpu_kdtree_export_dfs": [vp, vp, vp, vp, i32, i32],
    "ptgpu_kdtree_stats": [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)],
    "ptgpu_kdtree_destroy": [vp],
    "ptgpu_radiance_gpu": [vp, i32, C.POINTER(RadianceArgs)],
    "ptgpu_radiance_gpu_noalloc": [vp, i32, C.POINTER(RadianceArgs)],
    "ptgpu_radiance_box_gpu": [vp, i32, C.POINTER(RadianceArgs), i32, i32, i32, i32],
    "ptgpu_host_radiance_box": [C.POINTER(RadianceArgs), i32, i32, i32, i32, i32, i32, i32],
    "ptgpu_render_tiles": [vp, vp, vp, C.POINTER(RenderArgs), vp, i32, i32, i32],
    "ptgpu_hybrid_create": [vp, i32, C.POINTER(vp)],
    "ptgpu_hybrid_frame": [vp, vp, vp, C.POINTER(RenderArgs), i32, i32, C.POINTER(C.c_float), vp],
    "ptgpu_hybrid_destroy": [vp],
    "ptgpu_render": [vp, vp, vp, C.POINTER(RenderArgs)],
    "ptgpu_render_batch": [vp, vp, i32, vp],
    "ptgpu_server_create": [vp, C.c_size_t, i32, vp],
    "ptgpu_server_submit": [vp, C.POINTER(Request), vp],
    "ptgpu_server_tick": [vp, vp, vp],
    "ptgpu_server_finished": [vp, vp, i32, vp],
    "ptgpu_server_result": [vp, i32, vp, vp, vp],
    "ptgpu_server_stats": [vp, vp, vp, vp, vp],
    "ptgpu_server_destroy": [vp],
    "ptgpu_clbvh_device_bytes": [vp, vp],
    "ptgpu_autotune": [vp, vp, C.POINTER(RenderArgs), i32, i32, C.POINTER(Tuning)],
    "ptgpu_scene_read": [C.c_char_p, C.c_char_p, i32, i32, C.POINTER(vp), C.POINTER(u32), C.POINTER(u32), vp],
    "ptgpu_write_ppm": [C.c_char_p, vp, i32, i32],
    "ptgpu_graph_begin": [vp],
    "ptgpu_graph_end": [vp, C.POINTER(vp)],
    "ptgpu_graph_launch": [vp, vp],
    "ptgpu_graph_destroy": [vp],
    "ptgpu_memcpy_async": [vp, vp, vp, C.c_size_t],
    "ptgpu_pinned_alloc": [vp, C.c_size_t, C.POINTER(vp)],
    "ptgpu_pinned_free": [vp],
    "ptgpu_autotune_batch": [vp, vp, i32, i32, C.POINTER(Tuning)],
    "ptgpu_trace_closest": [vp, vp, vp, vp, u32, vp, i32, vp, vp],
    "ptgpu_trace_any": [vp, vp, vp, vp, u32, vp, vp, i32, vp],
    "ptgpu_update_camera": [vp],
}


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libptgpu.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(str(p))
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_char_p if name == "ptgpu_last_error" else (None if name == "ptgpu_host_free" else i32)
    sizes = [i32() for _ in range(6)]
    lib.ptgpu_layout_sizes(*[C.byref(s) for s in sizes])
    got = [s.value for s in sizes]
    want = [SHAPE_DTYPE.itemsize, CAMERA_DTYPE.itemsize, BVHNODE_DTYPE.itemsize, KDNODE_DTYPE.itemsize,
            RAY_DTYPE.itemsize, 24]
    if got != want:
        raise RuntimeError(f"layout mismatch between libptgpu.so {got} and layout.py {want}")
    if path is None:
        _LIB = lib
    return lib


def _check(status: int) -> None:
    if status != 0:
        msg = _LIB.ptgpu_last_error().decode(errors="replace") if _LIB else "library not loaded"
        raise PtgpuError(status, msg)


def _ptr(a: np.ndarray) -> int:
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


def device_count() -> int:
    lib = load_library()
    c = i32()
    _check(lib.ptgpu_device_count(C.byref(c)))
    return c.value


def read_scene(path, base_dir=None, width: int = 640, height: int = 480):
    """ReadTxt: (shapes SHAPE_DTYPE array, lightCnt, camera CAMERA_DTYPE array with the basis not yet
    computed -- pass it to update_camera) from a reference text scene; models resolve against base_dir."""
    lib = load_library()
    p, n, lc = vp(), u32(), u32()
    cam = np.zeros(1, dtype=CAMERA_DTYPE)
    _check(lib.ptgpu_scene_read(str(path).encode(), str(base_dir).encode() if base_dir else None, int(width),
                                int(height), C.byref(p), C.byref(n), C.byref(lc), _ptr(cam)))
    try:
        shapes = np.empty(n.value, dtype=SHAPE_DTYPE)
        C.memmove(_ptr(shapes), p.value, shapes.nbytes)
    finally:
        lib.ptgpu_host_free(p.value)
    return shapes, lc.value, cam


def write_ppm(path, pixels: np.ndarray, width: int, height: int) -> None:
    """The reference's PPM writer over packed RGBA pixels (width*height uint32)."""
    px = np.ascontiguousarray(pixels, dtype=np.uint32)
    if px.size != width * height:
        raise ValueError("pixels: one uint32 per pixel")
    _check(load_library().ptgpu_write_ppm(str(path).encode(), _ptr(px), int(width), int(height)))


def update_camera(cam: np.ndarray) -> np.ndarray:
    """UpdateCamera (displayfunc.cpp:732-756) in place on a 1-element CAMERA_DTYPE array."""
    lib = load_library()
    _check(lib.ptgpu_update_camera(_ptr(cam)))
    return cam


class DeviceArray:
    """A device allocation owned by a Context."""

    def __init__(self, ctx: "Context", nbytes: int, dtype=np.uint8, shape=None):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.shape = shape if shape is not None else (self.nbytes // max(1, self.dtype.itemsize),)
        p = vp()
        _check(ctx.lib.ptgpu_malloc(ctx.handle, max(4, self.nbytes), C.byref(p)))
        self.ptr = p.value
        ctx._adopt(self)

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("host array larger than the allocation")
        _check(self.ctx.lib.ptgpu_memcpy_h2d(self.ctx.handle, self.ptr, _ptr(a), a.nbytes))
        return self

    def download(self, dtype=None, shape=None) -> np.ndarray:
        dt = np.dtype(dtype) if dtype is not None else self.dtype
        shp = shape if shape is not None else (self.nbytes // dt.itemsize,)
        out = np.empty(shp, dtype=dt)
        _check(self.ctx.lib.ptgpu_memcpy_d2h(self.ctx.handle, _ptr(out), self.ptr, out.nbytes))
        return out

    def copy_from(self, other: "DeviceArray") -> "DeviceArray":
        """Stream-ordered device-to-device copy."""
        _check(self.ctx.lib.ptgpu_memcpy_d2d(self.ctx.handle, self.ptr, other.ptr, min(self.nbytes, other.nbytes)))
        return self

    def zero(self) -> "DeviceArray":
        _check(self.ctx.lib.ptgpu_memset_d(self.ctx.handle, self.ptr, 0, self.nbytes))
        return self

    def free(self) -> None:
        if self.ptr and self.ctx.handle:
            _check(self.ctx.lib.ptgpu_free(self.ctx.handle, self.ptr))
        self.ptr = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


ENGINES = {"auto": 0, "megakernel": 1, "wavefront": 2, "lanes": 3}


class Context:
    """One device + one HIP stream (replaces the OpenCL context/queue, smallptGPU.cpp:286-441)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        self.handle = None
        self._children = weakref.WeakSet()
        h = vp()
        _check(self.lib.ptgpu_ctx_create(device, None, C.byref(h)))
        self.handle = h.value

    def _adopt(self, obj) -> None:
        self._children.add(obj)

    def close(self) -> None:
        """Releases every handle and allocation made through this context, then the context."""
        if self.handle:
            for child in list(self._children):
                try:
                    child.close()
                except Exception:
                    pass
            self.lib.ptgpu_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_engine(self, engine: str) -> None:
        """'auto', 'megakernel' or 'wavefront' (same results)."""
        _check(self.lib.ptgpu_ctx_set_engine(self.handle, ENGINES[engine]))

    def set_tuning(self, name: str, value: int) -> None:
        _check(self.lib.ptgpu_ctx_set_tuning(self.handle, name.encode(), int(value)))

    def set_max_depth(self, d: int) -> None:
        _check(self.lib.ptgpu_ctx_set_max_depth(self.handle, d))

    def synchronize(self) -> None:
        _check(self.lib.ptgpu_synchronize(self.handle))

    def alloc(self, nbytes: int, dtype=np.uint8, shape=None) -> DeviceArray:
        return DeviceArray(self, nbytes, dtype, shape)

    def to_device(self, a: np.ndarray) -> DeviceArray:
        a = np.ascontiguousarray(a)
        return DeviceArray(self, a.nbytes, a.dtype, a.shape).upload(a)

    def copy_calibration(self, dst: DeviceArray, src: DeviceArray, nbytes: int) -> None:
        _check(self.lib.ptgpu_copy_calibration(self.handle, dst.ptr, src.ptr, nbytes))

    def graph_begin(self) -> None:
        """Start recording this context's library calls into a HIP graph (see ptgpu_graph_begin)."""
        _check(self.lib.ptgpu_graph_begin(self.handle))

    def graph_end(self) -> "Graph":
        g = vp()
        _check(self.lib.ptgpu_graph_end(self.handle, C.byref(g)))
        return Graph(self, g.value)

    def pinned(self, nbytes: int, dtype=np.uint8) -> "PinnedArray":
        return PinnedArray(self, nbytes, dtype)

    def memcpy_async(self, dst, src, nbytes: int) -> None:
        """Stream-ordered copy between DeviceArrays and PinnedArrays (recordable into a graph)."""
        _check(self.lib.ptgpu_memcpy_async(self.handle, dst.ptr, src.ptr, int(nbytes)))

    def record(self, slot: int) -> None:
        _check(self.lib.ptgpu_event_record(self.handle, slot))

    def wait_event(self, other: "Context", slot: int) -> None:
        """Later work on this context waits for event `slot` last recorded on `other`."""
        _check(self.lib.ptgpu_ctx_wait_event(self.handle, other.handle, slot))

    def elapsed_ms(self, a: int, b: int) -> float:
        ms = f32()
        _check(self.lib.ptgpu_event_elapsed_ms(self.handle, a, b, C.byref(ms)))
        return ms.value


class Graph:
    """A recorded per-frame call sequence (ptgpu_graph_*); launch() replays it on the context's stream."""

    def __init__(self, ctx: "Context", handle):
        self.ctx, self.handle = ctx, handle
        ctx._adopt(self)

    def launch(self) -> None:
        _check(self.ctx.lib.ptgpu_graph_launch(self.ctx.handle, self.handle))

    def close(self) -> None:
        if self.handle:
            self.ctx.lib.ptgpu_graph_destroy(self.handle)
        self.handle = None


class PinnedArray:
    """Page-locked host memory (ptgpu_pinned_alloc) with a numpy view, for recordable copies."""

    def __init__(self, ctx: "Context", nbytes: int, dtype=np.uint8):
        p = vp()
        _check(ctx.lib.ptgpu_pinned_alloc(ctx.handle, int(nbytes), C.byref(p)))
        self.ctx, self.ptr, self.nbytes = ctx, p.value, int(nbytes)
        self.array = np.ctypeslib.as_array((C.c_ubyte * self.nbytes).from_address(self.ptr)).view(dtype)
        ctx._adopt(self)

    def close(self) -> None:
        if self.ptr:
            self.array = None
            self.ctx.lib.ptgpu_pinned_free(self.ptr)
        self.ptr = None


class CLBVH:
    """Mirror of class CLBVH (CLBVH.h:39-47).

    shapes: a host SHAPE_DTYPE array (index/morton_code/b are written back into it, as the
    reference does) or a DeviceArray holding n shapes (updated in place on the device)."""

    def __init__(self, ctx: Context, shapes, n: int | None = None, digit_bits: int = 8, traversal: str = "morton"):
        self.ctx = ctx
        h = vp()
        if isinstance(shapes, DeviceArray):
            self.n = int(n if n is not None else shapes.nbytes // SHAPE_DTYPE.itemsize)
            _check(ctx.lib.ptgpu_clbvh_create_device(ctx.handle, shapes.ptr, self.n, C.byref(h)))
            self.d_shapes = shapes
        else:
            if shapes.dtype != SHAPE_DTYPE or not shapes.flags["C_CONTIGUOUS"]:
                raise ValueError("shapes must be a C-contiguous SHAPE_DTYPE array")
            self.n = int(n if n is not None else len(shapes))
            _check(ctx.lib.ptgpu_clbvh_create(ctx.handle, _ptr(shapes), self.n, C.byref(h)))
            self.d_shapes = None
        self.handle = h.value
        ctx._adopt(self)
        if digit_bits != 8:
            _check(ctx.lib.ptgpu_clbvh_set_digit_bits(self.handle, digit_bits))
        if traversal != "morton":
            self.set_traversal(traversal)

    @classmethod
    def create_batch(cls, ctx: Context, shapes: list, digit_bits: int = 8) -> list["CLBVH"]:
        """Handles over many device shape arrays (DeviceArray each), all built in one batched sequence of
        launches (ptgpu_clbvh_create_batch); each equals CLBVH(ctx, shapes[s]) + buildRadixTree + buildBVHTree."""
        k = len(shapes)
        ptrs = (vp * max(k, 1))(*[a.ptr for a in shapes])
        counts = (i32 * max(k, 1))(*[a.nbytes // SHAPE_DTYPE.itemsize for a in shapes])
        out = (vp * max(k, 1))()
        if digit_bits != 8:   # the digit width is a per-handle setting: create, set, then build
            hs = [cls(ctx, a) for a in shapes]
            for h in hs:
                h.set_digit_bits(digit_bits)
            cls.build_batch(ctx, hs)
            return hs
        _check(ctx.lib.ptgpu_clbvh_create_batch(ctx.handle, ptrs, counts, k, out))
        hs = []
        for s in range(k):
            h = cls.__new__(cls)
            h.ctx, h.n, h.d_shapes, h.handle = ctx, int(counts[s]), shapes[s], out[s]
            ctx._adopt(h)
            hs.append(h)
        return hs

    @staticmethod
    def build_batch(ctx: Context, handles: list["CLBVH"]) -> None:
        """Rebuild every handle from its current device shapes in one batched sequence of launches
        (ptgpu_clbvh_build_batch): reset + buildRadixTree + buildBVHTree per handle."""
        arr = (vp * max(len(handles), 1))(*[h.handle for h in handles])
        _check(ctx.lib.ptgpu_clbvh_build_batch(ctx.handle, arr, len(handles)))

    def set_height_readback(self, readback: bool) -> None:
        """False: after the first height read back, later builds are not waited for; renders size their
        stack for that height + 4 and check the built height on the device (ptgpu_clbvh_set_height_readback)."""
        _check(self.ctx.lib.ptgpu_clbvh_set_height_readback(self.handle, 1 if readback else 0))

    def set_traversal(self, topology: str) -> None:
        """'morton' (Morton-keyed LBVH, default) or 'reference' (the exported index-keyed tree)."""
        _check(self.ctx.lib.ptgpu_clbvh_set_traversal_topology(self.handle, {"reference": 0, "morton": 1}[topology]))

    def set_digit_bits(self, bits: int) -> None:
        _check(self.ctx.lib.ptgpu_clbvh_set_digit_bits(self.handle, bits))

    def reset(self, n: int | None = None) -> None:
        _check(self.ctx.lib.ptgpu_clbvh_reset(self.handle, int(n if n is not None else self.n)))

    def buildRadixTree(self) -> None:
        _check(self.ctx.lib.ptgpu_clbvh_build_radix_tree(self.handle))

    def buildBVHTree(self) -> None:
        _check(self.ctx.lib.ptgpu_clbvh_build_bvh_tree(self.handle))

    def optimize(self, treelet_size: int = 7) -> None:
        _check(self.ctx.lib.ptgpu_clbvh_optimize(self.handle, treelet_size))

    def getTrees(self):
        """Returns copies of (btn[N-1], btl[N]) as BVHNODE_DTYPE arrays."""
        a, b = vp(), vp()
        _check(self.ctx.lib.ptgpu_clbvh_get_trees(self.handle, C.byref(a), C.byref(b)))
        nn = max(self.n - 1, 0)
        btn = np.empty(nn, dtype=BVHNODE_DTYPE)
        btl = np.empty(self.n, dtype=BVHNODE_DTYPE)
        if nn:
            C.memmove(_ptr(btn), a.value, btn.nbytes)
        C.memmove(_ptr(btl), b.value, btl.nbytes)
        return btn, btl

    def export_traversal(self):
        """(refs4 [N-1,4] = left, right, lo, hi; boxes [N-1,2,6]; height) of the traversal tree."""
        m = max(self.n - 1, 0)
        refs = np.zeros((max(m, 1), 4), dtype=np.int32)
        boxes = np.zeros((max(m, 1), 2, 6), dtype=np.float32)
        hgt = i32()
        _check(self.ctx.lib.ptgpu_clbvh_export_traversal(self.handle, _ptr(refs), _ptr(boxes), C.byref(hgt)))
        return refs[:m], boxes[:m], hgt.value

    def device_trees(self):
        a, b = vp(), vp()
        _check(self.ctx.lib.ptgpu_clbvh_device_trees(self.handle, C.byref(a), C.byref(b)))
        return a.value, b.value

    def sorted_indices(self) -> np.ndarray:
        p = vp()
        _check(self.ctx.lib.ptgpu_clbvh_device_sorted(self.handle, C.byref(p)))
        out = np.empty(self.n, dtype=np.int32)
        _check(self.ctx.lib.ptgpu_memcpy_d2h(self.ctx.handle, _ptr(out), p.value, out.nbytes))
        return out

    def stage_times(self) -> dict:
        ms = (f32 * 5)()
        _check(self.ctx.lib.ptgpu_clbvh_stage_times(self.handle, ms))
        return dict(zip(["morton", "sort", "topology", "refit", "optimize"], list(ms)))

    def stage_times6(self) -> dict:
        """As stage_times, with the refit split into the exported tree's refit and the traversal tree."""
        ms = (f32 * 6)()
        _check(self.ctx.lib.ptgpu_clbvh_stage_times6(self.handle, ms))
        return dict(zip(["morton", "sort", "topology", "refit", "traversal_tree", "optimize"], list(ms)))

    def close(self) -> None:
        if self.handle and self.ctx.handle:
            self.ctx.lib.ptgpu_clbvh_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class KDTree:
    """Mirror of class KDTree (KDTree.h:47-90): build() then getTrees()."""

    def __init__(self, ctx: Context | None, shapes: np.ndarray, sah: bool = False, max_depth: int = 24,
                 max_leaf: int = 4):
        """ctx=None builds the host tree only (exports work, rendering does not)."""
        if shapes.dtype != SHAPE_DTYPE:
            raise ValueError("shapes must be a SHAPE_DTYPE array")
        self.ctx = ctx
        self.lib = ctx.lib if ctx is not None else load_library()
        self.n = len(shapes)
        shapes = np.ascontiguousarray(shapes)
        h = vp()
        ch = ctx.handle if ctx is not None else None
        if sah:
            _check(self.lib.ptgpu_kdtree_build_sah(ch, _ptr(shapes), self.n, max_depth, max_leaf, C.byref(h)))
        else:
            _check(self.lib.ptgpu_kdtree_build(ch, _ptr(shapes), self.n, C.byref(h)))
        self.handle = h.value
        if ctx is not None:
            ctx._adopt(self)

    def stats(self) -> dict:
        v = [i32() for _ in range(4)]
        _check(self.lib.ptgpu_kdtree_stats(self.handle, *[C.byref(x) for x in v]))
        return dict(zip(["max_depth", "nodes", "leaves", "kn_len"], [x.value for x in v]))

    def getTrees(self):
        """KDTree::getTrees: (kng heap as KDNODE_DTYPE array, kn int32 array)."""
        pk, pn = vp(), vp()
        ck, cn = i32(), i32()
        _check(self.lib.ptgpu_kdtree_get_trees(self.handle, C.byref(pk), C.byref(ck), C.byref(pn), C.byref(cn)))
        kng = np.empty(ck.value, dtype=KDNODE_DTYPE)
        kn = np.empty(cn.value, dtype=np.int32)
        C.memmove(_ptr(kng), pk.value, kng.nbytes)
        C.memmove(_ptr(kn), pn.value, kn.nbytes)
        self.lib.ptgpu_host_free(pk.value)
        self.lib.ptgpu_host_free(pn.value)
        return kng, kn

    def export_dfs(self):
        st = self.stats()
        node5 = np.empty((st["nodes"], 5), dtype=np.int32)
        box = np.empty((st["nodes"], 6), dtype=np.float32)
        kn = np.empty(max(st["kn_len"], 1), dtype=np.int32)
        _check(self.lib.ptgpu_kdtree_export_dfs(self.handle, _ptr(node5), _ptr(box), _ptr(kn), st["nodes"],
                                                len(kn)))
        return node5, box, kn[:st["kn_len"]]

    def close(self) -> None:
        if self.handle and (self.ctx is None or self.ctx.handle):
            self.lib.ptgpu_kdtree_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _h(x):
    return x.handle if x is not None else None


def _render_args(colors, seeds, camera, shapes, shapeCnt, lightCnt, width, height, currentSample=0, spp=1,
                 max_depth=6, pixels=None, ray_counter=None, trav_stats=None) -> RenderArgs:
    return RenderArgs(colors.ptr, seeds.ptr, camera.ptr, shapes.ptr, shapeCnt, lightCnt, width, height, currentSample,
                      spp, max_depth, pixels.ptr if pixels else None, ray_counter.ptr if ray_counter else None,
                      trav_stats.ptr if trav_stats else None)


def render(ctx: Context, bvh: CLBVH | None, kd: KDTree | None, *, colors: DeviceArray, seeds: DeviceArray,
           camera: DeviceArray, shapes: DeviceArray, shapeCnt: int, lightCnt: int, width: int, height: int,
           currentSample: int = 0, spp: int = 1, max_depth: int = 6, pixels: DeviceArray | None = None,
           ray_counter: DeviceArray | None = None, trav_stats: DeviceArray | None = None) -> None:
    """`spp` RadianceGPU passes in one launch over a prepared BVH / KD tree (or brute force)."""
    a = _render_args(colors, seeds, camera, shapes, shapeCnt, lightCnt, width, height, currentSample, spp, max_depth,
                     pixels, ray_counter, trav_stats)
    _check(ctx.lib.ptgpu_render(ctx.handle, _h(bvh), _h(kd), C.byref(a)))


def render_tiles(ctx: Context, bvh: CLBVH | None, kd: KDTree | None, tiles: np.ndarray, tile_w: int, tile_h: int,
                 **render_kwargs) -> None:
    """`render` over the listed tiles only: tiles (n, 2) int32 top-left pixels of tile_w x tile_h tiles."""
    t = np.ascontiguousarray(tiles, dtype=np.int32).reshape(-1, 2)
    d_t = ctx.to_device(t) if len(t) else None
    a = _render_args(**render_kwargs)
    _check(ctx.lib.ptgpu_render_tiles(ctx.handle, _h(bvh), _h(kd), C.byref(a), d_t.ptr if d_t else None, len(t),
                                      int(tile_w), int(tile_h)))
    if d_t:
        ctx.synchronize()
        d_t.free()


class HybridStats(C.Structure):
    _fields_ = [("gpu_tiles", i32), ("cpu_tiles", i32), ("gpu_ms", C.c_float), ("cpu_ms", C.c_float),
                ("total_ms", C.c_float)]


class Hybrid:
    """The reference's CPU_PARTRENDERING frame (DrawAllBoxes): tiles split between the GPU and `threads`
    host threads by the ratio rcpu, which each frame updates from the measured time per tile."""

    def __init__(self, ctx: Context, threads: int):
        self.ctx = ctx
        h = vp()
        _check(ctx.lib.ptgpu_hybrid_create(ctx.handle, int(threads), C.byref(h)))
        self.handle = h.value
        ctx._children.add(self)

    def frame(self, bvh: CLBVH | None, kd: KDTree | None, bwidth: int, bheight: int, rcpu: float,
              **render_kwargs) -> tuple:
        """One frame; returns (the updated rcpu, stats dict)."""
        a = _render_args(**render_kwargs)
        r = C.c_float(rcpu)
        st = HybridStats()
        _check(self.ctx.lib.ptgpu_hybrid_frame(self.handle, _h(bvh), _h(kd), C.byref(a), int(bwidth), int(bheight),
                                               C.byref(r), C.byref(st)))
        return r.value, {k: getattr(st, k) for k, _ in HybridStats._fields_}

    def close(self) -> None:
        if self.handle and self.ctx.handle:
            self.ctx.lib.ptgpu_hybrid_destroy(self.handle)
        self.handle = None


def host_radiance_box(accel: int, colors: np.ndarray, seeds: np.ndarray, camera: np.ndarray, shapes: np.ndarray,
                      lightCnt: int, width: int, height: int, currentSample: int, pixels: np.ndarray, x: int, y: int,
                      bwidth: int, bheight: int, max_depth: int = 6, btn=None, btl=None, kng=None, kn=None,
                      threads: int = 1) -> None:
    """DrawBox on the host (ptgpu_host_radiance_box): one sample of one box, in place on host arrays."""
    lib = load_library()
    p = lambda x: _ptr(x) if x is not None else None  # noqa: E731
    a = RadianceArgs(p(colors), p(seeds), p(camera), p(shapes), len(shapes), lightCnt, p(btn), p(btl), p(kng),
                     len(kng) if kng is not None else 0, p(kn), len(kn) if kn is not None else 0, width, height,
                     currentSample, p(pixels))
    _check(lib.ptgpu_host_radiance_box(C.byref(a), accel, max_depth, x, y, bwidth, bheight, threads))


def autotune(ctx: Context, bvh: CLBVH, *, frames_per_build: int = 1, trials: int = 2, **render_kwargs) -> dict:
    """ptgpu_autotune: digit width, launch shape and treelet size for this scene and frame (the keyword
    arguments of `render`); applied to ctx and bvh, which is left holding a tree built with them."""
    a = _render_args(**render_kwargs)
    t = Tuning()
    _check(ctx.lib.ptgpu_autotune(ctx.handle, _h(bvh), C.byref(a), int(frames_per_build), int(trials), C.byref(t)))
    return t.as_dict()


def autotune_batch(ctx: Context, jobs: list, trials: int = 2) -> dict:
    """ptgpu_autotune_batch: launch shape of ptgpu_render_batch on these jobs (dicts as for render_batch)."""
    arr = _job_table(jobs)
    t = Tuning()
    _check(ctx.lib.ptgpu_autotune_batch(ctx.handle, C.cast(arr, vp), len(jobs), int(trials), C.byref(t)))
    return t.as_dict()


class Server:
    """Continuous batching of render requests (ptgpu_server_*): submit frames as they arrive, call
    tick() to admit and advance every admitted frame by spp_per_tick samples in one launch."""

    def __init__(self, ctx: Context, hbm_budget: int, spp_per_tick: int = 1):
        self.ctx = ctx
        h = vp()
        _check(ctx.lib.ptgpu_server_create(ctx.handle, int(hbm_budget), int(spp_per_tick), C.byref(h)))
        self.handle = h.value
        self._keep = {}     # host arrays a queued request points at
        self._size = {}
        ctx._children.add(self)

    def submit(self, scene_key: int, shapes: np.ndarray, light_cnt: int, camera: np.ndarray, width: int, height: int,
               spp: int, max_depth: int = 5, seeds: np.ndarray | None = None) -> int:
        if shapes.dtype.itemsize != SHAPE_DTYPE.itemsize or camera.dtype.itemsize != CAMERA_DTYPE.itemsize:
            raise ValueError("shapes / camera must use the reference layouts")
        shapes = np.ascontiguousarray(shapes)
        r = Request()
        r.scene_key = int(scene_key)
        r.h_shapes = _ptr(shapes)
        r.shapeCnt = len(shapes)
        r.lightCnt = int(light_cnt)
        C.memmove(r.camera, np.ascontiguousarray(camera).ctypes.data, 76)
        r.width, r.height, r.spp, r.max_depth = int(width), int(height), int(spp), int(max_depth)
        if seeds is not None:
            seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
            if seeds.size != 2 * width * height:
                raise ValueError("seeds: 2 per pixel")
        r.h_seeds = _ptr(seeds) if seeds is not None else None
        rid = i32()
        _check(self.ctx.lib.ptgpu_server_submit(self.handle, C.byref(r), C.byref(rid)))
        self._keep[rid.value] = (shapes, seeds)
        self._size[rid.value] = (int(width), int(height))
        return rid.value

    def tick(self) -> tuple:
        a, q = i32(), i32()
        _check(self.ctx.lib.ptgpu_server_tick(self.handle, C.byref(a), C.byref(q)))
        return a.value, q.value

    def finished(self) -> list:
        out = []
        ids = (i32 * 256)()
        while True:
            n = i32()
            _check(self.ctx.lib.ptgpu_server_finished(self.handle, ids, 256, C.byref(n)))
            out.extend(ids[:n.value])
            if n.value < 256:
                return out

    def result(self, rid: int, colors: bool = True, pixels: bool = True, seeds: bool = False):
        w, h = self._size[rid]
        c = np.empty(w * h * 3, np.float32) if colors else None
        p = np.empty(w * h, np.uint32) if pixels else None
        s = np.empty(w * h * 2, np.uint32) if seeds else None
        _check(self.ctx.lib.ptgpu_server_result(self.handle, rid, _ptr(c) if c is not None else None,
                                                _ptr(p) if p is not None else None,
                                                _ptr(s) if s is not None else None))
        self._keep.pop(rid, None)
        self._size.pop(rid, None)
        return c, p, s

    def stats(self) -> dict:
        rays = (C.c_ulonglong * 2)()
        used, scenes, builds = C.c_size_t(), i32(), C.c_ulonglong()
        _check(self.ctx.lib.ptgpu_server_stats(self.handle, rays, C.byref(used), C.byref(scenes), C.byref(builds)))
        return {"closest": rays[0], "anyhit": rays[1], "hbm_used": used.value, "scenes": scenes.value,
                "builds": builds.value}

    def close(self) -> None:
        if self.handle and self.ctx.handle:
            self.ctx.lib.ptgpu_server_destroy(self.handle)
        self.handle = None
        self._keep.clear()


def _job_table(jobs: list):
    arr = (RenderJob * max(len(jobs), 1))()
    for k, j in enumerate(jobs):
        px = j.get("pixels")
        arr[k].bvh = _h(j.get("bvh"))
        arr[k].kd = _h(j.get("kd"))
        arr[k].args = RenderArgs(j["colors"].ptr, j["seeds"].ptr, j["camera"].ptr, j["shapes"].ptr, j["shapeCnt"],
                                 j["lightCnt"], j["width"], j["height"], j.get("currentSample", 0), j.get("spp", 1),
                                 j.get("max_depth", 6), px.ptr if px else None, None, None)
    return arr


def render_batch(ctx: Context, jobs: list, ray_counter: DeviceArray | None = None) -> None:
    """Many frames in one launch.  `jobs`: dicts with the keyword arguments of `render` plus exactly
    one of `bvh` / `kd`; each job's outputs equal its own `render` call."""
    arr = _job_table(jobs)
    _check(ctx.lib.ptgpu_render_batch(ctx.handle, C.cast(arr, vp) if jobs else None, len(jobs),
                                      ray_counter.ptr if ray_counter else None))


def radiance_gpu(ctx: Context, accel: int, colors, seeds, camera, shapes, shapeCnt, lightCnt, btn=None, btl=None,
                 kng=None, kngCnt=0, kn=None, knCnt=0, width=0, height=0, currentSample=0, pixels=None,
                 noalloc: bool = False, box=None) -> None:
    """One RadianceGPU launch with the reference argument list (device pointers or DeviceArrays).
    noalloc: through ptgpu_radiance_gpu_noalloc (the library's static scratch; allocates nothing).
    box = (x, y, bwidth, bheight): RadianceBoxGPU over that box only."""
    def p(x):
        if x is None:
            return None
        return x.ptr if isinstance(x, DeviceArray) else int(x)
    a = RadianceArgs(p(colors), p(seeds), p(camera), p(shapes), shapeCnt, lightCnt, p(btn), p(btl), p(kng), kngCnt,
                     p(kn), knCnt, width, height, currentSample, p(pixels))
    if box is not None:
        _check(ctx.lib.ptgpu_radiance_box_gpu(ctx.handle, accel, C.byref(a), *[int(v) for v in box]))
        return
    fn = ctx.lib.ptgpu_radiance_gpu_noalloc if noalloc else ctx.lib.ptgpu_radiance_gpu
    _check(fn(ctx.handle, accel, C.byref(a)))


def trace_closest(ctx: Context, bvh, kd, shapes: DeviceArray, shapeCnt: int, rays: DeviceArray, n: int,
                  t_out: DeviceArray, id_out: DeviceArray) -> None:
    _check(ctx.lib.ptgpu_trace_closest(ctx.handle, _h(bvh), _h(kd), shapes.ptr, shapeCnt, rays.ptr, n, t_out.ptr,
                                       id_out.ptr))


def trace_any(ctx: Context, bvh, kd, shapes: DeviceArray, shapeCnt: int, rays: DeviceArray, maxt: DeviceArray,
              n: int, hit_out: DeviceArray) -> None:
    _check(ctx.lib.ptgpu_trace_any(ctx.handle, _h(bvh), _h(kd), shapes.ptr, shapeCnt, rays.ptr, maxt.ptr, n,
                                   hit_out.ptr))
