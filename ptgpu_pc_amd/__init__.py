"""ptgpu_pc_amd -- gfx950 (MI355X) implementation of PTGPU_PC's ray-acceleration and radiance path.

The compute lives in libptgpu.so (hand-written HIP kernels behind the C ABI in include/ptgpu.h);
this package is the host-side mirror of the reference interface plus scene generators.
"""
from .layout import (BVHNODE_DTYPE, CAMERA_DTYPE, DIFF, KDNODE_DTYPE, RAY_DTYPE, REFR, SHAPE_DTYPE, SPEC,
                     SPHERE, TRIANGLE, default_seeds, empty_shapes, make_camera)
from .lib import (ACCEL_BVH, ACCEL_KD, ACCEL_NONE, CLBVH, Context, DeviceArray, Hybrid, KDTree, PtgpuError, Server,
                  autotune, autotune_batch, device_count, host_radiance_box, load_library, radiance_gpu, read_scene,
                  render, render_batch, render_tiles,
                  trace_any, trace_closest, update_camera, write_ppm)

__all__ = [
    "BVHNODE_DTYPE", "CAMERA_DTYPE", "DIFF", "KDNODE_DTYPE", "RAY_DTYPE", "REFR", "SHAPE_DTYPE", "SPEC", "SPHERE",
    "TRIANGLE", "default_seeds", "empty_shapes", "make_camera", "ACCEL_BVH", "ACCEL_KD", "ACCEL_NONE", "CLBVH",
    "Context", "DeviceArray", "Hybrid", "KDTree", "PtgpuError", "Server", "autotune", "autotune_batch", "device_count",
    "host_radiance_box", "load_library", "radiance_gpu", "read_scene", "render", "render_batch", "render_tiles", "trace_any", "trace_closest",
    "update_camera", "write_ppm",
]
