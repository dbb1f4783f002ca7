// device_util.hpp -- small device helpers shared by the kernels (wave64, gfx950).
#pragma once

#include "common.hpp"

namespace ptgpu {

constexpr int kWave = 64;

// The reference's select forms (AccelCommon.h:24-32, BVH.cl:78-86): (a<b)?a:b keeps b on ties and
// NaN, so they are written out instead of v_min/v_max to keep signed-zero and NaN behaviour.
__device__ __forceinline__ float sel_min(float a, float b) { return (a < b) ? a : b; }
__device__ __forceinline__ float sel_max(float a, float b) { return (a > b) ? a : b; }

__device__ __forceinline__ Bound merge_bounds(const Bound &b1, const Bound &b2) {   // BVH.cl:112-122
    Bound b3;
    b3.min_x = sel_min(b1.min_x, b2.min_x); b3.max_x = sel_max(b1.max_x, b2.max_x);
    b3.min_y = sel_min(b1.min_y, b2.min_y); b3.max_y = sel_max(b1.max_y, b2.max_y);
    b3.min_z = sel_min(b1.min_z, b2.min_z); b3.max_z = sel_max(b1.max_z, b2.max_z);
    return b3;
}

__device__ __forceinline__ Bound load_bound(const Bound *p) {
    const float *f = reinterpret_cast<const float *>(p);
    Bound b;
    b.min_x = f[0]; b.max_x = f[1]; b.min_y = f[2]; b.max_y = f[3]; b.min_z = f[4]; b.max_z = f[5];
    return b;
}

__device__ __forceinline__ unsigned long long lanemask_lt() {
    const unsigned lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Device-scope relaxed atomics for the sort's decoupled look-back: each status word carries its own
// value, so no ordering with other memory is needed (an acquire / release at device scope would
// write back or invalidate the L2 on gfx950).
__device__ __forceinline__ unsigned ld_relaxed(const unsigned *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(unsigned *p, unsigned v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Fill kernel (used instead of hipMemsetAsync so that sequences stay graph-capture safe).
__global__ void k_fill_u32(unsigned *p, unsigned v, size_t n);
int fill_u32(hipStream_t s, void *p, unsigned v, size_t n_words);
int fill_bytes(hipStream_t s, void *p, unsigned char v, size_t n_bytes);

}  // namespace ptgpu
