// ctx.cpp -- context, memory, events, errors, layout self-check and UpdateCamera.
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>

#include "device_util.hpp"

namespace ptgpu {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return e == hipErrorOutOfMemory ? PTGPU_ERR_OOM : PTGPU_ERR_HIP;
}

int DevBuf::ensure(size_t need) {
    if (need <= bytes && p) return PTGPU_OK;
    release();
    size_t want = need < 256 ? 256 : need;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        bytes = 0;
        return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
    }
    bytes = want;
    return PTGPU_OK;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

void BatchChannel::release() {
    table.release();
    counter.release();
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr;
    pinned_bytes = 0;
    if (copied) (void)hipEventDestroy(copied);
    if (ready) (void)hipEventDestroy(ready);
    copied = ready = nullptr;
    for (int i = 0; i < 2; i++) {
        if (sub[i]) (void)hipStreamDestroy(sub[i]);
        if (join[i]) (void)hipEventDestroy(join[i]);
        sub[i] = nullptr;
        join[i] = nullptr;
    }
    if (fork) (void)hipEventDestroy(fork);
    fork = nullptr;
}

}  // namespace ptgpu

using namespace ptgpu;

extern "C" {

const char *ptgpu_last_error(void) { return g_err; }

int ptgpu_layout_sizes(int *shape, int *camera, int *bvhnode, int *kdnode, int *ray, int *bound) {
    if (shape) *shape = (int)sizeof(Shape);
    if (camera) *camera = (int)sizeof(Camera);
    if (bvhnode) *bvhnode = (int)sizeof(BVHNodeGPU);
    if (kdnode) *kdnode = (int)sizeof(KDNodeGPU);
    if (ray) *ray = (int)sizeof(Ray);
    if (bound) *bound = (int)sizeof(Bound);
    return PTGPU_OK;
}

int ptgpu_device_count(int *count) {
    PT_CHECK(count, PTGPU_ERR_INVALID, "count is NULL");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    *count = c;
    return PTGPU_OK;
}

int ptgpu_ctx_create(int device, void *stream, ptgpu_ctx **out) {
    PT_CHECK(out, PTGPU_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        set_error("no HIP device is visible");
        return PTGPU_ERR_NO_DEVICE;
    }
    PT_CHECK(device >= 0 && device < count, PTGPU_ERR_INVALID, "device %d out of range [0,%d)", device, count);
    PT_HIP(hipSetDevice(device));
    ptgpu_ctx *c = new ptgpu_ctx();
    c->device = device;
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            return hip_fail(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
        }
        c->own_stream = true;
    }
    for (auto &e : c->ev) {
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) {
            ptgpu_ctx_destroy(c);
            return hip_fail(r, "hipEventCreate", __FILE__, __LINE__);
        }
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cu_count = prop.multiProcessorCount;
    if (c->cu_count <= 0) c->cu_count = 256;
    // the allocation-free entry's scratch address on this device (checked again by that entry,
    // which reports a failed look-up; a context that never uses it works without it)
    if (ptgpu::noalloc_scratch_init(device) != PTGPU_OK) (void)hipGetLastError();   // keep no stale HIP error
    // status words of recorded renders (allocated here: a recording must not be the first to allocate)
    int s = c->rec_status.ensure(16);
    if (s == PTGPU_OK) s = ptgpu::fill_u32(c->stream, c->rec_status.p, 0u, 4);
    if (s != PTGPU_OK) {
        ptgpu_ctx_destroy(c);
        return s;
    }
    *out = c;
    return PTGPU_OK;
}

int ptgpu_ctx_destroy(ptgpu_ctx *ctx) {
    if (!ctx) return PTGPU_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->tmp_nodes.release();
    ctx->tmp_nodes4.release();
    ctx->tmp_prims.release();
    ctx->tmp_lights.release();
    ctx->tmp_scan.release();
    ctx->tmp_counter.release();
    ctx->tmp_parents.release();
    ctx->tmp_tiles.release();
    ctx->rec_status.release();
    ptgpu::release_stages(ctx->rec_stages);
    ctx->batch.release();
    ctx->build_staging.release();
    wf_release(ctx);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PTGPU_OK;
}

int ptgpu_synchronize(ptgpu_ctx *ctx) {
    PT_CHECK(ctx, PTGPU_ERR_INVALID, "ctx is NULL");
    PT_NO_CAPTURE(ctx, "ptgpu_synchronize");
    PT_HIP(hipStreamSynchronize(ctx->stream));
    // what the device-side checks of recorded renders found since the last synchronisation
    unsigned w[2] = {0, 0};
    PT_HIP(hipMemcpyAsync(w, ctx->rec_status.p, 8, hipMemcpyDeviceToHost, ctx->stream));
    PT_HIP(hipStreamSynchronize(ctx->stream));
    if (w[0]) {
        PT_TRY(ptgpu::fill_u32(ctx->stream, ctx->rec_status.p, 0u, 2));
        PT_HIP(hipStreamSynchronize(ctx->stream));
        PT_CHECK(!(w[0] & 1u), PTGPU_ERR_HIP, "a recorded render found a failed LBVH build and rendered nothing");
        PT_CHECK(false, PTGPU_ERR_UNSUPPORTED,
                 "a recorded render found a tree of height %u, deeper than the traversal stack it was recorded "
                 "with, and rendered nothing (render the scene once outside the recording, then record again)",
                 w[1]);
    }
    return PTGPU_OK;
}

int ptgpu_ctx_set_engine(ptgpu_ctx *ctx, int engine) {
    PT_CHECK(ctx, PTGPU_ERR_INVALID, "ctx is NULL");
    PT_CHECK(engine >= 0 && engine <= 3, PTGPU_ERR_INVALID,
             "engine %d (0 auto, 1 lockstep megakernel, 2 wavefront, 3 per-lane-state megakernel)", engine);
    ctx->engine = engine;
    return PTGPU_OK;
}

int ptgpu_ctx_set_tuning(ptgpu_ctx *ctx, const char *name, int value) {
    PT_CHECK(ctx && name, PTGPU_ERR_INVALID, "null argument");
    if (!strcmp(name, "regroup")) {
        PT_CHECK(value >= 1 && value <= 64, PTGPU_ERR_INVALID, "regroup %d not in [1, 64]", value);
        ctx->regroup = value;
    } else if (!strcmp(name, "blocks_per_cu")) {
        PT_CHECK(value >= 1 && value <= 64, PTGPU_ERR_INVALID, "blocks_per_cu %d not in [1, 64]", value);
        ctx->blocks_per_cu = value;
    } else if (!strcmp(name, "batch_blocks_per_cu")) {
        PT_CHECK(value >= 1 && value <= 64, PTGPU_ERR_INVALID, "batch_blocks_per_cu %d not in [1, 64]", value);
        ctx->batch_blocks_per_cu = value;
    } else if (!strcmp(name, "batch_regroup")) {
        PT_CHECK(value >= 1 && value <= 64, PTGPU_ERR_INVALID, "batch_regroup %d not in [1, 64]", value);
        ctx->batch_regroup = value;
    } else if (!strcmp(name, "quad")) {
        PT_CHECK(value >= 0 && value <= 16, PTGPU_ERR_INVALID, "quad %d not in [0, 16]", value);
        ctx->quad = value;
    } else if (!strcmp(name, "wide")) {
        PT_CHECK(value >= 0 && value <= 2, PTGPU_ERR_INVALID, "wide %d not in {0, 1, 2}", value);
        ctx->wide = value;
    } else if (!strcmp(name, "short_stack")) {
        PT_CHECK(value >= 0 && value <= 2, PTGPU_ERR_INVALID, "short_stack %d not in {0, 1, 2}", value);
        ctx->short_stack = value;
    } else if (!strcmp(name, "stack")) {
        PT_CHECK(value >= 0 && value <= 64, PTGPU_ERR_INVALID, "stack %d not in [0, 64]", value);
        ctx->stack_min = value;
    } else {
        PT_CHECK(false, PTGPU_ERR_INVALID, "unknown tuning parameter '%s'", name);
    }
    return PTGPU_OK;
}

int ptgpu_ctx_set_max_depth(ptgpu_ctx *ctx, int max_depth) {
    PT_CHECK(ctx, PTGPU_ERR_INVALID, "ctx is NULL");
    PT_CHECK(max_depth >= 0 && max_depth < 1024, PTGPU_ERR_INVALID, "max_depth %d out of range", max_depth);
    ctx->max_depth = max_depth;
    return PTGPU_OK;
}

int ptgpu_malloc(ptgpu_ctx *ctx, size_t bytes, void **d_out) {
    PT_CHECK(ctx && d_out, PTGPU_ERR_INVALID, "null argument");
    *d_out = nullptr;
    PT_HIP(hipSetDevice(ctx->device));
    PT_HIP(hipMalloc(d_out, bytes ? bytes : 4));
    return PTGPU_OK;
}

int ptgpu_free(ptgpu_ctx *ctx, void *d_ptr) {
    PT_CHECK(ctx, PTGPU_ERR_INVALID, "ctx is NULL");
    if (!d_ptr) return PTGPU_OK;
    PT_NO_CAPTURE(ctx, "ptgpu_free");
    PT_HIP(hipStreamSynchronize(ctx->stream));
    PT_HIP(hipFree(d_ptr));
    return PTGPU_OK;
}

int ptgpu_memcpy_h2d(ptgpu_ctx *ctx, void *d_dst, const void *h_src, size_t bytes) {
    PT_CHECK(ctx && d_dst && h_src, PTGPU_ERR_INVALID, "null argument");
    PT_NO_CAPTURE(ctx, "ptgpu_memcpy_h2d (use ptgpu_memcpy_async from pinned memory)");
    PT_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    PT_HIP(hipStreamSynchronize(ctx->stream));
    return PTGPU_OK;
}

int ptgpu_memcpy_d2h(ptgpu_ctx *ctx, void *h_dst, const void *d_src, size_t bytes) {
    PT_CHECK(ctx && h_dst && d_src, PTGPU_ERR_INVALID, "null argument");
    PT_NO_CAPTURE(ctx, "ptgpu_memcpy_d2h (use ptgpu_memcpy_async into pinned memory)");
    PT_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    PT_HIP(hipStreamSynchronize(ctx->stream));
    return PTGPU_OK;
}

int ptgpu_memset_d(ptgpu_ctx *ctx, void *d_dst, int value, size_t bytes) {
    PT_CHECK(ctx && d_dst, PTGPU_ERR_INVALID, "null argument");
    // a fill kernel, not hipMemsetAsync, so that the call can be recorded into a graph
    return ptgpu::fill_bytes(ctx->stream, d_dst, (unsigned char)value, bytes);
}

int ptgpu_memcpy_d2d(ptgpu_ctx *ctx, void *d_dst, const void *d_src, size_t bytes) {
    PT_CHECK(ctx && d_dst && d_src, PTGPU_ERR_INVALID, "null argument");
    PT_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return PTGPU_OK;
}

void ptgpu_host_free(void *h_ptr) { free(h_ptr); }

int ptgpu_event_record(ptgpu_ctx *ctx, int slot) {
    PT_CHECK(ctx && slot >= 0 && slot < 16, PTGPU_ERR_INVALID, "bad event slot %d", slot);
    PT_HIP(hipEventRecord(ctx->ev[slot], ctx->stream));
    return PTGPU_OK;
}

int ptgpu_ctx_wait_event(ptgpu_ctx *waiter, ptgpu_ctx *signaller, int slot) {
    PT_CHECK(waiter && signaller && slot >= 0 && slot < 16, PTGPU_ERR_INVALID, "bad argument");
    PT_HIP(hipStreamWaitEvent(waiter->stream, signaller->ev[slot], 0));
    return PTGPU_OK;
}

int ptgpu_event_elapsed_ms(ptgpu_ctx *ctx, int a, int b, float *ms) {
    PT_CHECK(ctx && ms && a >= 0 && a < 16 && b >= 0 && b < 16, PTGPU_ERR_INVALID, "bad event slots");
    PT_HIP(hipEventSynchronize(ctx->ev[b]));
    PT_HIP(hipEventElapsedTime(ms, ctx->ev[a], ctx->ev[b]));
    return PTGPU_OK;
}

// UpdateCamera, displayfunc.cpp:732-756 (host; the kernels receive the finished basis).
int ptgpu_update_camera(Camera *cam) {
    PT_CHECK(cam, PTGPU_ERR_INVALID, "cam is NULL");
    PT_CHECK(cam->height != 0.f, PTGPU_ERR_INVALID, "camera height is 0 (set width/height first)");
    float xd = sinf(cam->yaw) * cosf(cam->pitch);
    float yd = sinf(cam->pitch);
    float zd = cosf(cam->yaw) * cosf(cam->pitch);
    Vec dir = {-1.f * xd, -1.f * yd, -1.f * zd};
    float l = 1.f / sqrtf(dir.x * dir.x + dir.y * dir.y + dir.z * dir.z);
    cam->dir = {l * dir.x, l * dir.y, l * dir.z};
    const Vec up = {0.f, 1.f, 0.f};
    const float fov = (kPi / 180.f) * 45.f;
    Vec x = {cam->dir.y * up.z - cam->dir.z * up.y, cam->dir.z * up.x - cam->dir.x * up.z,
             cam->dir.x * up.y - cam->dir.y * up.x};
    l = 1.f / sqrtf(x.x * x.x + x.y * x.y + x.z * x.z);
    x = {l * x.x, l * x.y, l * x.z};
    float k = cam->width * fov / cam->height;
    cam->x = {k * x.x, k * x.y, k * x.z};
    Vec y = {cam->x.y * cam->dir.z - cam->x.z * cam->dir.y, cam->x.z * cam->dir.x - cam->x.x * cam->dir.z,
             cam->x.x * cam->dir.y - cam->x.y * cam->dir.x};
    l = 1.f / sqrtf(y.x * y.x + y.y * y.y + y.z * y.z);
    y = {l * y.x, l * y.y, l * y.z};
    cam->y = {fov * y.x, fov * y.y, fov * y.z};
    return PTGPU_OK;
}

}  // extern "C"
