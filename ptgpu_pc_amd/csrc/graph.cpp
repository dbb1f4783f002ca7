// graph.cpp -- HIP graphs of a per-frame call sequence (no reference counterpart).
//
// ptgpu_graph_begin starts stream capture on the context's stream; the library calls that follow
// are recorded instead of run, and ptgpu_graph_end instantiates the recording.  ptgpu_graph_launch
// replays the whole sequence as one submission.  A replay re-runs the recorded kernels with their
// recorded arguments and copies, so per-frame inputs that change (camera, seeds) travel through
// device buffers that the sequence itself fills, e.g. with ptgpu_memcpy_async from pinned host
// memory that the caller rewrites between replays (after the previous replay has finished).
// Calls that wait for the device refuse to be recorded (PTGPU_ERR_UNSUPPORTED); every fill in the
// recordable calls is a kernel, never a memset node.
#include "common.hpp"

#include <cstring>

struct ptgpu_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int device = 0;
    std::vector<ptgpu::RecordedStage> stages;   // tables the recorded copies read, freed with the graph
};

namespace ptgpu {

void release_stages(std::vector<RecordedStage> &v) {
    for (RecordedStage &r : v) {
        if (r.dev) (void)hipFree(r.dev);
        if (r.host) (void)hipHostFree(r.host);
    }
    v.clear();
}

int record_buffer(ptgpu_ctx *ctx, size_t bytes, void **d_out) {
    RecordedStage r;
    PT_HIP(hipMalloc(&r.dev, bytes ? bytes : 4));
    ctx->rec_stages.push_back(r);
    *d_out = r.dev;
    return PTGPU_OK;
}

int record_copy(ptgpu_ctx *ctx, hipStream_t stream, void *d_dst, const void *src, size_t bytes) {
    if (!bytes) return PTGPU_OK;
    RecordedStage r;
    PT_HIP(hipHostMalloc(&r.host, bytes, hipHostMallocDefault));
    ctx->rec_stages.push_back(r);
    std::memcpy(r.host, src, bytes);
    PT_HIP(hipMemcpyAsync(d_dst, r.host, bytes, hipMemcpyHostToDevice, stream));
    return PTGPU_OK;
}

int record_stage(ptgpu_ctx *ctx, hipStream_t stream, const void *src, size_t copy_bytes, size_t alloc_bytes,
                 void **d_out) {
    PT_CHECK(copy_bytes <= alloc_bytes, PTGPU_ERR_INVALID, "staged table larger than its buffer");
    PT_TRY(record_buffer(ctx, alloc_bytes, d_out));
    return record_copy(ctx, stream, *d_out, src, copy_bytes);
}

}  // namespace ptgpu

using namespace ptgpu;

extern "C" {

int ptgpu_graph_begin(ptgpu_ctx *ctx) {
    PT_CHECK(ctx, PTGPU_ERR_INVALID, "ctx is NULL");
    PT_CHECK(!ctx->capturing, PTGPU_ERR_INVALID, "a graph is already being recorded on this context");
    PT_HIP(hipSetDevice(ctx->device));
    PT_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    ctx->capturing = true;
    return PTGPU_OK;
}

int ptgpu_graph_end(ptgpu_ctx *ctx, ptgpu_graph **out) {
    PT_CHECK(ctx && out, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(ctx->capturing, PTGPU_ERR_INVALID, "ptgpu_graph_begin was not called");
    *out = nullptr;
    ctx->capturing = false;
    std::vector<RecordedStage> stages;
    stages.swap(ctx->rec_stages);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
    if (ec != hipSuccess || !g) {
        (void)hipStreamSynchronize(ctx->stream);
        release_stages(stages);
        if (ec != hipSuccess) return hip_fail(ec, "hipStreamEndCapture", __FILE__, __LINE__);
        PT_CHECK(false, PTGPU_ERR_HIP, "the recording was invalidated (a call waited for the device)");
    }
    hipGraphExec_t e = nullptr;
    const hipError_t r = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    if (r != hipSuccess) {
        (void)hipGraphDestroy(g);
        release_stages(stages);
        return hip_fail(r, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    ptgpu_graph *pg = new ptgpu_graph();
    pg->graph = g;
    pg->exec = e;
    pg->device = ctx->device;
    pg->stages.swap(stages);
    *out = pg;
    return PTGPU_OK;
}

int ptgpu_graph_launch(ptgpu_ctx *ctx, ptgpu_graph *g) {
    PT_CHECK(ctx && g, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(!ctx->capturing, PTGPU_ERR_INVALID, "cannot replay a graph while recording one");
    PT_CHECK(g->device == ctx->device, PTGPU_ERR_INVALID, "graph recorded on device %d, context on %d", g->device,
             ctx->device);
    PT_HIP(hipGraphLaunch(g->exec, ctx->stream));
    return PTGPU_OK;
}

int ptgpu_graph_destroy(ptgpu_graph *g) {
    if (!g) return PTGPU_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (!g->stages.empty()) {
        (void)hipSetDevice(g->device);
        (void)hipDeviceSynchronize();   // a replay still running reads these tables
        release_stages(g->stages);
    }
    delete g;
    return PTGPU_OK;
}

int ptgpu_memcpy_async(ptgpu_ctx *ctx, void *dst, const void *src, size_t bytes) {
    PT_CHECK(ctx && dst && src, PTGPU_ERR_INVALID, "null argument");
    PT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream));
    return PTGPU_OK;
}

int ptgpu_pinned_alloc(ptgpu_ctx *ctx, size_t bytes, void **h_out) {
    PT_CHECK(ctx && h_out, PTGPU_ERR_INVALID, "null argument");
    *h_out = nullptr;
    PT_HIP(hipSetDevice(ctx->device));
    PT_HIP(hipHostMalloc(h_out, bytes ? bytes : 4, hipHostMallocDefault));
    return PTGPU_OK;
}

int ptgpu_pinned_free(void *h_ptr) {
    if (!h_ptr) return PTGPU_OK;
    PT_HIP(hipHostFree(h_ptr));
    return PTGPU_OK;
}

}  // extern "C"
