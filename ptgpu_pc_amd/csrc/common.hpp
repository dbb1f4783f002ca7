// common.hpp -- shared definitions for the gfx950 ray-acceleration and radiance library.
//
// Device layouts used by the traversal engine (derived from the reference layouts at build time):
//   TNode (64 B)  one internal node of a binary tree with BOTH child boxes stored in the parent, so
//                 one 64-byte fetch decides which children a ray enters.
//   Prim  (48 B)  one primitive in traversal (leaf) order: triangle v0, e1 = v1-v0, e2 = v2-v0
//                 (the same FP32 differences TriangleIntersect forms at rendering_kernel.cl:72-74),
//                 or sphere centre + radius; the original shape index rides in p0.w.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/ptgpu.h"

// ---- layout checks against the reference declarations (offsets computed from geom.h etc.) ----
static_assert(sizeof(Vec) == 12, "Vec");
static_assert(sizeof(Ray) == 24, "Ray");
static_assert(sizeof(Bound) == 24, "Bound");
static_assert(sizeof(Shape) == 104, "Shape");
static_assert(offsetof(Shape, t) == 4 && offsetof(Shape, e) == 40 && offsetof(Shape, c) == 52, "Shape");
static_assert(offsetof(Shape, refl) == 64 && offsetof(Shape, index) == 68, "Shape");
static_assert(offsetof(Shape, morton_code) == 72 && offsetof(Shape, b) == 76, "Shape");
static_assert(offsetof(Shape, area) == 100, "Shape");
static_assert(sizeof(Camera) == 76 && offsetof(Camera, dir) == 40 && offsetof(Camera, y) == 64, "Camera");
static_assert(sizeof(BVHNodeGPU) == 60 && offsetof(BVHNodeGPU, bound) == 28, "BVHNodeGPU");
static_assert(offsetof(BVHNodeGPU, cost) == 52 && offsetof(BVHNodeGPU, area) == 56, "BVHNodeGPU");
static_assert(sizeof(KDNodeGPU) == 36 && offsetof(KDNodeGPU, bound) == 8, "KDNodeGPU");
static_assert(offsetof(KDNodeGPU, min) == 32 && offsetof(KDNodeGPU, max) == 34, "KDNodeGPU");

namespace ptgpu {

constexpr float kEpsilon = 0.01f;                       // geom.h:29
constexpr float kPi = 3.14159265358979323846f;          // geom.h:30

struct alignas(16) TNode {
    float4 a;   // c0.min_x c0.min_y c0.min_z c0.max_x
    float4 b;   // c0.max_y c0.max_z c1.min_x c1.min_y
    float4 c;   // c1.min_z c1.max_x c1.max_y c1.max_z
    int4 r;     // ref0 ref1 cnt0 cnt1 : ref >= 0 internal node, ref < 0 leaf with first prim ~ref
};
static_assert(sizeof(TNode) == 64, "TNode");

// Four-wide traversal node, one per internal TNode: the children of binary node i's children
// (a leaf child stays a leaf), so a walk from the root that always jumps two binary levels
// touches only every other level.  Boxes are stored per axis, child k in component k.
struct alignas(16) TNode4 {
    float4 lox, loy, loz;
    float4 hix, hiy, hiz;
    int4 ref;   // >= 0 internal (TNode4 index), < 0 leaf with first prim ~ref, kEmptyRef unused
    int4 cnt;   // leaf primitive counts
};
static_assert(sizeof(TNode4) == 128, "TNode4");
constexpr int kEmptyRef = (int)0x80000000;
constexpr float kFarBox = 1e30f;   // coordinates of an unused TNode4 slot

// The same four-wide node in half the bytes (four 16-byte loads instead of eight): each child box
// as binary16 bounds rounded outward (lo toward -inf, hi toward +inf), so it contains the fp32 box
// and every ray that hits the fp32 box hits it too; the walk may visit a few more nodes, never
// fewer, and the hits are the same.  One 16-byte word per child k: (x, y, z, ref), each axis as
// lo in the low half and hi in the high half, so one lane can fetch one child in one load.  Leaves
// are runs of primitives ending in kLastInLeaf (trees with leaf_runs only).  An unused slot holds
// +inf bounds, which no ray hits.
struct alignas(16) TNode4H {
    uint4 c[4];
};
static_assert(sizeof(TNode4H) == 64, "TNode4H");

struct alignas(16) Prim {
    float4 p0;  // v0.xyz | centre.xyz, w = shape index (bits)
    float4 p1;  // e1.xyz | (rad, 0, 0), w = type (bits 0-7) | kLastInLeaf
    float4 p2;  // e2.xyz, w = type again (the test branches on it, so all three loads issue together)
};
static_assert(sizeof(Prim) == 48, "Prim");
// Set in Prim::p1.w on the last primitive of each leaf range when the leaf ranges are disjoint, so
// a traversal can hold a whole leaf as the index of its first primitive.
constexpr int kLastInLeaf = 0x100;

// A prepared traversal structure on the device.
struct DevAccel {
    const TNode *nodes = nullptr;
    const TNode4 *nodes4 = nullptr;   // optional four-wide view of `nodes` (same indices)
    const TNode4H *nodes4h = nullptr; // optional binary16 copy of nodes4 (leaf_runs trees only)
    const Prim *prims = nullptr;
    int root = 0;          // >= 0 internal node; < 0: the whole scene is one leaf ~root
    int root_cnt = 0;      // primitive count when the root is a leaf
    int nprims = 0;
    int nnodes = 0;        // number of TNodes (node refs are checked against it)
    int nshapes = 0;       // shapeCnt: any-hit queries ignore shape nshapes-1 (IntersectP, Q10)
    int depth = 0;         // tree height: deepest leaf in edges from the root (bounds the stack)
    int mode = PTGPU_ACCEL_NONE;
    int leaf_runs = 0;     // 1: every leaf range is disjoint and its last Prim carries kLastInLeaf
    // A traversal recorded into a graph right after a build, before its height could be read back:
    // `depth` is then a bound, and the build's status words (error word, -, height) are checked
    // against it on the device before any traversal.  NULL when `depth` is the tree's height.
    const unsigned *height_check = nullptr;
};

// ---- errors ----
void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define PT_HIP(expr)                                                                   \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) return ::ptgpu::hip_fail(_e, #expr, __FILE__, __LINE__); \
    } while (0)

#define PT_CHECK(cond, code, ...)            \
    do {                                     \
        if (!(cond)) {                       \
            ::ptgpu::set_error(__VA_ARGS__); \
            return (code);                   \
        }                                    \
    } while (0)

// Refuses a call that waits for the device while the context's stream is being captured into a graph.
#define PT_NO_CAPTURE(ctx, what)                                                                       \
    PT_CHECK(!(ctx)->capturing, PTGPU_ERR_UNSUPPORTED,                                                 \
             "%s waits for the device and cannot be recorded into a graph (prepare it before "        \
             "ptgpu_graph_begin)", what)

#define PT_TRY(expr)                  \
    do {                              \
        int _s = (expr);              \
        if (_s != PTGPU_OK) return _s; \
    } while (0)

// Growable device buffer owned by a handle.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
    template <class T> T *as() const { return static_cast<T *>(p); }
};

}  // namespace ptgpu

namespace ptgpu {
// One stream's worth of batched-render launches: the device job table, its pinned staging copy (reused
// once the previous copy is done), and the launch's work counter.
struct BatchChannel {
    hipStream_t stream = nullptr;
    DevBuf table, counter;
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    hipEvent_t copied = nullptr;   // the last table upload from `pinned` has finished
    hipEvent_t ready = nullptr;    // scene preparation on the context's stream, waited on by `stream`
    // jobs are rendered in one launch per traversal-stack class; the classes after the first run on
    // these streams, forked from and joined back into `stream` (default priority, created on first use)
    hipStream_t sub[2] = {nullptr, nullptr};
    hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
    void release();
};
}  // namespace ptgpu

namespace ptgpu {
// A table uploaded by a call that is being recorded into a graph: a pinned host copy and a device
// buffer owned by the recording (then by its graph), so that every replay copies the table as it was
// recorded whatever eager calls do with the reusable staging buffers in between.
struct RecordedStage {
    void *host = nullptr;
    void *dev = nullptr;
};
void release_stages(std::vector<RecordedStage> &v);
}  // namespace ptgpu

namespace ptgpu {
// Tables of a batched LBVH build (per-scene buffers, block offsets, clears) and their pinned
// upload copy; reused by the next batch once the previous one has been waited for.
struct BuildStaging {
    DevBuf table;
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    void release();
};
}  // namespace ptgpu

struct ptgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t ev[16] = {};
    int cu_count = 0;
    int max_depth = 6;   // MAX_DEPTH (geom.h:32) for ptgpu_radiance_gpu
    int engine = 0;      // 0 auto, 1 megakernel (lockstep queries), 2 wavefront, 3 megakernel (per-lane states)
    int regroup = 16;   // per-lane-state megakernel: ready lanes that trigger a shading pass
    int quad = 16;      // per-lane-state megakernel, binary16 nodes: traversing lanes at or below which four lanes share a query
    int batch_regroup = 32;
    int blocks_per_cu = 2;        // single-frame renders: few waves shorten each pixel's chain
    int batch_blocks_per_cu = 3;  // batched renders: throughput-bound, more waves per SIMD
    int wide = 2;        // per-lane-state megakernel: 0 binary nodes, 1 four-wide, 2 four-wide binary16 (when they exist)
    int stack_min = 0;
    int short_stack = 1;  // 48-entry binary16 walks on the 36-entry class (0 off; 2 every binary16 walk on 8 entries, a test)   // smallest traversal stack class a render or query launch uses (0: the tree's height decides)
    void *wf_workspace = nullptr;
    // scratch for raw-layout renders and light lists
    ptgpu::DevBuf tmp_nodes, tmp_nodes4, tmp_prims, tmp_lights, tmp_scan, tmp_counter;
    ptgpu::DevBuf tmp_parents;           // parent map and height of a caller's BVH (ptgpu_radiance_gpu)
    ptgpu::DevBuf tmp_tiles;             // tile origins of RadianceBoxGPU launches
    ptgpu::DevBuf rec_status;            // recorded renders' device-side checks: flags, height seen
    ptgpu::BatchChannel batch;            // ptgpu_render_batch launches (on the context's stream)
    ptgpu::BuildStaging build_staging;    // batched LBVH builds
    bool capturing = false;               // between ptgpu_graph_begin and ptgpu_graph_end
    std::vector<ptgpu::RecordedStage> rec_stages;   // tables staged by the calls being recorded
};

namespace ptgpu {
// Staging for a call being recorded on `stream`, owned by the recording (then by its graph):
// record_buffer allocates device memory; record_copy keeps a pinned copy of `src` and records its
// upload into d_dst, so every replay copies what was recorded; record_stage does both.
int record_buffer(ptgpu_ctx *ctx, size_t bytes, void **d_out);
int record_copy(ptgpu_ctx *ctx, hipStream_t stream, void *d_dst, const void *src, size_t bytes);
int record_stage(ptgpu_ctx *ctx, hipStream_t stream, const void *src, size_t copy_bytes, size_t alloc_bytes,
                 void **d_out);
}  // namespace ptgpu

namespace ptgpu {
// light list: indices of the first `light_cnt` shapes with non-zero emission, in index order
// (the scan order of SampleLights, rendering_kernel.cl:415-423).  Writes *found on the host.
int build_light_list(ptgpu_ctx *ctx, const Shape *d_shapes, int n, int light_cnt, DevBuf &out_lights,
                     DevBuf &scan_tmp, int *found);
// brute-force layout: prims in shape-index order
int build_index_prims(ptgpu_ctx *ctx, const Shape *d_shapes, int n, DevBuf &out_prims);
// four-wide nodes for a prepared binary structure (no-op without internal nodes); sets acc->nodes4
int build_wide(ptgpu_ctx *ctx, DevAccel *acc, DevBuf &out_nodes4);
// batched render on a channel: jobs resolved on the context's stream, launched on the channel's
int launch_batch(ptgpu_ctx *ctx, BatchChannel &ch, const ptgpu_render_job *jobs, int njobs,
                 unsigned long long *d_ray_counter);
// light list cached in a KD handle (as the LBVH handle caches its own)
int kdtree_lights(ptgpu_kdtree *kd, const Shape *d_shapes, int n, int light_cnt, const int **lights, int *nlights);
// the same with caller-placed buffers (no allocation)
int build_index_prims_into(ptgpu_ctx *ctx, const Shape *d_shapes, int n, Prim *prims);
size_t wide_bytes(const DevAccel &acc);
int build_wide_into(ptgpu_ctx *ctx, DevAccel *acc, TNode4 *w4);   // room for wide_bytes(*acc)
size_t light_scan_bytes(int n);
int build_light_list_into(ptgpu_ctx *ctx, const Shape *d_shapes, int n, int light_cnt, int *out, unsigned *scan,
                          int *found);
size_t ref_parents_bytes(int n);
int convert_bvh_ref_into(ptgpu_ctx *ctx, const Shape *d_shapes, int n, const BVHNodeGPU *d_btn,
                         const BVHNodeGPU *d_btl, TNode *nodes, Prim *prims, int *parents, DevAccel *acc);
int convert_kd_ref_into(ptgpu_ctx *ctx, const Shape *d_shapes, int n, const KDNodeGPU *d_kng, int kngCnt,
                        const int *d_kn, int knCnt, TNode *nodes, Prim *prims, DevAccel *acc);
// reference-layout trees -> traversal layout
int convert_bvh_ref(ptgpu_ctx *ctx, const Shape *d_shapes, int n, const BVHNodeGPU *d_btn,
                    const BVHNodeGPU *d_btl, DevBuf &nodes, DevBuf &prims, DevAccel *acc);
int convert_kd_ref(ptgpu_ctx *ctx, const Shape *d_shapes, int n, const KDNodeGPU *d_kng, int kngCnt,
                   const int *d_kn, int knCnt, DevBuf &nodes, DevBuf &prims, DevAccel *acc);
// render / trace over a prepared structure
struct RenderLaunch {
    Vec *colors;
    unsigned int *seeds;
    const Camera *camera;
    const Shape *shapes;
    unsigned int shapeCnt;
    const int *lights;
    int nlights;
    int width, height, currentSample, spp, max_depth;
    unsigned int *pixels;
    unsigned long long *ray_counter;
    unsigned long long *trav_stats;
    unsigned *work_counter = nullptr;   // device word for the persistent waves' pixel counter; NULL: the context's
    const int2 *tiles = nullptr;        // device tile origins: render only these tile_w x tile_h tiles
    int ntiles = 0, tile_w = 0, tile_h = 0;
    bool no_alloc = false;   // the allocation-free entry: kernels that need no scratch memory only
};
int launch_render(ptgpu_ctx *ctx, const DevAccel &acc, const RenderLaunch &L);
// ptgpu_render's checks and preparation (traversal structure, light list) without the launch
int prepare_render(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const ptgpu_render_args *a, DevAccel *acc,
                   RenderLaunch *L);
// ptgpu_render's preparation, then a launch over the listed tiles (device int2 origins, tile_w x tile_h
// each); acc and L receive the resolved structure and launch description (host tiles use them too)
int render_tiles(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const ptgpu_render_args *a, const int2 *d_tiles,
                 int ntiles, int tile_w, int tile_h, DevAccel *acc, RenderLaunch *L);
// device address of this GPU's allocation-free scratch, looked up once (device current)
int noalloc_scratch_init(int device);
int launch_render_wavefront(ptgpu_ctx *ctx, const DevAccel &acc, const RenderLaunch &L);
int wf_supported(int nlights);
void wf_release(ptgpu_ctx *ctx);
int launch_trace(ptgpu_ctx *ctx, const DevAccel &acc, unsigned int shapeCnt, const Ray *d_rays,
                 const float *d_maxt, int n, float *d_t, int *d_id, int *d_hit, bool any);
}  // namespace ptgpu
