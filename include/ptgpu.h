/*
 * ptgpu.h -- C ABI of the MI355X (gfx950) ray-acceleration and radiance library.
 *
 * Every structure below keeps the byte layout of the PTGPU_PC reference so that buffers produced by
 * the reference's host code (Shape arrays, Camera, BVHNodeGPU/KDNodeGPU trees, seed and colour
 * buffers) can be handed to this library unchanged.  Every entry point returns an int status
 * (PTGPU_OK == 0); nothing throws across the boundary.  ptgpu_last_error() returns a readable
 * message for the last failure on the calling thread.
 *
 * Pointer conventions: arguments named d_* are device (HBM) pointers, h_* are host pointers.
 * Work is ordered on the context's HIP stream; calls that return host data synchronise it.
 * A context is single-threaded: one host thread drives one context at a time.
 */
#ifndef PTGPU_H
#define PTGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------------------------ */
/* Layout types (bit-identical to the reference).                                              */
/* ------------------------------------------------------------------------------------------ */

/* vec.h:27-29 */
typedef struct { float x, y, z; } Vec;
/* geom.h:48-50 */
typedef struct { Vec o, d; } Ray;
/* geom.h:55-57 */
enum Refl { DIFF = 0, SPEC = 1, REFR = 2 };
/* geom.h:59-62 */
typedef struct { float rad; Vec p; } Sphere;
/* geom.h:68-70 */
typedef struct { Vec p1, p2, p3; } Triangle;
/* geom.h:72 */
typedef enum tagType { SPHERE = 0, TRIANGLE = 1 } Type;
/* geom.h:73 -- min and max are interleaved per axis */
typedef struct { float min_x, max_x, min_y, max_y, min_z, max_z; } Bound;
/* geom.h:75-91 -- 104 bytes */
typedef struct {
    Type type;
    union { Triangle t; Sphere s; };
    Vec e, c;            /* emission, colour */
    enum Refl refl;
    int index;
    unsigned int morton_code;
    Bound b;
    float area;
} Shape;
/* camera.h:29-38 -- 76 bytes */
typedef struct {
    Vec orig, target;
    float pitch, yaw;
    float width, height;
    Vec dir, x, y;
} Camera;
/* BVHNodeGPU.h:4-16 -- 60 bytes */
typedef struct BVHNodeGPU {
    int min, max;
    int nShape;
    int leaf;
    int nLeft, nRight, nParent;
    Bound bound;
    float cost;
    float area;
} BVHNodeGPU;
/* KDNodeGPU.h:4-14 -- 36 bytes (int16 indices) */
typedef struct KDNodeGPU {
    short leaf;
    short nLeft, nRight, nParent;
    Bound bound;
    short min;
    short max;
} KDNodeGPU;

/* ------------------------------------------------------------------------------------------ */
/* Status codes                                                                                */
/* ------------------------------------------------------------------------------------------ */
enum {
    PTGPU_OK = 0,
    PTGPU_ERR_INVALID = 1,      /* bad argument (null pointer, size out of range, wrong state) */
    PTGPU_ERR_HIP = 2,          /* a HIP runtime call failed */
    PTGPU_ERR_OOM = 3,          /* device allocation failed or admission budget exceeded */
    PTGPU_ERR_UNSUPPORTED = 4,  /* the request cannot be represented (e.g. int16 KD export overflow) */
    PTGPU_ERR_NO_DEVICE = 5     /* no HIP device is visible */
};

/* Acceleration modes: the reference's compile-time ACCELSTR (geom.h:35) as a runtime value. */
enum { PTGPU_ACCEL_NONE = 0, PTGPU_ACCEL_BVH = 1, PTGPU_ACCEL_KD = 2 };

typedef struct ptgpu_ctx ptgpu_ctx;
typedef struct ptgpu_clbvh ptgpu_clbvh;
typedef struct ptgpu_kdtree ptgpu_kdtree;

/* ------------------------------------------------------------------------------------------ */
/* Library, context, memory                                                                    */
/* ------------------------------------------------------------------------------------------ */

/* Size of each layout type as compiled into the library (layout self-check for FFI callers). */
int ptgpu_layout_sizes(int *shape, int *camera, int *bvhnode, int *kdnode, int *ray, int *bound);
const char *ptgpu_last_error(void);
int ptgpu_device_count(int *count);

/* Replaces the OpenCL context/queue pair made in SetUpOpenCL (smallptGPU.cpp:286-441).
 * `stream` may be NULL (the library creates a default-priority stream) or a hipStream_t. */
int ptgpu_ctx_create(int device, void *stream, ptgpu_ctx **out);
int ptgpu_ctx_destroy(ptgpu_ctx *ctx);
int ptgpu_synchronize(ptgpu_ctx *ctx);
/* MAX_DEPTH (geom.h:32) used by ptgpu_radiance_gpu: the reference fixes it at compile time (6);
 * here it is a per-context value, default 6. */
int ptgpu_ctx_set_max_depth(ptgpu_ctx *ctx, int max_depth);
/* Path-tracing schedule: 0 auto (3 when a tree is walked, else 1), 1 megakernel with wave-lockstep
 * queries, 2 wavefront (needs a tree and lightCnt <= 8), 3 megakernel with per-lane states (a lane
 * whose query ends waits for shading while the rest of its wave keeps traversing, over four-wide
 * nodes).  All give the same bits. */
int ptgpu_ctx_set_engine(ptgpu_ctx *ctx, int engine);
/* Launch tuning, no effect on results: "regroup" (1..64, ready lanes that start a shading pass in
 * engine 3, default 16), "blocks_per_cu" (1..64, persistent workgroups per CU, default 2), "wide"
 * (engine 3's node format: 0 binary, 1 four-wide fp32, 2 four-wide with binary16 boxes rounded
 * outward, half the bytes per node, where the tree's leaves allow it; default 2), "quad" (0..16,
 * with wide 2: when at most this many lanes of a wave are traversing, four lanes share each query,
 * one child box and leaf each; default 16, 0 off); "stack" (0..64, the smallest traversal stack
 * class a launch uses: the launch takes the first class (24 or 64 entries for binary nodes, 32, 48 or 64
 * for four-wide nodes) that holds both this many entries and the tree's height; default 0, the height
 * alone decides; a larger class only costs LDS, so this is a test knob for every kernel instance);
 * "short_stack" (0..2: with wide 2, a tree that needs the 48-entry stack is walked on a 36-entry one,
 * so four workgroups fit on a CU instead of three, and a walk that outgrows it is finished against
 * every primitive, with the same result; 1 the default, 0 off: about 2.6% less time for one frame
 * alone, 4% less throughput with frames in flight at C2; 2 walks every binary16 tree on an 8-entry
 * stack, a test setting);
 * "batch_regroup" and "batch_blocks_per_cu" (the same for ptgpu_render_batch, defaults 32 and 3). */
int ptgpu_ctx_set_tuning(ptgpu_ctx *ctx, const char *name, int value);
int ptgpu_malloc(ptgpu_ctx *ctx, size_t bytes, void **d_out);
int ptgpu_free(ptgpu_ctx *ctx, void *d_ptr);
int ptgpu_memcpy_h2d(ptgpu_ctx *ctx, void *d_dst, const void *h_src, size_t bytes); /* synchronous */
int ptgpu_memcpy_d2h(ptgpu_ctx *ctx, void *h_dst, const void *d_src, size_t bytes); /* synchronous */
int ptgpu_memset_d(ptgpu_ctx *ctx, void *d_dst, int value, size_t bytes);          /* stream-ordered (a fill kernel) */
int ptgpu_memcpy_d2d(ptgpu_ctx *ctx, void *d_dst, const void *d_src, size_t bytes); /* stream-ordered */
/* Host buffers returned by ptgpu_kdtree_get_trees are released with this. */
void ptgpu_host_free(void *h_ptr);

/* A plain device copy kernel (bytes a multiple of 16) with known traffic, for calibrating hardware
 * memory-traffic counters. */
int ptgpu_copy_calibration(ptgpu_ctx *ctx, void *d_dst, const void *d_src, size_t bytes);
/* Stream-ordered timing: record a named event slot (0..15) and read the elapsed ms between two. */
int ptgpu_event_record(ptgpu_ctx *ctx, int slot);
int ptgpu_event_elapsed_ms(ptgpu_ctx *ctx, int slot_begin, int slot_end, float *ms);
/* No reference counterpart.  Work enqueued on `waiter` after this call waits for event `slot` last
 * recorded on `signaller` (two contexts on one device run on two streams; e.g. build the next frame's
 * tree on one while the other renders).  A render on one context of a handle built on another
 * waits for the handle's build by itself; reusing a handle for a rebuild while another context
 * still renders with it needs this wait. */
int ptgpu_ctx_wait_event(ptgpu_ctx *waiter, ptgpu_ctx *signaller, int slot);

/* ------------------------------------------------------------------------------------------ */
/* LBVH builder -- replaces class CLBVH (CLBVH.h:39-47, CLBVH.cpp:80-366)                      */
/* ------------------------------------------------------------------------------------------ */

/* CLBVH::CLBVH (CLBVH.cpp:80-181).  Uploads h_shapes, computes per-shape AABBs, scene centroid
 * bounds and 30-bit Morton codes on the GPU, and -- as the reference does -- writes `index`,
 * `morton_code` and `b` back into the caller's h_shapes array.  n >= 1. */
int ptgpu_clbvh_create(ptgpu_ctx *ctx, Shape *h_shapes, int n, ptgpu_clbvh **out);
/* Same, with a device-resident Shape array (no PCIe traffic).  d_shapes is updated in place
 * (index, morton_code, b) and must stay valid for the lifetime of the handle. */
int ptgpu_clbvh_create_device(ptgpu_ctx *ctx, Shape *d_shapes, int n, ptgpu_clbvh **out);
/* Re-run the constructor's work (bounds + Morton) on the same device Shape array, e.g. after the
 * caller moved vertices; keeps allocations.  n must not exceed the handle's capacity. */
int ptgpu_clbvh_reset(ptgpu_clbvh *h, int n);
/* CLBVH::buildRadixTree (CLBVH.cpp:207-238; kernel BVH.cl:141-208). */
int ptgpu_clbvh_build_radix_tree(ptgpu_clbvh *h);
/* CLBVH::buildBVHTree (CLBVH.cpp:245-311; kernel BVH.cl:214-258): stable descending sort by Morton
 * key, leaf fill and bottom-up bound refit. */
int ptgpu_clbvh_build_bvh_tree(ptgpu_clbvh *h);
/* Batched builds (no single reference counterpart: the reference builds one CLBVH at a time,
 * smallptGPU.cpp:1462-1470).  For every handle, the same work and results as
 * ptgpu_clbvh_reset(h, n) + ptgpu_clbvh_build_radix_tree + ptgpu_clbvh_build_bvh_tree on the
 * handle's device shapes, with each stage run as one launch over all scenes; bit-identical to
 * building each scene alone.  Handles must be distinct, belong to ctx and share one sort digit
 * width.  Returns after the builds have finished (one synchronisation for the batch). */
int ptgpu_clbvh_build_batch(ptgpu_ctx *ctx, ptgpu_clbvh *const *handles, int count);
/* ptgpu_clbvh_create_device for `count` scenes followed by ptgpu_clbvh_build_batch: out[s] is a
 * handle over d_shapes[s] (counts[s] shapes, caller-owned, updated in place) with its trees built.
 * On failure every out[s] is NULL. */
int ptgpu_clbvh_create_batch(ptgpu_ctx *ctx, Shape *const *d_shapes, const int *counts, int count,
                             ptgpu_clbvh **out);
/* CLBVH::optimize (CLBVH.cpp:318-355; kernel BVH.cl:514-543): treelet restructuring with up to
 * `treelet_size` (2..7) leaves per treelet, SAH constants Ci=1.2, Ct=1.0 (clheader.h:14-15). */
int ptgpu_clbvh_optimize(ptgpu_clbvh *h, int treelet_size);
/* CLBVH::getTrees (CLBVH.cpp:357-366): blocking readback into host mirrors owned by the handle
 * (borrowed pointers, valid until the next build call or destroy). */
int ptgpu_clbvh_get_trees(ptgpu_clbvh *h, BVHNodeGPU **h_btn, BVHNodeGPU **h_btl);
/* Zero-copy access to the device trees (N-1 internal nodes, N leaves). */
int ptgpu_clbvh_device_trees(ptgpu_clbvh *h, const BVHNodeGPU **d_btn, const BVHNodeGPU **d_btl);
/* Device memory held by the handle (its own shapes copy included), in bytes. */
int ptgpu_clbvh_device_bytes(ptgpu_clbvh *h, size_t *bytes);
/* Device pointer to sorted_geometry_indices (CLBVH.cpp:257), N ints. */
int ptgpu_clbvh_device_sorted(ptgpu_clbvh *h, const int **d_sorted);
/* Tree used by the traversal layout (ptgpu_render / ptgpu_trace_*) after the next build_bvh_tree:
 * 0 = the exported index-keyed Karras tree (BVH.cl:127-134), 1 = the Morton-keyed LBVH over the same
 * sorted leaves (default; about half the SAH cost).  btn/btl are the index-keyed tree either way. */
int ptgpu_clbvh_set_traversal_topology(ptgpu_clbvh *h, int mode);
/* No reference counterpart.  readback 1 (default): the first traversal or treelet pass after a build
 * waits for the build and reads the new tree's height back to size the traversal stack.  readback 0:
 * once a height has been read for a build of the same kind (plain or optimised), later ones are not
 * waited for: the stack is sized for that height plus 4 (the bound recorded renders use) and the render
 * kernels check the built height on the device; a frame whose tree is deeper is not rendered and the
 * next ptgpu_synchronize on its context returns PTGPU_ERR_UNSUPPORTED.  For pipelines that rebuild
 * trees of a steady height frame after frame, so the host never waits for a build. */
int ptgpu_clbvh_set_height_readback(ptgpu_clbvh *h, int readback);
/* Read back the traversal tree (N-1 internal nodes): refs4[i] = {left ref, right ref, lo, hi} (ref < 0 is
 * leaf ~k; [lo, hi] the node's leaf range) and boxes12[i] = left child Bound, right child Bound; *height =
 * root height.  For tests and tools. */
int ptgpu_clbvh_export_traversal(ptgpu_clbvh *h, int *h_refs4, float *h_boxes12, int *height);
/* Radix-sort digit width for the next build_bvh_tree (4 or 8 bits; default 8). */
int ptgpu_clbvh_set_digit_bits(ptgpu_clbvh *h, int bits);
/* Per-stage device times of the last build in ms: morton, sort, topology, refit, optimize (refit
 * includes the Morton-keyed traversal tree).  A batched build leaves them at 0. */
int ptgpu_clbvh_stage_times(ptgpu_clbvh *h, float *ms5);
/* The same with the refit split: morton, sort, topology, refit of the exported tree, traversal tree,
 * optimize. */
int ptgpu_clbvh_stage_times6(ptgpu_clbvh *h, float *ms6);
int ptgpu_clbvh_destroy(ptgpu_clbvh *h);

/* ------------------------------------------------------------------------------------------ */
/* KD tree -- replaces class KDTree (KDTree.h:88-89, KDTree.cpp:89-270)                        */
/* ------------------------------------------------------------------------------------------ */

/* KDTree::build (KDTree.cpp:89-182): the reference's mean-split object partition, built on the
 * host with the reference's floating-point order; also prepares the device traversal layout.
 * ctx may be NULL: the tree is then built on the host only (exports work, rendering does not). */
int ptgpu_kdtree_build(ptgpu_ctx *ctx, const Shape *h_shapes, int n, ptgpu_kdtree **out);
/* SAH variant with a depth cap (extension; no reference counterpart in the mean-split builder). */
int ptgpu_kdtree_build_sah(ptgpu_ctx *ctx, const Shape *h_shapes, int n, int max_depth, int max_leaf,
                           ptgpu_kdtree **out);
/* KDTree::getTrees (KDTree.cpp:257-270): BFS heap export of 2^(maxdepth+1) KDNodeGPU plus the
 * primitive list.  Ownership of both buffers passes to the caller (free with ptgpu_host_free).
 * Returns PTGPU_ERR_UNSUPPORTED when the tree does not fit the int16 fields of KDNodeGPU. */
int ptgpu_kdtree_get_trees(ptgpu_kdtree *kd, KDNodeGPU **h_kng, int *kngCnt, int **h_kn, int *knCnt);
/* Depth-first export of the tree in int32 form: node5[k] = {leaf, left, right, start, count} and
 * box[k] for node k (pre-order, left before right); kn lists shape indices of the leaves in order. */
int ptgpu_kdtree_export_dfs(ptgpu_kdtree *kd, int *node5, Bound *box, int *kn, int cap_nodes, int cap_kn);
/* Tree statistics: max depth, number of nodes, number of leaves, primitive-list length. */
int ptgpu_kdtree_stats(ptgpu_kdtree *kd, int *max_depth, int *nodes, int *leaves, int *kn_len);
int ptgpu_kdtree_destroy(ptgpu_kdtree *kd);

/* ------------------------------------------------------------------------------------------ */
/* Rendering -- replaces RadianceGPU (rendering_kernel.cl:1072-1150) and DrawFrame             */
/* (smallptGPU.cpp:1010-1087)                                                                   */
/* ------------------------------------------------------------------------------------------ */

/* The RadianceGPU argument list in its original order (smallptGPU.cpp:1013-1031).  All pointers
 * are device pointers.  btn/btl are used when accel == PTGPU_ACCEL_BVH, kng/kngCnt/kn/knCnt when
 * accel == PTGPU_ACCEL_KD. */
typedef struct {
    Vec *colors;
    unsigned int *seeds;
    const Camera *camera;
    const Shape *shapes;
    unsigned int shapeCnt;
    unsigned int lightCnt;
    const BVHNodeGPU *btn;
    const BVHNodeGPU *btl;
    const KDNodeGPU *kng;
    int kngCnt;
    const int *kn;
    int knCnt;
    int width;
    int height;
    int currentSample;
    unsigned int *pixels;
} ptgpu_radiance_args;

/* One RadianceGPU launch: one sample per pixel, running-mean accumulation into colors, RGBA8 pack
 * into pixels, seeds written back.  Trees are read in the reference layout. */
int ptgpu_radiance_gpu(ptgpu_ctx *ctx, int accel, const ptgpu_radiance_args *args);

/* RadianceBoxGPU (rendering_kernel.cl:1152-1236, launched by ExecuteBoxKernel smallptGPU.cpp:636-650): one
 * RadianceGPU sample for the pixels of one box, x..x+bwidth-1 by y..y+bheight-1 of the frame (pixels
 * past the frame's edge are skipped), on whole-frame buffers.  The reference's seed row offset
 * (y + ywithinbox - 1) is not reproduced: each pixel uses its own seeds, as in RadianceGPU. */
int ptgpu_radiance_box_gpu(ptgpu_ctx *ctx, int accel, const ptgpu_radiance_args *args, int x, int y, int bwidth,
                           int bheight);
/* DrawBox (smallptGPU.cpp:652-723) on the host: the same box rendered by the calling thread and up to
 * `threads` - 1 more, on host buffers (every pointer in args a host pointer), with MAX_DEPTH max_depth.
 * The kernels' arithmetic is used, so the box gets the bits ptgpu_radiance_box_gpu gives it. */
int ptgpu_host_radiance_box(const ptgpu_radiance_args *args, int accel, int max_depth, int x, int y, int bwidth,
                            int bheight, int threads);

/* No reference counterpart: ptgpu_radiance_gpu for callers that cannot allocate device memory (code
 * set up for HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * arguments, work and bits as ptgpu_radiance_gpu; the call itself allocates nothing.  The traversal
 * layout converted from the caller's tree, its parent map, the light list and the pixel counter live
 * in one static device array of 128 MiB per GPU (held until the process ends), enough for about
 * 400,000 shapes in BVH mode; a frame whose tree needs more is refused with PTGPU_ERR_UNSUPPORTED
 * before any device work.  Every call on one GPU shares that array: calls that overlap on one GPU
 * (other streams, other contexts or other threads) race.  The wavefront engine (engine 2) is refused:
 * it allocates its queues. */
int ptgpu_radiance_gpu_noalloc(ptgpu_ctx *ctx, int accel, const ptgpu_radiance_args *args);

/* Fast path: render `spp` consecutive samples (currentSample .. currentSample+spp-1) in one
 * launch, traversing the handle's prepared device layout.  Results are identical to `spp`
 * successive ptgpu_radiance_gpu calls.  Exactly one of bvh/kd may be non-NULL, or both NULL for
 * the brute-force mode.  d_camera is a device Camera; d_shapes the device Shape array. */
typedef struct {
    Vec *colors;
    unsigned int *seeds;
    const Camera *camera;
    const Shape *shapes;
    unsigned int shapeCnt;
    unsigned int lightCnt;
    int width;
    int height;
    int currentSample;
    int spp;
    int max_depth;            /* MAX_DEPTH (geom.h:32); paths have max_depth+1 segments */
    unsigned int *pixels;     /* may be NULL: skip the RGBA8 pack */
    unsigned long long *ray_counter; /* may be NULL: device u64[2] += {closest, any-hit} queries */
    unsigned long long *trav_stats;  /* may be NULL: device u64[2] += {node visits, primitive tests};
                                        selects a counting variant of the kernel (slower) */
} ptgpu_render_args;
int ptgpu_render(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const ptgpu_render_args *args);

/* ptgpu_render over the listed tiles only: d_tile_origins holds ntiles (x, y) int pairs in device
 * memory, each the top-left pixel of a tile_w x tile_h tile (pixels past the frame's edge are
 * skipped).  All tiles go out in one launch; no reference counterpart (the reference launches
 * RadianceBoxGPU once per box). */
int ptgpu_render_tiles(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const ptgpu_render_args *args,
                       const int *d_tile_origins, int ntiles, int tile_w, int tile_h);

/* The reference's CPU_PARTRENDERING frame (DrawAllBoxes, smallptGPU.cpp:725-884): the frame is cut into
 * bwidth x bheight tiles in row order; the first goes to the GPU, and each later one to the host when
 * (GPU tiles) / (host tiles, counted from 1) has reached *rcpu, else to the GPU.  The GPU tiles go out
 * in one launch (ptgpu_render_tiles) while `threads` host threads render the host tiles (the host
 * side of ptgpu_host_radiance_box over the same prepared tree); the host tiles' rows are then copied
 * into the frame buffers.  *rcpu becomes (host ms per host tile) / (GPU ms per GPU tile), the
 * reference's update, for the next frame.  args is one ptgpu_render call's (spp samples from
 * currentSample; no counters); the frame gets the bits a ptgpu_render of it gives. */
typedef struct ptgpu_hybrid ptgpu_hybrid;
typedef struct {
    int gpu_tiles;
    int cpu_tiles;
    float gpu_ms;     /* the GPU tiles' launch, HIP events */
    float cpu_ms;     /* the host tiles, wall time */
    float total_ms;   /* the whole frame, wall time */
} ptgpu_hybrid_stats;
int ptgpu_hybrid_create(ptgpu_ctx *ctx, int threads, ptgpu_hybrid **out);
int ptgpu_hybrid_frame(ptgpu_hybrid *h, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const ptgpu_render_args *args,
                       int bwidth, int bheight, float *rcpu, ptgpu_hybrid_stats *stats);
int ptgpu_hybrid_destroy(ptgpu_hybrid *h);

/* Many frames in one launch (no reference counterpart: the batched form of the path).  Each job is
 * one ptgpu_render call's work -- a prepared tree (exactly one of bvh / kd, each with at least two
 * shapes), its scene and its frame buffers -- and the lanes of the launch take pixels from all jobs,
 * so frames of different scenes, sizes and sample counts are traversed together.  Every job's
 * colors / seeds / pixels come out bit-identical to its own ptgpu_render call.  Jobs must not share
 * output buffers.  args.ray_counter / args.trav_stats of the jobs are ignored; d_ray_counter (may be
 * NULL) receives the batch's {closest, any-hit} query counts. */
typedef struct {
    ptgpu_clbvh *bvh;
    ptgpu_kdtree *kd;
    ptgpu_render_args args;
} ptgpu_render_job;
int ptgpu_render_batch(ptgpu_ctx *ctx, const ptgpu_render_job *jobs, int njobs, unsigned long long *d_ray_counter);

/* ---- Scene ingest and image output (host only) ----
 * ReadTxt (displayfunc.cpp:144-505): reads the reference's text scene (gamma/sky line, camera, materials,
 * spheres, triangles and OBJ models, the first `objects` entries) into a malloc'ed Shape array (free
 * with ptgpu_host_free) in the reference's order, with its areas, DIFF/SPEC/REFR mapping and light
 * count.  Model paths are resolved against base_dir (NULL: the working directory, as the reference
 * does).  camera gets orig, target, width and height with pitch = yaw = 0; call ptgpu_update_camera.
 * OBJ numbers are converted as the reference's bundled OBJ reader does, so vertices are bit-identical. */
int ptgpu_scene_read(const char *path, const char *base_dir, int width, int height, Shape **h_shapes,
                     unsigned int *shapeCnt, unsigned int *lightCnt, Camera *camera);
/* The 'p' key's image writer (smallptGPU.cpp:1238-1258): ASCII PPM, top image row first, the r, g, b
 * bytes of each packed pixel. */
int ptgpu_write_ppm(const char *path, const unsigned int *h_pixels, int width, int height);

/* ---- HIP graphs of a per-frame call sequence (no reference counterpart) ----
 * Between ptgpu_graph_begin and ptgpu_graph_end the library calls on ctx are recorded into a HIP
 * graph instead of run; ptgpu_graph_launch replays the recording as one submission on ctx's stream.
 * A replay re-runs the recorded kernels and copies with their recorded arguments: inputs that change
 * per frame (camera, seeds) go through device buffers that the recorded sequence fills, e.g. with
 * ptgpu_memcpy_async from pinned memory the caller rewrites between replays (once the previous
 * replay has finished).  Recordable: ptgpu_render on a prepared handle (built, traversed once, same
 * lightCnt as before), ptgpu_memcpy_async, ptgpu_memcpy_d2d, ptgpu_memset_d, and the build calls of
 * an existing handle (whose first traversal afterwards is not recordable).  Calls that wait for the
 * device return PTGPU_ERR_UNSUPPORTED while recording and record nothing. */
typedef struct ptgpu_graph ptgpu_graph;
int ptgpu_graph_begin(ptgpu_ctx *ctx);
int ptgpu_graph_end(ptgpu_ctx *ctx, ptgpu_graph **out);
int ptgpu_graph_launch(ptgpu_ctx *ctx, ptgpu_graph *g);
int ptgpu_graph_destroy(ptgpu_graph *g);
/* Stream-ordered copy in any direction (host memory must be pinned for a recorded copy). */
int ptgpu_memcpy_async(ptgpu_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Page-locked host memory (hipHostMalloc). */
int ptgpu_pinned_alloc(ptgpu_ctx *ctx, size_t bytes, void **h_out);
int ptgpu_pinned_free(void *h_ptr);

/* ---- Autotuning (no reference counterpart) ----
 * Chooses the settings that change speed only -- every value gives bit-identical images, hits and
 * trees -- by timing real work on the caller's scene and frame (HIP events, fastest of `trials`).
 * Trial renders use scratch copies of the frame buffers: the caller's colors, seeds and pixels are
 * left as they were. */
typedef struct {
    int blocks_per_cu;        /* ptgpu_ctx_set_tuning names */
    int regroup;
    int wide;
    int batch_blocks_per_cu;
    int batch_regroup;
    int digit_bits;           /* LBVH sort digit width (4 or 8) */
    int treelet;              /* treelet restructuring size, 0 = none */
    float render_ms;          /* fastest trial render with the chosen settings */
    float build_ms;           /* fastest rebuild (+ optimise) with the chosen settings */
} ptgpu_tuning;
/* Tunes the digit width, the single-frame launch shape (blocks per CU 1..4, regroup 8/16/32, node
 * format: wide 2, 1 or 0) and the treelet size (none, 5, 7; minimising rebuild + optimise +
 * frames_per_build renders).  frames_per_build 0: every frame rebuilds its tree on another stream
 * while the previous frame renders (two frames in flight), so the render time alone is minimised,
 * over treelet sizes whose rebuild + optimise takes at most 1/8 of the render.  Applies the result
 * to ctx and h and leaves h holding a tree built (and optimised) with it.  args: one ptgpu_render
 * call's arguments on h (h: >= 2 shapes). */
int ptgpu_autotune(ptgpu_ctx *ctx, ptgpu_clbvh *h, const ptgpu_render_args *args, int frames_per_build,
                   int trials, ptgpu_tuning *out);
/* Tunes batch_blocks_per_cu (2..4) and batch_regroup (16/32/64) on a ptgpu_render_batch job list
 * and applies them to ctx. */
int ptgpu_autotune_batch(ptgpu_ctx *ctx, const ptgpu_render_job *jobs, int njobs, int trials,
                         ptgpu_tuning *out);

/* ---- Serving: continuous batching of render requests (no reference counterpart) ----
 * A request renders one frame of one scene: `spp` samples from sample 0.  Requests queue in
 * submission order.  ptgpu_server_tick admits queued requests, head of line first, while their HBM
 * footprint fits the budget (the scene's shapes + LBVH unless the scene key is resident, plus the
 * frame's colors/seeds/pixels), builds the scenes they bring in, and renders min(spp_per_tick,
 * remaining) samples of every admitted request in one batched launch.  Up to two launches are in
 * flight (on two streams): a tick retires the launches that have finished, and waits for the oldest
 * only when it has nothing new to launch.  A scene stays resident while requests use it and afterwards until its memory is needed
 * (least recently used first).  A finished frame stays on the device until ptgpu_server_result
 * copies it out.  Every result is bit-identical to one ptgpu_render of the request alone.
 * h_shapes and h_seeds must stay valid until the tick that admits the request returns.  Requests
 * on one scene key must use the same shapes and lightCnt. */
typedef struct ptgpu_server ptgpu_server;
typedef struct {
    unsigned long long scene_key;
    const Shape *h_shapes;
    unsigned int shapeCnt;        /* >= 2 */
    unsigned int lightCnt;
    Camera camera;                /* UpdateCamera already applied */
    int width;
    int height;
    int spp;
    int max_depth;
    const unsigned int *h_seeds;  /* 2 per pixel, or NULL: seeds[2g] = seeds[2g+1] = max(g, 2) */
} ptgpu_request;
int ptgpu_server_create(ptgpu_ctx *ctx, size_t hbm_budget, int spp_per_tick, ptgpu_server **out);
int ptgpu_server_submit(ptgpu_server *s, const ptgpu_request *r, int *id);
int ptgpu_server_tick(ptgpu_server *s, int *n_active, int *n_queued);
/* ids of the requests finished since the last call (at most cap) */
int ptgpu_server_finished(ptgpu_server *s, int *ids, int cap, int *n);
/* copies a finished frame out (any pointer may be NULL) and releases its device memory */
int ptgpu_server_result(ptgpu_server *s, int id, float *h_colors, unsigned int *h_pixels, unsigned int *h_seeds);
/* {closest, any-hit} queries so far, HBM in use, resident scenes, LBVH builds so far */
int ptgpu_server_stats(ptgpu_server *s, unsigned long long *rays2, size_t *hbm_used, int *scenes_resident,
                       unsigned long long *builds);
int ptgpu_server_destroy(ptgpu_server *s);

/* Batch ray queries over the same traversal engine (the reference's Intersect and IntersectP,
 * rendering_kernel.cl:194-395).  d_rays: n Rays.  Closest: d_t[i], d_id[i] (id = original shape
 * index, t = 0 and id = -1 on a miss; ties in t resolve to the higher index).  Any-hit: d_hit[i] =
 * 1 if a shape with index < shapeCnt-1 is hit at 0 != t < d_maxt[i]. */
int ptgpu_trace_closest(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const Shape *d_shapes,
                        unsigned int shapeCnt, const Ray *d_rays, int n, float *d_t, int *d_id);
int ptgpu_trace_any(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const Shape *d_shapes,
                    unsigned int shapeCnt, const Ray *d_rays, const float *d_maxt, int n, int *d_hit);

/* UpdateCamera (displayfunc.cpp:732-756) on the host: fills dir/x/y from yaw, pitch, width, height. */
int ptgpu_update_camera(Camera *cam);

#ifdef __cplusplus
}
#endif

#endif /* PTGPU_H */
