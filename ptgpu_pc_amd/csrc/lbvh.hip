// lbvh.hip -- LBVH build for gfx950: scene bounds + 30-bit Morton codes, onesweep LSD radix sort
// (stable, descending by Morton code), Karras radix-tree topology over leaf indices, bottom-up
// bound refit, the Morton-keyed traversal tree, and the C ABI that replaces class CLBVH
// (CLBVH.h:39-47).  The stage bodies live in lbvh_device.hpp; the kernels here run one scene.
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "device_util.hpp"
#include "lbvh.hpp"
#include "lbvh_device.hpp"

namespace ptgpu {

// -------------------------------------------------------------------------------------------
// utilities
// -------------------------------------------------------------------------------------------
__global__ void k_fill_u32(unsigned *p, unsigned v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_u8(unsigned char *p, unsigned char v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// Plain streaming copy (16-byte vectors, grid-stride): a kernel of known traffic for calibrating
// the FETCH_SIZE / WRITE_SIZE counters.
__global__ __launch_bounds__(256) void k_copy_calib(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

int fill_bytes(hipStream_t s, void *p, unsigned char v, size_t n_bytes) {
    if (((uintptr_t)p & 3) == 0 && (n_bytes & 3) == 0) return fill_u32(s, p, 0x01010101u * v, n_bytes / 4);
    if (n_bytes == 0) return PTGPU_OK;
    size_t blocks = (n_bytes + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)blocks), dim3(256), 0, s, (unsigned char *)p, v, n_bytes);
    PT_HIP(hipGetLastError());
    return PTGPU_OK;
}

int fill_u32(hipStream_t s, void *p, unsigned v, size_t n_words) {
    if (n_words == 0) return PTGPU_OK;
    size_t blocks = (n_words + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, (unsigned *)p, v, n_words);
    PT_HIP(hipGetLastError());
    return PTGPU_OK;
}

__global__ __launch_bounds__(kBoundsThreads) void k_bounds(const Shape *shapes, int n, float *partials,
                                                           ShapeBox *boxes) {
    bounds_body(shapes, n, partials, boxes, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(kBoundsThreads) void k_bounds_final(const float *partials, int nblocks, float *scene6) {
    bounds_final_body(partials, nblocks, scene6);
}

__global__ __launch_bounds__(256) void k_morton(Shape *shapes, int n, const float *scene6, unsigned *keys,
                                                const ShapeBox *boxes) {
    morton_body(shapes, n, scene6, keys, boxes, blockIdx.x);
}

template <int BITS>
__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const unsigned *keys, int n, unsigned *hist) {
    sort_hist_body<BITS>(keys, n, hist, blockIdx.x, gridDim.x);
}

template <int BITS>
__global__ __launch_bounds__(kSortThreads) void k_sort_scan(unsigned *hist) {
    sort_scan_body<BITS>(hist);
}

template <int BITS, int IN, int OUT>
__global__ __launch_bounds__(kSortThreads) void k_sort_pass(const unsigned *__restrict__ keys_in,
                                                            const unsigned *__restrict__ vals_in,
                                                            unsigned *__restrict__ keys_out,
                                                            unsigned *__restrict__ vals_out, int n, int shift,
                                                            const unsigned *__restrict__ digit_base,
                                                            unsigned *status, unsigned *tile_counter,
                                                            unsigned *err) {
    __shared__ unsigned s_tile;
    if (threadIdx.x == 0) s_tile = atomicAdd(tile_counter, 1u);
    __syncthreads();
    sort_pass_body<BITS, IN, OUT>(keys_in, vals_in, keys_out, vals_out, n, shift, digit_base, status,
                                  __builtin_amdgcn_readfirstlane(s_tile), err);
}

__global__ __launch_bounds__(256) void k_karras(int n, BVHNodeGPU *btn, BVHNodeGPU *btl) {
    karras_body(n, btn, btl, blockIdx.x);
}

template <bool LEAF, bool TNODES>
__global__ __launch_bounds__(256) void k_refit_levels(const ShapeBox *boxes, const unsigned *sorted, int n, int s0,
                                                      int levels, const float *in_bounds, float *out_bounds,
                                                      BVHNodeGPU *btn, BVHNodeGPU *btl, TNode *tnodes, Prim *prims) {
    refit_levels_body<LEAF>(boxes, sorted, n, s0, levels, in_bounds, out_bounds, btn, btl, tnodes, prims, TNODES,
                            blockIdx.x);
}

__global__ __launch_bounds__(256) void k_karras_morton(const unsigned *keys, int n, int4 *mnode, int *mpar_int,
                                                       int *mpar_leaf) {
    karras_morton_body(keys, n, mnode, mpar_int, mpar_leaf, blockIdx.x);
}

__global__ __launch_bounds__(kMChunk) void k_mrefit_chunk(int n, const int4 *mnode, const int *mpar_int,
                                                      const int *mpar_leaf, const BVHNodeGPU *btl, TNode *tnodes,
                                                      float *gbi, int *gh, int2 *xlist, unsigned *xcnt,
                                                      unsigned *root_height, unsigned *err) {
    mrefit_chunk_body(n, mnode, mpar_int, mpar_leaf, btl, tnodes, gbi, gh, xlist, xcnt, root_height, err, blockIdx.x);
}

__global__ __launch_bounds__(1024) void k_mrefit_mid(const int4 *mnode, const int *mpar_int, const BVHNodeGPU *btl,
                                                     TNode *tnodes, float *gbi, int *gh, const int2 *xlist,
                                                     const unsigned *xcnt, unsigned *gcnt, int2 *xtop,
                                                     unsigned *xtop_count, unsigned xtop_cap, unsigned *root_height,
                                                     unsigned *err) {
    mrefit_mid_body(mnode, mpar_int, btl, tnodes, gbi, gh, xlist, xcnt, gcnt, xtop, xtop_count, xtop_cap, root_height,
                    err, blockIdx.x);
}

__global__ __launch_bounds__(1024) void k_mrefit_top(const int4 *mnode, const int *mpar_int, const BVHNodeGPU *btl,
                                                     TNode *tnodes, float *gbi, int *gh, const int2 *xtop,
                                                     const unsigned *xtop_count, unsigned xtop_cap, unsigned *gcnt,
                                                     unsigned *root_height, unsigned *err) {
    mrefit_top_body(mnode, mpar_int, btl, tnodes, gbi, gh, xtop, xtop_count, xtop_cap, gcnt, root_height, err);
}

}  // namespace ptgpu

// ---------------------------------------------------------------------------------------------
// C ABI (class CLBVH)
// ---------------------------------------------------------------------------------------------
using namespace ptgpu;

static int clbvh_alloc(ptgpu_clbvh *h, int n) {
    const size_t N = (size_t)n;
    const size_t tiles = (N + kSortTile - 1) / kSortTile;
    PT_TRY(h->keysA.ensure(N * 4));
    PT_TRY(h->valsA.ensure(N * 4));
    PT_TRY(h->pairs.ensure(N * 16));   // the sort's packed (key, value) words, two buffers of N
    PT_TRY(h->btn.ensure((N > 1 ? N - 1 : 1) * sizeof(BVHNodeGPU)));
    PT_TRY(h->btl.ensure(N * sizeof(BVHNodeGPU)));
    PT_TRY(h->tnodes.ensure((N > 1 ? N - 1 : 1) * sizeof(TNode)));
    PT_TRY(h->prims.ensure(N * sizeof(Prim)));
    PT_TRY(h->boxes.ensure(N * sizeof(ShapeBox)));
    // sort scratch: histograms (<= 8 passes x 256), status words for every pass, counters
    PT_TRY(h->hist.ensure(8 * 256 * 4));
    PT_TRY(h->status.ensure(8 * tiles * 256 * 4 + 64));
    PT_TRY(h->misc.ensure(4096 * 6 * 4 + 256));
    return PTGPU_OK;
}

static inline float *misc_scene6(ptgpu_clbvh *h) { return h->scene6(); }
static inline unsigned *misc_counters(ptgpu_clbvh *h) { return h->counters(); }
static inline float *misc_partials(ptgpu_clbvh *h) { return h->partials(); }

namespace ptgpu {
int clbvh_level_scratch(ptgpu_clbvh *h, float **halves, size_t *per) {
    *per = (((size_t)h->n + kRefitItems - 1) / kRefitItems + 1) * 6;
    PT_TRY(h->opt_tmp.ensure(2 * *per * sizeof(float)));
    halves[0] = h->opt_tmp.as<float>();
    halves[1] = h->opt_tmp.as<float>() + *per;
    return PTGPU_OK;
}

int clbvh_morton_scratch(ptgpu_clbvh *h, MortonScratch *m) {
    const size_t N = (size_t)h->n;
    PT_TRY(h->mnode.ensure((N > 1 ? N - 1 : 1) * sizeof(int4)));
    PT_TRY(h->mpar.ensure((2 * N + 2) * sizeof(int)));
    PT_TRY(h->gbi.ensure((N > 1 ? N - 1 : 1) * 6 * sizeof(float) + (N + 2) * sizeof(int)));
    const size_t nsuper = (N + kMSuper - 1) / kMSuper;
    const size_t xcap = nsuper * kXCap, tcap = mtop_cap((unsigned)nsuper);
    PT_TRY(h->xlist.ensure((xcap + tcap) * sizeof(int2) + (nsuper + N + 2) * sizeof(unsigned)));
    m->mnode = h->mnode.as<int4>();
    m->mpar_int = h->mpar.as<int>();
    m->mpar_leaf = h->mpar.as<int>() + N;
    m->gbi = h->gbi.as<float>();
    m->gh = reinterpret_cast<int *>(m->gbi + (N > 1 ? N - 1 : 1) * 6);
    m->xlist = h->xlist.as<int2>();
    m->xtop = m->xlist + xcap;
    m->xtop_cap = (unsigned)tcap;
    m->xcnt = reinterpret_cast<unsigned *>(m->xtop + tcap);
    m->gcnt = m->xcnt + nsuper;
    m->nsuper = (unsigned)nsuper;
    return PTGPU_OK;
}
}  // namespace ptgpu

static int clbvh_construct(ptgpu_clbvh *h) {
    hipStream_t st = h->ctx->stream;
    const int n = h->n;
    PT_HIP(hipEventRecord(h->ev[0], st));
    int blocks = (n + kBoundsThreads - 1) / kBoundsThreads;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_bounds, dim3(blocks), dim3(kBoundsThreads), 0, st, (const Shape *)h->d_shapes, n,
                       misc_partials(h), h->boxes.as<ShapeBox>());
    PT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(kBoundsThreads), 0, st, (const float *)misc_partials(h), blocks,
                       misc_scene6(h));
    PT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_morton, dim3((n + 255) / 256), dim3(256), 0, st, h->d_shapes, n, misc_scene6(h),
                       h->keysA.as<unsigned>(), (const ShapeBox *)h->boxes.as<ShapeBox>());
    PT_HIP(hipGetLastError());
    PT_HIP(hipEventRecord(h->ev[1], st));
    h->timed[0] = true;
    h->state = 1;
    return PTGPU_OK;
}

// Refit of the index-keyed tree: ceil(log2 N) levels in passes of kRefitLog levels.
static int clbvh_refit_levels(ptgpu_clbvh *h) {
    hipStream_t st = h->ctx->stream;
    const int n = h->n;
    int L = 0;
    while ((1ll << L) < n) L++;
    size_t per = 0;
    float *buf[2];
    PT_TRY(clbvh_level_scratch(h, buf, &per));
    int s0 = 0, pass = 0;
    const float *in = nullptr;
    do {
        const int levels = (L - s0) < kRefitLog ? (L - s0) : kRefitLog;
        const long long m = ((long long)n + (1ll << s0) - 1) >> s0;
        const unsigned blocks = (unsigned)((m + kRefitItems - 1) / kRefitItems);
        float *out = (s0 + levels < L) ? buf[pass & 1] : nullptr;
        const bool tn = h->trav_topology == 0;
        if (s0 == 0) {
            if (tn)
                hipLaunchKernelGGL((k_refit_levels<true, true>), dim3(blocks), dim3(256), 0, st, h->boxes.as<ShapeBox>(), h->sorted,
                                   n, 0, levels, (const float *)nullptr, out, h->btn.as<BVHNodeGPU>(),
                                   h->btl.as<BVHNodeGPU>(), h->tnodes.as<TNode>(), h->prims.as<Prim>());
            else
                hipLaunchKernelGGL((k_refit_levels<true, false>), dim3(blocks), dim3(256), 0, st, h->boxes.as<ShapeBox>(),
                                   h->sorted, n, 0, levels, (const float *)nullptr, out, h->btn.as<BVHNodeGPU>(),
                                   h->btl.as<BVHNodeGPU>(), h->tnodes.as<TNode>(), h->prims.as<Prim>());
        } else {
            if (tn)
                hipLaunchKernelGGL((k_refit_levels<false, true>), dim3(blocks), dim3(256), 0, st, h->boxes.as<ShapeBox>(),
                                   h->sorted, n, s0, levels, in, out, h->btn.as<BVHNodeGPU>(), h->btl.as<BVHNodeGPU>(),
                                   h->tnodes.as<TNode>(), h->prims.as<Prim>());
            else
                hipLaunchKernelGGL((k_refit_levels<false, false>), dim3(blocks), dim3(256), 0, st, h->boxes.as<ShapeBox>(),
                                   h->sorted, n, s0, levels, in, out, h->btn.as<BVHNodeGPU>(), h->btl.as<BVHNodeGPU>(),
                                   h->tnodes.as<TNode>(), h->prims.as<Prim>());
        }
        PT_HIP(hipGetLastError());
        in = out;
        s0 += levels;
        pass++;
    } while (s0 < L);
    return PTGPU_OK;
}

// Morton-keyed traversal tree (topology + two-pass refit); also records the tree height.
static int clbvh_traversal_morton(ptgpu_clbvh *h) {
    hipStream_t st = h->ctx->stream;
    const int n = h->n;
    const size_t N = (size_t)n;
    MortonScratch ms;
    PT_TRY(clbvh_morton_scratch(h, &ms));
    int *mpar_int = ms.mpar_int, *mpar_leaf = ms.mpar_leaf;
    float *gbi = ms.gbi;
    int *gh = ms.gh;
    int2 *xlist = ms.xlist;
    unsigned *gcnt = ms.gcnt;
    unsigned *ctr = misc_counters(h);   // [9] error word, [10] arrivals for pass 3, [11] root height
    PT_TRY(fill_u32(st, ctr + 10, 0u, 2));
    if (n > 1) {
        PT_TRY(fill_u32(st, ms.xcnt, 0u, ms.nsuper + N - 1));   // per-super-chunk counts, then gcnt
        hipLaunchKernelGGL(k_karras_morton, dim3((n - 1 + 255) / 256), dim3(256), 0, st, h->keysA.as<unsigned>(), n,
                           h->mnode.as<int4>(), mpar_int, mpar_leaf);
        PT_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_mrefit_chunk, dim3((n + kMChunk - 1) / kMChunk), dim3(kMChunk), 0, st, n,
                           h->mnode.as<int4>(), mpar_int, mpar_leaf, h->btl.as<BVHNodeGPU>(), h->tnodes.as<TNode>(),
                           gbi, gh, xlist, ms.xcnt, ctr + 11, ctr + 9);
        PT_HIP(hipGetLastError());
        if (n > kMChunk) {
            hipLaunchKernelGGL(k_mrefit_mid, dim3(ms.nsuper), dim3(1024), 0, st, h->mnode.as<int4>(), mpar_int,
                               h->btl.as<BVHNodeGPU>(), h->tnodes.as<TNode>(), gbi, gh, (const int2 *)xlist,
                               (const unsigned *)ms.xcnt, gcnt, ms.xtop, ctr + 10, ms.xtop_cap, ctr + 11, ctr + 9);
            PT_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_mrefit_top, dim3(1), dim3(1024), 0, st, h->mnode.as<int4>(), mpar_int,
                               h->btl.as<BVHNodeGPU>(), h->tnodes.as<TNode>(), gbi, gh, (const int2 *)ms.xtop,
                               (const unsigned *)(ctr + 10), ms.xtop_cap, gcnt, ctr + 11, ctr + 9);
            PT_HIP(hipGetLastError());
        }
    }
    return PTGPU_OK;
}

template <int BITS>
static int clbvh_sort(ptgpu_clbvh *h) {
    using C = SortCfg<BITS>;
    hipStream_t st = h->ctx->stream;
    const int n = h->n;
    const size_t tiles = ((size_t)n + kSortTile - 1) / kSortTile;
    unsigned *hist = h->hist.as<unsigned>();
    unsigned *status = h->status.as<unsigned>();
    unsigned *ctr = misc_counters(h);
    PT_TRY(fill_u32(st, hist, 0u, (size_t)C::kPasses * C::kRadix));
    PT_TRY(fill_u32(st, status, 0u, (size_t)C::kPasses * tiles * C::kRadix));
    PT_TRY(fill_u32(st, ctr + 1, 0u, 9));
    int hblocks = (int)((n + kSortThreads * 8 - 1) / (kSortThreads * 8));
    if (hblocks > 2048) hblocks = 2048;
    if (hblocks < 1) hblocks = 1;
    hipLaunchKernelGGL(k_sort_hist<BITS>, dim3(hblocks), dim3(kSortThreads), 0, st, h->keysA.as<unsigned>(), n, hist);
    PT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sort_scan<BITS>, dim3(1), dim3(kSortThreads), 0, st, hist);
    PT_HIP(hipGetLastError());
    // The first pass reads the Morton keys (values: the shape indices) and writes packed (key, value)
    // words; the middle passes go between the two packed buffers; the last writes the sorted keys back
    // over the Morton keys and the sorted values beside them (an even number of passes).
    static_assert(C::kPasses % 2 == 0, "the passes end in the packed buffer the first one wrote");
    unsigned *kA = h->keysA.as<unsigned>(), *vA = h->valsA.as<unsigned>();
    unsigned *pk[2] = {h->pairs.as<unsigned>(), h->pairs.as<unsigned>() + 2 * (size_t)n};
    for (int p = 0; p < C::kPasses; p++) {
        unsigned *pst_w = status + (size_t)p * tiles * C::kRadix;
        const dim3 g((unsigned)tiles), b(kSortThreads);
        if (p == 0)
            hipLaunchKernelGGL((k_sort_pass<BITS, kSplit, kPacked>), g, b, 0, st, kA, (const unsigned *)nullptr, pk[0],
                               (unsigned *)nullptr, n, 0, hist, pst_w, ctr + 1, ctr + 9);
        else if (p == C::kPasses - 1)
            hipLaunchKernelGGL((k_sort_pass<BITS, kPacked, kSplit>), g, b, 0, st, pk[(p - 1) & 1],
                               (const unsigned *)nullptr, kA, vA, n, p * BITS, hist + p * C::kRadix, pst_w,
                               ctr + 1 + p, ctr + 9);
        else
            hipLaunchKernelGGL((k_sort_pass<BITS, kPacked, kPacked>), g, b, 0, st, pk[(p - 1) & 1],
                               (const unsigned *)nullptr, pk[p & 1], (unsigned *)nullptr, n, p * BITS,
                               hist + p * C::kRadix, pst_w, ctr + 1 + p, ctr + 9);
        PT_HIP(hipGetLastError());
    }
    h->sorted = vA;
    return PTGPU_OK;
}

namespace ptgpu {
// The Morton-keyed tree's height is known on the device after the refit; the traversal stack is
// sized from it, so the first use after a build reads it back (one 4-byte copy).
int clbvh_prepare_traversal(ptgpu_clbvh *h, bool recording, bool exact) {
    const int known = h->known_height[h->optimized ? 1 : 0];
    const bool no_readback = recording || h->ctx->capturing || (h->speculate && known > 0);
    if (h->height_pending && exact)
        PT_CHECK(!recording && !h->ctx->capturing, PTGPU_ERR_UNSUPPORTED,
                 "this call needs the tree height, which a recording cannot read back");
    if (h->height_pending && !exact && no_readback) {
        // Recorded into a graph, or a handle set to skip the read-back: size the stack for a bound
        // and let the traversal check the built tree against it on the device (error word,
        // height).  The read-back stays pending for a use that needs the height itself.
        int bound = known > 0 ? known + kRecordedHeightMargin : kRecordedHeightDefault;
        // a known height the four-wide stack holds keeps the bound within it (the margin is a guess;
        // the four-wide walk is the fast one)
        if (known > 0 && known <= kRecordedHeightDefault && bound > kRecordedHeightDefault)
            bound = kRecordedHeightDefault;
        h->acc.depth = bound;
        h->acc.height_check = misc_counters(h) + 9;
    } else if (h->height_pending) {
        unsigned w[3] = {0, 0, 0};   // error word, pass-3 arrivals, height
        PT_HIP(hipMemcpyAsync(w, misc_counters(h) + 9, 12, hipMemcpyDeviceToHost, h->ctx->stream));
        PT_HIP(hipStreamSynchronize(h->ctx->stream));
        PT_CHECK(w[0] == 0, PTGPU_ERR_HIP, "LBVH build failed (error word %u: 1 sort look-back spin limit, 2 "
                 "traversal-tree list overflow, 4 climb beyond the height bound)", w[0]);
        h->acc.depth = (int)w[2];
        h->acc.height_check = nullptr;
        h->known_height[h->optimized ? 1 : 0] = (int)w[2];
        h->height_pending = false;
    }
    if (h->wide_pending) {
        PT_TRY(build_wide(h->ctx, &h->acc, h->tnodes4));
        h->wide_pending = false;
    }
    return PTGPU_OK;
}
}  // namespace ptgpu

int ptgpu::clbvh_new(ptgpu_ctx *ctx, int n, ptgpu_clbvh **out) {
    ptgpu_clbvh *h = new ptgpu_clbvh();
    h->ctx = ctx;
    h->n = n;
    h->cap = n;
    for (auto &e : h->ev) {
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) {
            ptgpu_clbvh_destroy(h);
            return hip_fail(r, "hipEventCreate", __FILE__, __LINE__);
        }
    }
    int s = clbvh_alloc(h, n);
    if (s == PTGPU_OK) s = fill_u32(ctx->stream, h->misc.p, 0u, h->misc.bytes / 4);
    if (s != PTGPU_OK) {
        ptgpu_clbvh_destroy(h);
        return s;
    }
    *out = h;
    return PTGPU_OK;
}

void ptgpu::clbvh_mark_built(ptgpu_clbvh *h) {
    h->radix_pending = false;
    const int n = h->n;
    h->state = 3;
    h->timed[4] = false;   // an optimisation timed before this build no longer describes the tree
    h->optimized = false;
    h->export_pending = false;
    h->acc.nodes = h->tnodes.as<TNode>();
    h->acc.prims = h->prims.as<Prim>();
    h->acc.nprims = n;
    h->acc.nnodes = n - 1;
    h->acc.nshapes = n;
    h->acc.mode = PTGPU_ACCEL_BVH;
    h->acc.leaf_runs = 1;
    h->acc.nodes4 = nullptr;
    h->acc.nodes4h = nullptr;
    h->wide_pending = n >= 2;   // four-wide nodes are derived at the first traversal after a build
    h->acc.height_check = nullptr;
    if (n >= 2) {
        h->acc.root = 0;
        h->acc.root_cnt = 0;
        int d = 0;
        while ((1ll << d) < n) d++;
        h->acc.depth = d;                 // index-keyed: ceil(log2 N); Morton-keyed: read back lazily
    } else {
        h->acc.root = ~0;
        h->acc.root_cnt = 1;
        h->acc.depth = 0;
    }
    h->height_pending = h->trav_topology == 1 && n >= 2;
}

extern "C" {

int ptgpu_copy_calibration(ptgpu_ctx *ctx, void *d_dst, const void *d_src, size_t bytes) {
    PT_CHECK(ctx && d_dst && d_src && bytes % 16 == 0, PTGPU_ERR_INVALID, "bad calibration copy arguments");
    const size_t n = bytes / 16;
    hipLaunchKernelGGL(k_copy_calib, dim3(ctx->cu_count * 8), dim3(256), 0, ctx->stream, (const uint4 *)d_src,
                       (uint4 *)d_dst, n);
    PT_HIP(hipGetLastError());
    return PTGPU_OK;
}

int ptgpu_clbvh_export_traversal(ptgpu_clbvh *h, int *h_refs4, float *h_boxes12, int *height) {
    PT_CHECK(h && h_refs4 && h_boxes12, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(h->state >= 3, PTGPU_ERR_INVALID, "build_bvh_tree has not run");
    PT_TRY(clbvh_prepare_traversal(h, false, true));
    const int m = h->n - 1;
    if (m <= 0) {
        if (height) *height = 0;
        return PTGPU_OK;
    }
    std::vector<TNode> tn((size_t)m);
    PT_HIP(hipMemcpyAsync(tn.data(), h->tnodes.p, (size_t)m * sizeof(TNode), hipMemcpyDeviceToHost, h->ctx->stream));
    std::vector<int4> mn;
    if (h->trav_topology == 1 && !h->optimized) {
        mn.resize((size_t)m);
        PT_HIP(hipMemcpyAsync(mn.data(), h->mnode.p, (size_t)m * sizeof(int4), hipMemcpyDeviceToHost, h->ctx->stream));
    }
    std::vector<BVHNodeGPU> bt;
    if (h->trav_topology == 0 && !h->optimized) {
        bt.resize((size_t)m);
        PT_HIP(hipMemcpyAsync(bt.data(), h->btn.p, (size_t)m * sizeof(BVHNodeGPU), hipMemcpyDeviceToHost,
                              h->ctx->stream));
    }
    PT_HIP(hipStreamSynchronize(h->ctx->stream));
    for (int i = 0; i < m; i++) {
        const TNode &t = tn[i];
        h_refs4[i * 4 + 0] = t.r.x;
        h_refs4[i * 4 + 1] = t.r.y;
        // leaf ranges [lo, hi] of the build's tree; an optimised tree has none (-1)
        h_refs4[i * 4 + 2] = h->optimized ? -1 : h->trav_topology == 1 ? mn[i].z : bt[i].min;
        h_refs4[i * 4 + 3] = h->optimized ? -1 : h->trav_topology == 1 ? mn[i].w : bt[i].max;
        const float v[12] = {t.a.x, t.a.w, t.a.y, t.b.x, t.a.z, t.b.y, t.b.z, t.c.y, t.b.w, t.c.z, t.c.x, t.c.w};
        for (int f = 0; f < 12; f++) h_boxes12[i * 12 + f] = v[f];   // two Bounds in reference order
    }
    if (height) *height = h->acc.depth;
    return PTGPU_OK;
}

int ptgpu_clbvh_set_height_readback(ptgpu_clbvh *h, int readback) {
    PT_CHECK(h, PTGPU_ERR_INVALID, "null handle");
    PT_CHECK(readback == 0 || readback == 1, PTGPU_ERR_INVALID, "readback %d (1 read back, 0 device-checked bound)",
             readback);
    h->speculate = readback == 0;
    return PTGPU_OK;
}

int ptgpu_clbvh_set_traversal_topology(ptgpu_clbvh *h, int mode) {
    PT_CHECK(h, PTGPU_ERR_INVALID, "null handle");
    PT_CHECK(mode == 0 || mode == 1, PTGPU_ERR_INVALID, "traversal topology %d (0 reference, 1 Morton)", mode);
    h->trav_topology = mode;
    return PTGPU_OK;
}

int ptgpu_clbvh_create(ptgpu_ctx *ctx, Shape *h_shapes, int n, ptgpu_clbvh **out) {
    PT_CHECK(ctx && h_shapes && out, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(n >= 1 && n < (1 << 30), PTGPU_ERR_INVALID, "shape count %d out of range [1, 2^30)", n);
    *out = nullptr;
    PT_HIP(hipSetDevice(ctx->device));
    ptgpu_clbvh *h = nullptr;
    PT_TRY(clbvh_new(ctx, n, &h));
    int s = h->own_shapes.ensure((size_t)n * sizeof(Shape));
    if (s == PTGPU_OK) {
        h->d_shapes = h->own_shapes.as<Shape>();
        hipError_t e = hipMemcpyAsync(h->d_shapes, h_shapes, (size_t)n * sizeof(Shape), hipMemcpyHostToDevice,
                                      ctx->stream);
        if (e != hipSuccess) s = hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__);
    }
    if (s == PTGPU_OK) s = clbvh_construct(h);
    if (s == PTGPU_OK) {
        // the reference writes index, morton_code and b into the caller's array (CLBVH.cpp:102-159)
        hipError_t e = hipMemcpyAsync(h_shapes, h->d_shapes, (size_t)n * sizeof(Shape), hipMemcpyDeviceToHost,
                                      ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) s = hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__);
    }
    if (s != PTGPU_OK) {
        ptgpu_clbvh_destroy(h);
        return s;
    }
    *out = h;
    return PTGPU_OK;
}

int ptgpu_clbvh_create_device(ptgpu_ctx *ctx, Shape *d_shapes, int n, ptgpu_clbvh **out) {
    PT_CHECK(ctx && d_shapes && out, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(n >= 1 && n < (1 << 30), PTGPU_ERR_INVALID, "shape count %d out of range [1, 2^30)", n);
    *out = nullptr;
    PT_HIP(hipSetDevice(ctx->device));
    ptgpu_clbvh *h = nullptr;
    PT_TRY(clbvh_new(ctx, n, &h));
    h->d_shapes = d_shapes;
    int s = clbvh_construct(h);
    if (s != PTGPU_OK) {
        ptgpu_clbvh_destroy(h);
        return s;
    }
    *out = h;
    return PTGPU_OK;
}

int ptgpu_clbvh_reset(ptgpu_clbvh *h, int n) {
    PT_CHECK(h, PTGPU_ERR_INVALID, "null handle");
    PT_CHECK(n >= 1 && n <= h->cap, PTGPU_ERR_INVALID, "n=%d exceeds the handle capacity %d", n, h->cap);
    h->n = n;
    h->optimized = false;
    h->export_pending = false;
    return clbvh_construct(h);
}

int ptgpu_clbvh_build_radix_tree(ptgpu_clbvh *h) {
    PT_CHECK(h && h->state >= 1, PTGPU_ERR_INVALID, "handle not constructed");
    hipStream_t st = h->ctx->stream;
    const int n = h->n;
    PT_HIP(hipEventRecord(h->ev[2], st));
    // The topology is a closed form of N (LCP = clz(i ^ j)): the refit writes whole records,
    // topology and bounds together, so the records are written here only when they are read
    // before a refit (radix_pending, clbvh_materialize_radix).
    h->radix_pending = true;
    PT_HIP(hipEventRecord(h->ev[3], st));
    h->timed[2] = true;
    h->state = 2;
    h->optimized = false;
    h->export_pending = false;
    return PTGPU_OK;
}

int ptgpu_clbvh_build_bvh_tree(ptgpu_clbvh *h) {
    PT_CHECK(h && h->state >= 2, PTGPU_ERR_INVALID, "build_radix_tree must run first");
    hipStream_t st = h->ctx->stream;
    const int n = h->n;
    PT_HIP(hipEventRecord(h->ev[4], st));
    if (h->digit_bits == 4) PT_TRY(clbvh_sort<4>(h));
    else PT_TRY(clbvh_sort<8>(h));
    PT_HIP(hipEventRecord(h->ev[5], st));
    PT_TRY(clbvh_refit_levels(h));
    h->radix_pending = false;
    PT_HIP(hipEventRecord(h->ev[9], st));
    if (h->trav_topology == 1) PT_TRY(clbvh_traversal_morton(h));
    PT_HIP(hipEventRecord(h->ev[6], st));
    h->timed[1] = true;
    h->timed[3] = true;
    clbvh_mark_built(h);
    return PTGPU_OK;
}

int ptgpu_clbvh_set_digit_bits(ptgpu_clbvh *h, int bits) {
    PT_CHECK(h, PTGPU_ERR_INVALID, "null handle");
    PT_CHECK(bits == 4 || bits == 8, PTGPU_ERR_INVALID, "digit width %d (4 or 8)", bits);
    h->digit_bits = bits;
    return PTGPU_OK;
}

// The exported arrays of a tree built only to the radix-tree stage (CLBVH.cpp:83-88 fills, then
// BVH.cl:127-208): k_karras writes them on demand.
static int clbvh_materialize_radix(ptgpu_clbvh *h) {
    if (!h->radix_pending || h->state != 2) return PTGPU_OK;
    hipStream_t st = h->ctx->stream;
    const int n = h->n;
    if (n == 1) PT_TRY(fill_u32(st, h->btl.p, 0xFFFFFFFFu, sizeof(BVHNodeGPU) / 4));
    if (n > 1) {
        hipLaunchKernelGGL(k_karras, dim3((n - 1 + 255) / 256), dim3(256), 0, st, n, h->btn.as<BVHNodeGPU>(),
                           h->btl.as<BVHNodeGPU>());
        PT_HIP(hipGetLastError());
    }
    h->radix_pending = false;
    return PTGPU_OK;
}

int ptgpu_clbvh_get_trees(ptgpu_clbvh *h, BVHNodeGPU **h_btn, BVHNodeGPU **h_btl) {
    PT_CHECK(h && h_btn && h_btl, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(h->state >= 2, PTGPU_ERR_INVALID, "no tree built");
    PT_TRY(clbvh_materialize_radix(h));
    PT_TRY(clbvh_export_optimized(h));
    const int n = h->n;
    h->h_btn.resize(n > 1 ? n - 1 : 0);
    h->h_btl.resize(n);
    if (n > 1)
        PT_HIP(hipMemcpyAsync(h->h_btn.data(), h->btn.p, (size_t)(n - 1) * sizeof(BVHNodeGPU), hipMemcpyDeviceToHost,
                              h->ctx->stream));
    PT_HIP(hipMemcpyAsync(h->h_btl.data(), h->btl.p, (size_t)n * sizeof(BVHNodeGPU), hipMemcpyDeviceToHost,
                          h->ctx->stream));
    unsigned err = 0;
    PT_HIP(hipMemcpyAsync(&err, misc_counters(h) + 9, 4, hipMemcpyDeviceToHost, h->ctx->stream));
    PT_HIP(hipStreamSynchronize(h->ctx->stream));
    PT_CHECK(err == 0, PTGPU_ERR_HIP, "LBVH build failed (error word %u: 1 sort look-back spin limit, 2 "
             "traversal-tree list overflow, 4 climb beyond the height bound)", err);
    *h_btn = h->h_btn.data();
    *h_btl = h->h_btl.data();
    return PTGPU_OK;
}

int ptgpu_clbvh_device_trees(ptgpu_clbvh *h, const BVHNodeGPU **d_btn, const BVHNodeGPU **d_btl) {
    PT_CHECK(h && d_btn && d_btl, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(h->state >= 2, PTGPU_ERR_INVALID, "no tree built");
    PT_TRY(clbvh_materialize_radix(h));
    PT_TRY(clbvh_export_optimized(h));
    *d_btn = h->btn.as<BVHNodeGPU>();
    *d_btl = h->btl.as<BVHNodeGPU>();
    return PTGPU_OK;
}

int ptgpu_clbvh_device_bytes(ptgpu_clbvh *h, size_t *bytes) {
    PT_CHECK(h && bytes, PTGPU_ERR_INVALID, "null argument");
    const DevBuf *bufs[] = {&h->own_shapes, &h->keysA, &h->pairs, &h->valsA, &h->btn, &h->btl,
                            &h->tnodes, &h->prims, &h->hist, &h->status, &h->misc, &h->lights, &h->scan,
                            &h->opt_tmp, &h->mnode, &h->mpar, &h->gbi, &h->xlist, &h->tnodes4, &h->treelet,
                            &h->boxes};
    size_t t = 0;
    for (const DevBuf *b : bufs) t += b->bytes;
    // the four-wide view is derived at the first traversal: count it before it exists
    if (!h->tnodes4.p && h->n > 1) t += (size_t)(h->n - 1) * (sizeof(TNode4) + sizeof(TNode4H));
    *bytes = t;
    return PTGPU_OK;
}

int ptgpu_clbvh_device_sorted(ptgpu_clbvh *h, const int **d_sorted) {
    PT_CHECK(h && d_sorted, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(h->state >= 3, PTGPU_ERR_INVALID, "build_bvh_tree has not run");
    *d_sorted = reinterpret_cast<const int *>(h->sorted);
    return PTGPU_OK;
}

int ptgpu_clbvh_stage_times(ptgpu_clbvh *h, float *ms5) {
    PT_CHECK(h && ms5, PTGPU_ERR_INVALID, "null argument");
    PT_HIP(hipStreamSynchronize(h->ctx->stream));
    const int pairs[5][2] = {{0, 1}, {4, 5}, {2, 3}, {5, 6}, {7, 8}};
    for (int k = 0; k < 5; k++) {
        ms5[k] = 0.f;
        if (!h->timed[k]) continue;
        PT_HIP(hipEventElapsedTime(&ms5[k], h->ev[pairs[k][0]], h->ev[pairs[k][1]]));
    }
    return PTGPU_OK;
}

int ptgpu_clbvh_stage_times6(ptgpu_clbvh *h, float *ms6) {
    PT_CHECK(h && ms6, PTGPU_ERR_INVALID, "null argument");
    PT_HIP(hipStreamSynchronize(h->ctx->stream));
    const int pairs[6][2] = {{0, 1}, {4, 5}, {2, 3}, {5, 9}, {9, 6}, {7, 8}};
    const int timed[6] = {0, 1, 2, 3, 3, 4};
    for (int k = 0; k < 6; k++) {
        ms6[k] = 0.f;
        if (!h->timed[timed[k]]) continue;
        PT_HIP(hipEventElapsedTime(&ms6[k], h->ev[pairs[k][0]], h->ev[pairs[k][1]]));
    }
    return PTGPU_OK;
}

int ptgpu_clbvh_destroy(ptgpu_clbvh *h) {
    if (!h) return PTGPU_OK;
    if (h->ctx) PT_NO_CAPTURE(h->ctx, "ptgpu_clbvh_destroy");
    if (h->ctx && h->ctx->stream) (void)hipStreamSynchronize(h->ctx->stream);
    h->keysA.release(); h->pairs.release(); h->valsA.release();
    h->btn.release(); h->btl.release(); h->tnodes.release(); h->prims.release();
    h->hist.release(); h->status.release(); h->misc.release(); h->own_shapes.release();
    h->lights.release(); h->scan.release(); h->opt_tmp.release();
    h->mnode.release(); h->mpar.release(); h->gbi.release(); h->xlist.release(); h->tnodes4.release();
    h->boxes.release();
    h->treelet.release();
    for (auto &e : h->ev)
        if (e) (void)hipEventDestroy(e);
    delete h;
    return PTGPU_OK;
}

}  // extern "C"
