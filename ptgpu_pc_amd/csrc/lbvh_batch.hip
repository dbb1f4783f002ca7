// lbvh_batch.hip -- many LBVH builds in one sequence of launches (ptgpu_clbvh_build_batch).
//
// Every stage of the build (bounds, Morton, Karras topology, radix-sort histogram / scan / passes,
// level refit, Morton-keyed traversal tree) runs as ONE launch over all scenes.  A per-stage table
// of block offsets (prefix sums of each scene's block count) is the work queue: block b of a launch
// finds its scene by binary search and plays block (b - offset) of that scene, running the same
// device body as the single-scene kernel, so each scene's result is bit-identical to building it
// alone.  The sort passes take tile ids from one atomic counter for the whole batch; a tile's
// decoupled look-back only waits on earlier tiles of its own scene, which hold smaller ids and have
// therefore already started.  Small scenes fill the chip together instead of leaving it idle
// behind a launch each, and the single-workgroup top pass of the traversal-tree refit runs for all
// scenes at once.  One 8-byte-per-scene readback at the end carries tree heights and sort errors.
#include <cstring>
#include <unordered_set>
#include <vector>

#include "device_util.hpp"
#include "lbvh.hpp"
#include "lbvh_device.hpp"

namespace ptgpu {
namespace {

// One scene's buffers and sizes, as the batched kernels see them.
struct BuildSeg {
    Shape *shapes;
    ShapeBox *boxes;
    unsigned *keysA, *valsA;
    unsigned *pairs;        // the sort's packed (key, value) words: two buffers of n
    unsigned *hist, *status;
    unsigned *ctr;          // counters of the handle's misc block (see ptgpu_clbvh::counters)
    float *partials, *scene6;
    BVHNodeGPU *btn, *btl;
    TNode *tnodes;
    Prim *prims;
    float *lvl;             // level-refit hand-off: two halves of lvl_per floats
    int4 *mnode;
    int *mpar_int, *mpar_leaf;
    float *gbi;
    int *gh;
    int2 *xlist, *xtop;
    unsigned *xcnt, *gcnt;
    unsigned xtop_cap, pad;
    long long lvl_per;
    int n, tiles, levels_total, topo;
};

struct FillSeg {
    unsigned *p;
    unsigned long long words;
    unsigned value, pad;
};

enum Stage { kStBounds, kStMorton, kStKarras, kStHist, kStTiles, kStRefit0, kStRefit1, kStRefit2, kStKMorton,
             kStMChunk, kStMSuper, kStages };
constexpr int kMaxRefitPasses = 3;   // ceil(30 / kRefitLog) levels for N < 2^30
constexpr int kFillWords = 4096;     // words cleared per block of the fill launch

// Largest s with starts[s] <= b (scenes with no blocks have equal neighbouring offsets).
__device__ __forceinline__ int seg_find(const int *starts, int nseg, int b) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void kb_fill(const FillSeg *f, const int *starts, int nf) {
    const int e = seg_find(starts, nf, blockIdx.x);
    const FillSeg F = f[e];
    const unsigned long long w0 = (unsigned long long)(blockIdx.x - starts[e]) * kFillWords;
    for (int j = threadIdx.x; j < kFillWords; j += 256) {
        const unsigned long long w = w0 + j;
        if (w < F.words) F.p[w] = F.value;
    }
}

__global__ __launch_bounds__(kBoundsThreads) void kb_bounds(const BuildSeg *g, const int *starts, int ns) {
    const int s = seg_find(starts, ns, blockIdx.x);
    const BuildSeg &S = g[s];
    bounds_body(S.shapes, S.n, S.partials, S.boxes, blockIdx.x - starts[s], starts[s + 1] - starts[s]);
}

// one block per scene
__global__ __launch_bounds__(kBoundsThreads) void kb_bounds_final(const BuildSeg *g, const int *starts) {
    const BuildSeg &S = g[blockIdx.x];
    bounds_final_body(S.partials, starts[blockIdx.x + 1] - starts[blockIdx.x], S.scene6);
}

__global__ __launch_bounds__(256) void kb_morton(const BuildSeg *g, const int *starts, int ns) {
    const int s = seg_find(starts, ns, blockIdx.x);
    const BuildSeg &S = g[s];
    morton_body(S.shapes, S.n, S.scene6, S.keysA, S.boxes, blockIdx.x - starts[s]);
}

__global__ __launch_bounds__(256) void kb_karras(const BuildSeg *g, const int *starts, int ns) {
    const int s = seg_find(starts, ns, blockIdx.x);
    const BuildSeg &S = g[s];
    karras_body(S.n, S.btn, S.btl, blockIdx.x - starts[s]);
}

template <int BITS>
__global__ __launch_bounds__(kSortThreads) void kb_sort_hist(const BuildSeg *g, const int *starts, int ns) {
    const int s = seg_find(starts, ns, blockIdx.x);
    const BuildSeg &S = g[s];
    sort_hist_body<BITS>(S.keysA, S.n, S.hist, blockIdx.x - starts[s], starts[s + 1] - starts[s]);
}

template <int BITS>
__global__ __launch_bounds__(kSortThreads) void kb_sort_scan(const BuildSeg *g) {
    sort_scan_body<BITS>(g[blockIdx.x].hist);
}

// Pass p of every scene's sort.  Pass 0 reads the Morton keys in keysA with implicit values
// 0..N-1 and writes packed (key, value) words; the middle passes go between the two packed buffers;
// the last (odd) pass writes the sorted keys and values to keysA/valsA.
template <int BITS, int IN, int OUT>
__global__ __launch_bounds__(kSortThreads) void kb_sort_pass(const BuildSeg *g, const int *starts, int ns, int p,
                                                             unsigned *tile_counter) {
    using C = SortCfg<BITS>;
    __shared__ unsigned s_tile;
    if (threadIdx.x == 0) s_tile = atomicAdd(tile_counter, 1u);
    __syncthreads();
    const int t = (int)__builtin_amdgcn_readfirstlane(s_tile);
    const int s = seg_find(starts, ns, t);
    const BuildSeg &S = g[s];
    unsigned *pk[2] = {S.pairs, S.pairs + 2 * (size_t)S.n};
    const unsigned *kin = IN == kPacked ? pk[(p - 1) & 1] : S.keysA;
    unsigned *kout = OUT == kPacked ? pk[p & 1] : S.keysA;
    sort_pass_body<BITS, IN, OUT>(kin, nullptr, kout, S.valsA, S.n, p * BITS, S.hist + p * C::kRadix,
                                  S.status + (size_t)p * S.tiles * C::kRadix, (unsigned)(t - starts[s]), S.ctr + 9);
}

// Pass p of every scene's level refit: levels [p*kRefitLog, min(L, (p+1)*kRefitLog)).
template <bool LEAF>
__global__ __launch_bounds__(256) void kb_refit(const BuildSeg *g, const int *starts, int ns, int p) {
    const int s = seg_find(starts, ns, blockIdx.x);
    const BuildSeg &S = g[s];
    const int L = S.levels_total, s0 = p * kRefitLog;
    const int levels = (L - s0) < kRefitLog ? (L - s0) : kRefitLog;
    const float *in = p == 0 ? nullptr : S.lvl + ((p - 1) & 1) * S.lvl_per;
    float *out = (s0 + levels < L) ? S.lvl + (p & 1) * S.lvl_per : nullptr;
    refit_levels_body<LEAF>(S.boxes, S.valsA, S.n, s0, levels, in, out, S.btn, S.btl, S.tnodes, S.prims, S.topo == 0,
                            blockIdx.x - starts[s]);
}

__global__ __launch_bounds__(256) void kb_karras_morton(const BuildSeg *g, const int *starts, int ns) {
    const int s = seg_find(starts, ns, blockIdx.x);
    const BuildSeg &S = g[s];
    karras_morton_body(S.keysA, S.n, S.mnode, S.mpar_int, S.mpar_leaf, blockIdx.x - starts[s]);
}

__global__ __launch_bounds__(kMChunk) void kb_mrefit_chunk(const BuildSeg *g, const int *starts, int ns) {
    const int s = seg_find(starts, ns, blockIdx.x);
    const BuildSeg &S = g[s];
    mrefit_chunk_body(S.n, S.mnode, S.mpar_int, S.mpar_leaf, S.btl, S.tnodes, S.gbi, S.gh, S.xlist, S.xcnt,
                      S.ctr + 11, S.ctr + 9, blockIdx.x - starts[s]);
}

__global__ __launch_bounds__(1024) void kb_mrefit_mid(const BuildSeg *g, const int *starts, int ns) {
    const int s = seg_find(starts, ns, blockIdx.x);
    const BuildSeg &S = g[s];
    mrefit_mid_body(S.mnode, S.mpar_int, S.btl, S.tnodes, S.gbi, S.gh, S.xlist, S.xcnt, S.gcnt, S.xtop, S.ctr + 10,
                    S.xtop_cap, S.ctr + 11, S.ctr + 9, (unsigned)(blockIdx.x - starts[s]));
}

// One workgroup per scene: the chunk-spanning nodes of every scene are refitted concurrently.
__global__ __launch_bounds__(1024) void kb_mrefit_top(const BuildSeg *g) {
    const BuildSeg &S = g[blockIdx.x];
    if (S.topo != 1 || S.n <= kMChunk) return;
    mrefit_top_body(S.mnode, S.mpar_int, S.btl, S.tnodes, S.gbi, S.gh, S.xtop, S.ctr + 10, S.xtop_cap, S.gcnt,
                    S.ctr + 11, S.ctr + 9);
}

// heights[s] = traversal-tree height, errs[s] = build error word
__global__ __launch_bounds__(256) void kb_gather(const BuildSeg *g, int ns, unsigned *out) {
    for (int s = threadIdx.x; s < ns; s += 256) {
        out[s] = g[s].ctr[11];
        out[ns + s] = g[s].ctr[9];
    }
}

inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

template <int BITS>
int build_batch(ptgpu_ctx *ctx, ptgpu_clbvh *const *hs, int ns) {
    using C = SortCfg<BITS>;
    hipStream_t st = ctx->stream;
    std::vector<BuildSeg> segs((size_t)ns);
    std::vector<int> starts((size_t)kStages * (ns + 1), 0);
    std::vector<FillSeg> fills;
    auto blocks = [&](int stage, int s) -> int & { return starts[(size_t)stage * (ns + 1) + s + 1]; };
    for (int s = 0; s < ns; s++) {
        ptgpu_clbvh *h = hs[s];
        const int n = h->n;
        BuildSeg &S = segs[s];
        std::memset(&S, 0, sizeof(S));
        S.shapes = h->d_shapes;
        S.boxes = h->boxes.as<ShapeBox>();
        S.keysA = h->keysA.as<unsigned>();
        S.valsA = h->valsA.as<unsigned>();
        S.pairs = h->pairs.as<unsigned>();
        S.hist = h->hist.as<unsigned>();
        S.status = h->status.as<unsigned>();
        S.ctr = h->counters();
        S.partials = h->partials();
        S.scene6 = h->scene6();
        S.btn = h->btn.as<BVHNodeGPU>();
        S.btl = h->btl.as<BVHNodeGPU>();
        S.tnodes = h->tnodes.as<TNode>();
        S.prims = h->prims.as<Prim>();
        float *halves[2];
        size_t per = 0;
        PT_TRY(clbvh_level_scratch(h, halves, &per));
        S.lvl = halves[0];
        S.lvl_per = (long long)per;
        S.n = n;
        S.tiles = (int)(((long long)n + kSortTile - 1) / kSortTile);
        int L = 0;
        while ((1ll << L) < n) L++;
        S.levels_total = L;
        S.topo = h->trav_topology;
        const bool mtree = S.topo == 1 && n > 1;
        if (mtree) {
            MortonScratch ms;
            PT_TRY(clbvh_morton_scratch(h, &ms));
            S.mnode = ms.mnode;
            S.mpar_int = ms.mpar_int;
            S.mpar_leaf = ms.mpar_leaf;
            S.gbi = ms.gbi;
            S.gh = ms.gh;
            S.xlist = ms.xlist;
            S.xtop = ms.xtop;
            S.xcnt = ms.xcnt;
            S.xtop_cap = ms.xtop_cap;
            S.gcnt = ms.gcnt;
        }
        // block counts: the single-scene launches' grids
        int bb = (n + kBoundsThreads - 1) / kBoundsThreads;
        blocks(kStBounds, s) = bb > 4096 ? 4096 : bb < 1 ? 1 : bb;
        blocks(kStMorton, s) = (n + 255) / 256;
        blocks(kStKarras, s) = 0;   // the refit writes whole btn/btl records (topology included)
        int hb = (int)((n + kSortThreads * 8 - 1) / (kSortThreads * 8));
        blocks(kStHist, s) = hb > 2048 ? 2048 : hb < 1 ? 1 : hb;
        blocks(kStTiles, s) = S.tiles;
        for (int p = 0; p < kMaxRefitPasses; p++) {
            const int s0 = p * kRefitLog;
            if (p > 0 && s0 >= L) continue;   // pass 0 always runs: it writes the leaves
            const long long m = ((long long)n + (1ll << s0) - 1) >> s0;
            blocks(kStRefit0 + p, s) = (int)((m + kRefitItems - 1) / kRefitItems);
        }
        blocks(kStKMorton, s) = mtree ? (n - 1 + 255) / 256 : 0;
        blocks(kStMChunk, s) = mtree ? (n + kMChunk - 1) / kMChunk : 0;
        blocks(kStMSuper, s) = mtree && n > kMChunk ? (int)((n + kMSuper - 1) / kMSuper) : 0;
        // clears done by the single-scene path (CLBVH.cpp:83-88 for btn/btl: kb_karras writes them,
        // except the lone leaf of a one-shape scene)
        if (n == 1) fills.push_back({(unsigned *)S.btl, (unsigned long long)sizeof(BVHNodeGPU) / 4, 0xFFFFFFFFu, 0u});
        fills.push_back({S.hist, (unsigned long long)C::kPasses * C::kRadix, 0u, 0u});
        fills.push_back({S.status, (unsigned long long)C::kPasses * S.tiles * C::kRadix, 0u, 0u});
        fills.push_back({S.ctr + 9, 3ull, 0u, 0u});
        if (mtree) fills.push_back({S.xcnt, (unsigned long long)(S.gcnt - S.xcnt) + (n - 1), 0u, 0u});
    }
    for (int k = 0; k < kStages; k++)
        for (int s = 0; s < ns; s++) starts[(size_t)k * (ns + 1) + s + 1] += starts[(size_t)k * (ns + 1) + s];
    auto total = [&](int stage) { return starts[(size_t)stage * (ns + 1) + ns]; };
    auto table = [&](int stage) { return (size_t)stage * (ns + 1); };

    // device / pinned layout: segs | starts | fill entries | fill offsets || tile counters | readback
    const int nf = (int)fills.size() + 1;
    const size_t o_starts = align16((size_t)ns * sizeof(BuildSeg));
    const size_t o_fills = align16(o_starts + starts.size() * sizeof(int));
    const size_t o_fstarts = align16(o_fills + (size_t)nf * sizeof(FillSeg));
    const size_t upload = align16(o_fstarts + (size_t)(nf + 1) * sizeof(int));
    const size_t o_ctr = upload;
    const size_t o_out = o_ctr + 64;
    const size_t dev_bytes = o_out + (size_t)2 * ns * sizeof(unsigned);
    BuildStaging &B = ctx->build_staging;
    const bool recording = ctx->capturing;
    std::vector<char> img;   // recorded: the tables' image, staged by record_stage
    char *dev = nullptr, *hp = nullptr;
    if (recording) {
        img.assign(upload, 0);
        hp = img.data();
    } else {
        PT_TRY(B.table.ensure(dev_bytes));
        if (B.pinned_bytes < dev_bytes) {
            if (B.pinned) PT_HIP(hipHostFree(B.pinned));
            B.pinned = nullptr;
            B.pinned_bytes = 0;
            PT_HIP(hipHostMalloc(&B.pinned, dev_bytes * 2, hipHostMallocDefault));
            B.pinned_bytes = dev_bytes * 2;
        }
        dev = B.table.as<char>();
        hp = static_cast<char *>(B.pinned);
    }
    if (recording) {   // a device buffer of the recording's own; its image is uploaded below
        void *d = nullptr;
        PT_TRY(record_buffer(ctx, dev_bytes, &d));
        dev = static_cast<char *>(d);
    }
    unsigned *d_ctr = reinterpret_cast<unsigned *>(dev + o_ctr);
    unsigned *d_out = reinterpret_cast<unsigned *>(dev + o_out);
    fills.push_back({d_ctr, (unsigned long long)C::kPasses, 0u, 0u});   // the batch's tile counters
    std::vector<int> fstarts((size_t)nf + 1, 0);
    for (int e = 0; e < nf; e++)
        fstarts[e + 1] = fstarts[e] + (int)((fills[e].words + kFillWords - 1) / kFillWords);
    std::memcpy(hp, segs.data(), (size_t)ns * sizeof(BuildSeg));
    std::memcpy(hp + o_starts, starts.data(), starts.size() * sizeof(int));
    std::memcpy(hp + o_fills, fills.data(), (size_t)nf * sizeof(FillSeg));
    std::memcpy(hp + o_fstarts, fstarts.data(), fstarts.size() * sizeof(int));
    if (recording) {
        PT_TRY(record_copy(ctx, st, dev, hp, upload));
    } else {
        PT_HIP(hipMemcpyAsync(dev, hp, upload, hipMemcpyHostToDevice, st));
    }
    const BuildSeg *d_segs = reinterpret_cast<const BuildSeg *>(dev);
    const int *d_starts = reinterpret_cast<const int *>(dev + o_starts);

    if (fstarts[nf] > 0) {
        hipLaunchKernelGGL(kb_fill, dim3(fstarts[nf]), dim3(256), 0, st, reinterpret_cast<const FillSeg *>(dev + o_fills),
                           reinterpret_cast<const int *>(dev + o_fstarts), nf);
        PT_HIP(hipGetLastError());
    }
#define PT_BATCH_LAUNCH(kernel, stage, threads, ...)                                                     \
    if (total(stage) > 0) {                                                                               \
        hipLaunchKernelGGL(kernel, dim3(total(stage)), dim3(threads), 0, st, d_segs, d_starts + table(stage), \
                           ns, ##__VA_ARGS__);                                                            \
        PT_HIP(hipGetLastError());                                                                        \
    }
    PT_BATCH_LAUNCH(kb_bounds, kStBounds, kBoundsThreads);
    hipLaunchKernelGGL(kb_bounds_final, dim3(ns), dim3(kBoundsThreads), 0, st, d_segs, d_starts + table(kStBounds));
    PT_HIP(hipGetLastError());
    PT_BATCH_LAUNCH(kb_morton, kStMorton, 256);
    PT_BATCH_LAUNCH(kb_karras, kStKarras, 256);
    PT_BATCH_LAUNCH(kb_sort_hist<BITS>, kStHist, kSortThreads);
    hipLaunchKernelGGL(kb_sort_scan<BITS>, dim3(ns), dim3(kSortThreads), 0, st, d_segs);
    PT_HIP(hipGetLastError());
    static_assert(C::kPasses % 2 == 0, "the passes end in the packed buffer the first one wrote");
    for (int p = 0; p < C::kPasses; p++) {
        if (p == 0) {
            PT_BATCH_LAUNCH((kb_sort_pass<BITS, kSplit, kPacked>), kStTiles, kSortThreads, p, d_ctr + p);
        } else if (p == C::kPasses - 1) {
            PT_BATCH_LAUNCH((kb_sort_pass<BITS, kPacked, kSplit>), kStTiles, kSortThreads, p, d_ctr + p);
        } else {
            PT_BATCH_LAUNCH((kb_sort_pass<BITS, kPacked, kPacked>), kStTiles, kSortThreads, p, d_ctr + p);
        }
    }
    PT_BATCH_LAUNCH(kb_refit<true>, kStRefit0, 256, 0);
    for (int p = 1; p < kMaxRefitPasses; p++) {
        PT_BATCH_LAUNCH(kb_refit<false>, kStRefit0 + p, 256, p);
    }
    PT_BATCH_LAUNCH(kb_karras_morton, kStKMorton, 256);
    PT_BATCH_LAUNCH(kb_mrefit_chunk, kStMChunk, kMChunk);
    PT_BATCH_LAUNCH(kb_mrefit_mid, kStMSuper, 1024);
#undef PT_BATCH_LAUNCH
    if (total(kStMChunk) > 0) {
        hipLaunchKernelGGL(kb_mrefit_top, dim3(ns), dim3(1024), 0, st, d_segs);
        PT_HIP(hipGetLastError());
    }
    if (recording) {
        // nothing is read back: each handle's height stays pending, and a recorded traversal checks
        // it (with the build's error word, both in the handle's counters) on the device
        for (int s = 0; s < ns; s++) {
            ptgpu_clbvh *h = hs[s];
            h->sorted = h->valsA.as<unsigned>();
            for (bool &t : h->timed) t = false;
            clbvh_mark_built(h);
        }
        return PTGPU_OK;
    }
    hipLaunchKernelGGL(kb_gather, dim3(1), dim3(256), 0, st, d_segs, ns, d_out);
    PT_HIP(hipGetLastError());
    unsigned *h_out = reinterpret_cast<unsigned *>(hp + o_out);
    PT_HIP(hipMemcpyAsync(h_out, d_out, (size_t)2 * ns * sizeof(unsigned), hipMemcpyDeviceToHost, st));
    PT_HIP(hipStreamSynchronize(st));

    bool sort_err = false;
    for (int s = 0; s < ns; s++) {
        ptgpu_clbvh *h = hs[s];
        h->sorted = h->valsA.as<unsigned>();
        for (bool &t : h->timed) t = false;
        clbvh_mark_built(h);
        if (h->height_pending) {
            h->acc.depth = (int)h_out[s];
            h->known_height[0] = (int)h_out[s];
            h->height_pending = false;
        }
        sort_err |= h_out[ns + s] != 0;
    }
    PT_CHECK(!sort_err, PTGPU_ERR_HIP, "LBVH build failed (sort look-back spin limit or traversal-tree list "
             "overflow)");
    return PTGPU_OK;
}

}  // namespace

void BuildStaging::release() {
    table.release();
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr;
    pinned_bytes = 0;
}

}  // namespace ptgpu

using namespace ptgpu;

extern "C" {

int ptgpu_clbvh_build_batch(ptgpu_ctx *ctx, ptgpu_clbvh *const *handles, int count) {
    PT_CHECK(ctx && (handles || count == 0), PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(count >= 0, PTGPU_ERR_INVALID, "negative scene count %d", count);
    if (count == 0) return PTGPU_OK;
    std::unordered_set<const ptgpu_clbvh *> seen;
    const int bits = handles[0] ? handles[0]->digit_bits : 8;
    for (int s = 0; s < count; s++) {
        const ptgpu_clbvh *h = handles[s];
        PT_CHECK(h, PTGPU_ERR_INVALID, "handle %d is null", s);
        PT_CHECK(h->ctx == ctx, PTGPU_ERR_INVALID, "handle %d belongs to another context", s);
        PT_CHECK(h->d_shapes && h->n >= 1, PTGPU_ERR_INVALID, "handle %d has no shapes", s);
        PT_CHECK(seen.insert(h).second, PTGPU_ERR_INVALID, "handle %d appears twice in the batch", s);
        PT_CHECK(h->digit_bits == bits, PTGPU_ERR_INVALID,
                 "handle %d sorts with %d-bit digits, handle 0 with %d: one digit width per batch", s, h->digit_bits,
                 bits);
    }
    PT_HIP(hipSetDevice(ctx->device));
    return bits == 4 ? build_batch<4>(ctx, handles, count) : build_batch<8>(ctx, handles, count);
}

int ptgpu_clbvh_create_batch(ptgpu_ctx *ctx, Shape *const *d_shapes, const int *counts, int count,
                             ptgpu_clbvh **out) {
    PT_CHECK(ctx && ((d_shapes && counts && out) || count == 0), PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(count >= 0, PTGPU_ERR_INVALID, "negative scene count %d", count);
    for (int s = 0; s < count; s++) out[s] = nullptr;
    for (int s = 0; s < count; s++) {
        PT_CHECK(d_shapes[s], PTGPU_ERR_INVALID, "scene %d: null shapes", s);
        PT_CHECK(counts[s] >= 1 && counts[s] < (1 << 30), PTGPU_ERR_INVALID, "scene %d: shape count %d out of range "
                 "[1, 2^30)", s, counts[s]);
    }
    PT_HIP(hipSetDevice(ctx->device));
    int st = PTGPU_OK;
    for (int s = 0; s < count && st == PTGPU_OK; s++) {
        st = clbvh_new(ctx, counts[s], &out[s]);
        if (st == PTGPU_OK) out[s]->d_shapes = d_shapes[s];
    }
    if (st == PTGPU_OK) st = ptgpu_clbvh_build_batch(ctx, out, count);
    if (st != PTGPU_OK) {
        for (int s = 0; s < count; s++) {
            ptgpu_clbvh_destroy(out[s]);
            out[s] = nullptr;
        }
    }
    return st;
}

}  // extern "C"
