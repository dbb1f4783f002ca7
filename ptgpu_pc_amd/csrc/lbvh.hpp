// lbvh.hpp -- state of one LBVH build (the object behind the C ABI's ptgpu_clbvh handle).
#pragma once

#include <vector>

#include "common.hpp"

namespace ptgpu {
// Makes the handle's traversal layout current (tree height, four-wide nodes).  `recording`: the
// caller's stream is being recorded into a graph, so nothing may be read back.  `exact`: the caller
// needs the tree's height itself, not a bound checked on the device (exports, batch queries).
int clbvh_prepare_traversal(ptgpu_clbvh *h, bool recording = false, bool exact = false);
// Height bound of a traversal recorded before the height could be read back: the height last read
// for a build of the same kind (plain or optimised) plus this margin, or the default without one.
constexpr int kRecordedHeightMargin = 4;
constexpr int kRecordedHeightDefault = 41;   // the deepest tree the 64-entry four-wide stack walks
int clbvh_export_optimized(ptgpu_clbvh *h);

// Scratch of the Morton-keyed traversal tree: topology, parents, and the refit's crossing queue.
struct MortonScratch {
    int4 *mnode;
    int *mpar_int, *mpar_leaf;
    float *gbi;      // bounds of internal nodes whose subtree spans 1024-leaf chunks
    int *gh;         // their heights
    int2 *xlist;     // (node, child) arrivals at chunk-spanning nodes, kXCap per super-chunk
    unsigned *xcnt;  // their count per super-chunk
    int2 *xtop;      // arrivals at nodes spanning super-chunks
    unsigned xtop_cap;
    unsigned *gcnt;  // visit counters of the passes over chunk-spanning nodes
    unsigned nsuper; // super-chunks (kMSuper leaves each)
};
int clbvh_morton_scratch(ptgpu_clbvh *h, MortonScratch *m);
// Bound buffers handed between the level passes of the index-keyed refit (two halves of `per` floats).
int clbvh_level_scratch(ptgpu_clbvh *h, float **halves, size_t *per);
int clbvh_new(ptgpu_ctx *ctx, int n, ptgpu_clbvh **out);
// Handle state after a full build (bounds, Morton, topology, sort, refit) has been enqueued.
void clbvh_mark_built(ptgpu_clbvh *h);
}  // namespace ptgpu

struct ptgpu_clbvh {
    ptgpu_ctx *ctx = nullptr;
    int n = 0, cap = 0;
    int state = 0;            // 1 constructed (bounds+Morton), 2 radix tree, 3 BVH refit done
    int digit_bits = 8;       // radix-sort digit width (4 or 8)
    bool optimized = false;
    Shape *d_shapes = nullptr;
    ptgpu::DevBuf own_shapes;
    ptgpu::DevBuf boxes;                     // per-shape box records (input order) for the build's passes
    bool radix_pending = false;              // btn/btl topology not yet written (the refit writes it)
    ptgpu::DevBuf keysA, valsA, pairs;
    ptgpu::DevBuf btn, btl, tnodes, prims;
    ptgpu::DevBuf hist, status, misc;
    ptgpu::DevBuf lights, scan, opt_tmp;
    ptgpu::DevBuf mnode, mpar, gbi, xlist;   // Morton-keyed traversal tree: topology, parents, refit scratch
    ptgpu::DevBuf tnodes4;                   // four-wide view of tnodes, rebuilt after each refit
    bool wide_pending = false;
    ptgpu::DevBuf treelet;                   // treelet optimiser state (kept for the lazy export)
    bool export_pending = false;             // btn/btl still describe the tree before optimize
    int trav_topology = 1;    // traversal tree: 0 = the exported index-keyed tree, 1 = Morton-keyed LBVH
    bool height_pending = false;
    int known_height[2] = {0, 0};   // last heights read back: [0] plain build, [1] after optimize
    bool speculate = false;         // no read-back once a height is known: a bound, checked on the device
    unsigned *sorted = nullptr;
    int lights_for = -1, nlights = 0;
    ptgpu::DevAccel acc;
    std::vector<BVHNodeGPU> h_btn, h_btl;
    // misc: [0..5] scene centroid bounds, words 8.. counters ([0] unused, [1..8] sort
    // tile counters, [9] sort look-back error, [10] chunk crossings, [11] traversal-tree height),
    // floats 64.. per-block bound partials
    float *scene6() const { return misc.as<float>(); }
    unsigned *counters() const { return misc.as<unsigned>() + 8; }
    float *partials() const { return misc.as<float>() + 64; }
    // device word holding the traversal tree's height (read back lazily by clbvh_prepare_traversal)
    unsigned *height_word() const { return counters() + 11; }
    // events: 0/1 morton, 2/3 topology, 4/5 sort, 5/6 refit, 7/8 optimize
    hipEvent_t ev[11] = {};   // 0-9 stage timing; 10: the traversal layout is ready (cross-stream use)
    bool timed[5] = {false, false, false, false, false};
};
