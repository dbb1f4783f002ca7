// treelet.hip -- treelet restructuring of an LBVH (CLBVH::optimize, CLBVH.cpp:318-355).
//
// The reference's kernelOptimize (BVH.cl:514-543 with 260-509) states the intent -- Karras & Aila,
// "Fast Parallel Construction of High-Quality Bounding Volume Hierarchies" (HPG 2013): grow a
// treelet to at most 7 leaves by largest surface area, find its SAH-optimal shape by dynamic
// programming over the 2^7 leaf subsets with Ci = 1.2, Ct = 1.0 (clheader.h:14-15), and rebuild
// it -- but is incomplete and never launched.  This is an implementation of that method:
//
//   * SAH cost of a subtree: C(leaf) = Ct * A(leaf) (one primitive per leaf),
//     C(node) = Ci * A(node) + C(left) + C(right), A = surface area.
//   * Nodes are processed deepest first, one launch per depth.  Nodes of equal depth own disjoint
//     subtrees, so their treelets never overlap, and every descendant's cost is final before its
//     ancestors look at it -- the bottom-up order of the paper without device-scope atomics.
//   * Eight treelet roots per wave, eight lanes each.  The group leader grows the treelet; the
//     lanes evaluate the subset unions and the DP by subset size; the leader rebuilds the treelet
//     from the DP choices, reusing the treelet's own internal nodes.  A treelet is rewritten only
//     when its optimal cost is lower than its current cost, so the tree's SAH cost never increases.
//   * A node record holds its children's boxes and subtree costs, so growing a treelet costs one
//     dependent fetch per expanded node, and a processed node hands its cost up by writing it into
//     its parent's record.
//
// The traversal layout (TNode, and the four-wide view) is rebuilt from the result; the exported
// btn/btl arrays are rewritten in the reference layout on first request (leaves renumbered in
// depth-first order, internal nodes in Karras numbering: a node over leaves [a, b] with split s has
// children s and s + 1).
#include <algorithm>
#include <cstring>
#include <vector>

#include "device_util.hpp"
#include "lbvh.hpp"

namespace ptgpu {

constexpr float kCi = 1.2f;     // clheader.h:14
constexpr float kCt = 1.0f;     // clheader.h:15
constexpr int kMaxTreelet = 7;
constexpr int kMaxLevels = 4096;   // trees taller than this are not restructured

struct OptNode {
    float cb[2][6];   // child boxes (lo xyz, hi xyz)
    float cc[2];      // child subtree costs: a leaf's from the start, an internal child's once its depth is done
    int c[2];         // >= 0 internal node, < 0 leaf ~k (k = primitive slot in sorted order)
    int parent;
    int depth;
    int pad[2];
};
static_assert(sizeof(OptNode) == 80, "OptNode");

__device__ __host__ inline void box_union(const float a[6], const float b[6], float out[6]) {
    for (int k = 0; k < 3; k++) {
        out[k] = a[k] < b[k] ? a[k] : b[k];
        out[3 + k] = a[3 + k] > b[3 + k] ? a[3 + k] : b[3 + k];
    }
}

__device__ __forceinline__ float surface_area(const float lo[3], const float hi[3]) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.f * (dx * dy + dy * dz + dz * dx);
}

// Optimiser state from the traversal layout: child refs, child boxes, leaf costs, parents.
__global__ __launch_bounds__(256) void k_opt_init(const TNode *__restrict__ tn, int m, OptNode *on) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const TNode t = tn[i];
    const float b0[6] = {t.a.x, t.a.y, t.a.z, t.a.w, t.b.x, t.b.y};   // lo xyz, hi xyz
    const float b1[6] = {t.b.z, t.b.w, t.c.x, t.c.y, t.c.z, t.c.w};
    OptNode &o = on[i];
    if (i == 0) o.parent = -1;
    for (int side = 0; side < 2; side++) {
        const int r = side ? t.r.y : t.r.x;
        const float *b = side ? b1 : b0;
        for (int f = 0; f < 6; f++) o.cb[side][f] = b[f];
        o.c[side] = r;
        o.cc[side] = r < 0 ? kCt * surface_area(b, b + 3) : 0.f;
        if (r >= 0 && r < m) on[r].parent = i;
    }
}

// Depth of every internal node (walk to the root) and the deepest one (one atomic per wave).
__global__ __launch_bounds__(256) void k_opt_depth(OptNode *on, int m, unsigned *max_depth) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned d = 0;
    if (i < m) {
        int p = on[i].parent;
        while (p >= 0 && d <= (unsigned)m) {
            d++;
            p = on[p].parent;
        }
        on[i].depth = (int)d;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned v = __shfl_down(d, o, kWave);
        d = v > d ? v : d;
    }
    if (__lane_id() == 0) atomicMax(max_depth, d);
}

// tree height after optimisation: deepest internal depth + 1 (the deepest leaf), for the lazy readback
__global__ void k_opt_height(const unsigned *max_depth, unsigned *height) { *height = *max_depth + 1u; }

// Treelets of one depth.  A workgroup is one wave of 64 lanes in 8 groups of 8: it scans 64-node
// chunks for the level's roots and restructures them 8 at a time, one per group.  The group leader
// grows the treelet (a chain of dependent loads); the group's lanes share the subset unions and the
// DP by subset size; the leader rebuilds.  Barriers stay wave-uniform: every group walks the same
// subset-size loop whatever its treelet's size.
constexpr int kGroup = 8;
constexpr int kGroups = 64 / kGroup;

__global__ __launch_bounds__(64) void k_treelet(const OptNode *depth_src, int m, int depth, int tsize, OptNode *on) {
    __shared__ float s_area[kGroups][1 << kMaxTreelet];
    __shared__ float s_cost[kGroups][1 << kMaxTreelet];
    __shared__ unsigned char s_best[kGroups][1 << kMaxTreelet];
    __shared__ float s_box[kGroups][kMaxTreelet][6];
    __shared__ int s_ref[kGroups][kMaxTreelet];
    __shared__ int s_int[kGroups][kMaxTreelet];
    __shared__ int s_nl[kGroups];
    __shared__ float s_cur[kGroups];
    const int lane = threadIdx.x, grp = lane / kGroup, sub = lane % kGroup;
    // a treelet leaf j: ref, box and subtree cost, taken from the record of its parent
    auto put_leaf = [&](int j, int ref, const float *box, float cost) {
        s_ref[grp][j] = ref;
        for (int f = 0; f < 6; f++) s_box[grp][j][f] = box[f];
        s_cost[grp][1 << j] = cost;
    };

    for (int base = blockIdx.x * 64; base < m; base += gridDim.x * 64) {
        const int i = base + lane;
        unsigned long long hit = __ballot(i < m && depth_src[i].depth == depth);
        while (hit) {
            // the next (up to) 8 roots of this chunk, one per group
            int root = -1;
            {
                unsigned long long h = hit;
                for (int g = 0; g < kGroups && h; g++) {
                    const int j = __ffsll((long long)h) - 1;
                    h &= h - 1;
                    if (g == grp) root = base + j;
                }
                hit = h;
            }
            if (root >= 0 && sub == 0) {
                // grow: replace the treelet leaf of largest area that is an internal node (one fetch:
                // its record carries its children's boxes and costs)
                int inl[kMaxTreelet], nl = 2, ni = 1;
                float ar[kMaxTreelet];
                inl[0] = root;
                const OptNode R = on[root];
                for (int j = 0; j < 2; j++) {
                    put_leaf(j, R.c[j], R.cb[j], R.cc[j]);
                    ar[j] = R.c[j] >= 0 ? surface_area(R.cb[j], R.cb[j] + 3) : -1.f;
                }
                while (nl < tsize) {
                    int best = -1;
                    float best_a = -1.f;
                    for (int j = 0; j < nl; j++)
                        if (ar[j] > best_a) {
                            best_a = ar[j];
                            best = j;
                        }
                    if (best < 0) break;
                    const int node = s_ref[grp][best];
                    inl[ni++] = node;
                    const OptNode N = on[node];
                    put_leaf(best, N.c[0], N.cb[0], N.cc[0]);
                    put_leaf(nl, N.c[1], N.cb[1], N.cc[1]);
                    ar[best] = N.c[0] >= 0 ? surface_area(N.cb[0], N.cb[0] + 3) : -1.f;
                    ar[nl] = N.c[1] >= 0 ? surface_area(N.cb[1], N.cb[1] + 3) : -1.f;
                    nl++;
                }
                float rb[6];
                box_union(R.cb[0], R.cb[1], rb);
                s_cur[grp] = kCi * surface_area(rb, rb + 3) + R.cc[0] + R.cc[1];   // the current shape's cost
                for (int j = 0; j < ni; j++) s_int[grp][j] = inl[j];
                s_nl[grp] = nl;
            }
            if (root < 0 && sub == 0) s_nl[grp] = 0;
            __syncthreads();
            const int nl = s_nl[grp];
            const int full = nl >= 3 ? (1 << nl) - 1 : 0;
            // areas of every subset's union box
            for (int s = sub + 1; s <= full; s += kGroup) {
                float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
                float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
                for (int j = 0; j < nl; j++) {
                    if (!((s >> j) & 1)) continue;
                    for (int a = 0; a < 3; a++) {
                        lo[a] = fminf(lo[a], s_box[grp][j][a]);
                        hi[a] = fmaxf(hi[a], s_box[grp][j][3 + a]);
                    }
                }
                s_area[grp][s] = surface_area(lo, hi);
            }
            __syncthreads();
            // optimal cost by subset size; a split is named by the side holding the lowest member
            for (int k = 2; k <= kMaxTreelet; k++) {
                for (int s = sub + 1; s <= full; s += kGroup) {
                    if (__popc((unsigned)s) != k) continue;
                    const int low = s & -s;
                    float best = __builtin_inff();
                    int bp = low;
                    for (int p = (s - 1) & s; p; p = (p - 1) & s) {
                        if (!(p & low)) continue;
                        const float c = s_cost[grp][p] + s_cost[grp][s ^ p];
                        if (c < best) {
                            best = c;
                            bp = p;
                        }
                    }
                    s_cost[grp][s] = kCi * s_area[grp][s] + best;
                    s_best[grp][s] = (unsigned char)bp;
                }
                __syncthreads();
            }
            if (root >= 0 && sub == 0) {
                float root_cost = s_cur[grp];
                if (nl >= 3 && s_cost[grp][full] < s_cur[grp] * (1.f - 1e-6f)) {
                    root_cost = s_cost[grp][full];
                    // rebuild top-down, reusing the treelet's internal nodes (the root keeps its slot)
                    int st_set[kMaxTreelet], st_node[kMaxTreelet], sp = 0, next = 1;
                    st_set[sp] = full;
                    st_node[sp++] = root;
                    while (sp > 0) {
                        --sp;
                        const int S = st_set[sp], node = st_node[sp];
                        const int parts[2] = {s_best[grp][S], S ^ s_best[grp][S]};
                        OptNode &o = on[node];
                        for (int side = 0; side < 2; side++) {
                            const int X = parts[side];
                            int cref;
                            if (__popc((unsigned)X) == 1) {
                                cref = s_ref[grp][__ffs(X) - 1];
                            } else {
                                cref = s_int[grp][next++];
                                st_set[sp] = X;
                                st_node[sp++] = cref;
                            }
                            if (cref >= 0) on[cref].parent = node;
                            float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
                            float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
                            for (int j = 0; j < nl; j++) {
                                if (!((X >> j) & 1)) continue;
                                for (int a = 0; a < 3; a++) {
                                    lo[a] = fminf(lo[a], s_box[grp][j][a]);
                                    hi[a] = fmaxf(hi[a], s_box[grp][j][3 + a]);
                                }
                            }
                            for (int a = 0; a < 3; a++) {
                                o.cb[side][a] = lo[a];
                                o.cb[side][3 + a] = hi[a];
                            }
                            o.c[side] = cref;
                            o.cc[side] = s_cost[grp][X];
                        }
                    }
                }
                // hand the subtree's cost to the parent's record (siblings write different slots)
                const int P = on[root].parent;
                if (P >= 0) on[P].cc[on[P].c[0] == root ? 0 : 1] = root_cost;
            }
            __syncthreads();
        }
    }
}

// Traversal layout from the optimised tree.
__global__ __launch_bounds__(256) void k_opt_emit(const OptNode *__restrict__ on, int m, TNode *tn) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const OptNode &o = on[i];
    const float *b0 = o.cb[0], *b1 = o.cb[1];   // lo xyz, hi xyz
    TNode t;
    t.a = make_float4(b0[0], b0[1], b0[2], b0[3]);
    t.b = make_float4(b0[4], b0[5], b1[0], b1[1]);
    t.c = make_float4(b1[2], b1[3], b1[4], b1[5]);
    t.r = make_int4(o.c[0], o.c[1], o.c[0] < 0 ? 1 : 0, o.c[1] < 0 ? 1 : 0);
    tn[i] = t;
}

namespace {

struct OptLayout {
    OptNode *nodes;
    unsigned *maxd;
};

int opt_layout(ptgpu_clbvh *h, OptLayout *L) {
    const size_t n = (size_t)h->n, m = n - 1;
    PT_TRY(h->treelet.ensure(m * sizeof(OptNode) + 256));
    char *p = static_cast<char *>(h->treelet.p);
    L->nodes = reinterpret_cast<OptNode *>(p);
    L->maxd = reinterpret_cast<unsigned *>(p + m * sizeof(OptNode));
    return PTGPU_OK;
}

}  // namespace

int clbvh_optimize(ptgpu_clbvh *h, int tsize) {
    const int n = h->n, m = n - 1;
    if (m < 2) return PTGPU_OK;   // at most two leaves: one shape only
    PT_TRY(clbvh_prepare_traversal(h));   // current height: the number of levels to visit
    const int levels = h->acc.depth;      // internal depths are 0 .. height - 1
    PT_CHECK(levels >= 1 && levels < kMaxLevels, PTGPU_ERR_UNSUPPORTED, "tree height %d", levels);
    hipStream_t st = h->ctx->stream;
    OptLayout L;
    PT_TRY(opt_layout(h, &L));
    const dim3 g((m + 255) / 256), b(256);
    hipLaunchKernelGGL(k_opt_init, g, b, 0, st, h->tnodes.as<TNode>(), m, L.nodes);
    PT_TRY(fill_u32(st, L.maxd, 0u, 1));
    hipLaunchKernelGGL(k_opt_depth, g, b, 0, st, L.nodes, m, L.maxd);
    PT_HIP(hipGetLastError());
    // persistent: each workgroup scans node chunks for the level's roots
    int grid = h->ctx->cu_count * 32;
    if (grid > (m + 63) / 64) grid = (m + 63) / 64;
    for (int d = levels - 1; d >= 0; d--)
        hipLaunchKernelGGL(k_treelet, dim3(grid), dim3(64), 0, st, L.nodes, m, d, tsize, L.nodes);
    PT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_opt_emit, g, b, 0, st, L.nodes, m, h->tnodes.as<TNode>());
    PT_TRY(fill_u32(st, L.maxd, 0u, 1));
    hipLaunchKernelGGL(k_opt_depth, g, b, 0, st, L.nodes, m, L.maxd);
    hipLaunchKernelGGL(k_opt_height, dim3(1), dim3(1), 0, st, L.maxd, h->height_word());
    PT_HIP(hipGetLastError());
    h->height_pending = true;   // new height read back before the next traversal
    h->acc.nodes4 = nullptr;
    h->acc.nodes4h = nullptr;
    h->wide_pending = true;
    h->export_pending = true;
    return PTGPU_OK;
}

// Reference-layout arrays for the optimised tree (host renumbering, on first request).
int clbvh_export_optimized(ptgpu_clbvh *h) {
    if (!h->export_pending) return PTGPU_OK;
    const int n = h->n, m = n - 1;
    hipStream_t st = h->ctx->stream;
    OptLayout L;
    PT_TRY(opt_layout(h, &L));
    std::vector<OptNode> on((size_t)m);
    std::vector<unsigned> sorted((size_t)n);
    PT_HIP(hipMemcpyAsync(on.data(), L.nodes, (size_t)m * sizeof(OptNode), hipMemcpyDeviceToHost, st));
    PT_HIP(hipMemcpyAsync(sorted.data(), h->sorted, (size_t)n * sizeof(unsigned), hipMemcpyDeviceToHost, st));
    PT_HIP(hipStreamSynchronize(st));
    // leaf counts per internal node (children before parents: reverse pre-order)
    std::vector<int> pre;
    pre.reserve((size_t)m);
    std::vector<int> stack{0};
    while (!stack.empty()) {
        const int v = stack.back();
        stack.pop_back();
        pre.push_back(v);
        if (on[v].c[1] >= 0) stack.push_back(on[v].c[1]);
        if (on[v].c[0] >= 0) stack.push_back(on[v].c[0]);
    }
    PT_CHECK((int)pre.size() == m, PTGPU_ERR_HIP, "optimised tree is not a tree (%zu of %d nodes reached)",
             pre.size(), m);
    std::vector<int> size((size_t)m, 0);
    for (auto it = pre.rbegin(); it != pre.rend(); ++it) {
        const OptNode &o = on[*it];
        size[*it] = (o.c[0] < 0 ? 1 : size[o.c[0]]) + (o.c[1] < 0 ? 1 : size[o.c[1]]);
    }
    // CLBVH.cpp:83-88 conventions: internal nodes zeroed, leaves 0xFF bytes; then the fields
    std::vector<BVHNodeGPU> btn((size_t)m), btl((size_t)n);
    std::memset(static_cast<void *>(btn.data()), 0, btn.size() * sizeof(BVHNodeGPU));
    std::memset(static_cast<void *>(btl.data()), 0xFF, btl.size() * sizeof(BVHNodeGPU));
    auto set_bound = [](BVHNodeGPU &d, const float *lo, const float *hi) {
        d.bound.min_x = lo[0]; d.bound.max_x = hi[0];
        d.bound.min_y = lo[1]; d.bound.max_y = hi[1];
        d.bound.min_z = lo[2]; d.bound.max_z = hi[2];
    };
    struct Item { int node, a, idx, parent; };
    std::vector<Item> work{{0, 0, 0, 0}};
    while (!work.empty()) {
        const Item w = work.back();
        work.pop_back();
        const OptNode &o = on[w.node];
        const int b = w.a + size[w.node] - 1;
        const int s = w.a + (o.c[0] < 0 ? 1 : size[o.c[0]]) - 1;   // last leaf of the left child
        BVHNodeGPU &d = btn[w.idx];
        d.min = w.a;
        d.max = b;
        d.nLeft = s;
        d.nRight = s + 1;
        d.nParent = w.parent;
        d.leaf = (o.c[0] < 0 || o.c[1] < 0) ? 2 : 0;
        float nb[6];
        box_union(o.cb[0], o.cb[1], nb);
        set_bound(d, nb, nb + 3);
        const int kids[2] = {o.c[0], o.c[1]};
        const int first[2] = {w.a, s + 1};
        const int slot[2] = {s, s + 1};
        for (int side = 0; side < 2; side++) {
            const int r = kids[side];
            if (r < 0) {
                BVHNodeGPU &l = btl[slot[side]];
                set_bound(l, o.cb[side], o.cb[side] + 3);
                l.nShape = (int)sorted[~r];
                l.nParent = w.idx;
            } else {
                work.push_back({r, first[side], slot[side], w.idx});
            }
        }
    }
    PT_HIP(hipMemcpyAsync(h->btn.p, btn.data(), btn.size() * sizeof(BVHNodeGPU), hipMemcpyHostToDevice, st));
    PT_HIP(hipMemcpyAsync(h->btl.p, btl.data(), btl.size() * sizeof(BVHNodeGPU), hipMemcpyHostToDevice, st));
    PT_HIP(hipStreamSynchronize(st));
    h->export_pending = false;
    return PTGPU_OK;
}

}  // namespace ptgpu

using namespace ptgpu;

extern "C" int ptgpu_clbvh_optimize(ptgpu_clbvh *h, int treelet_size) {
    PT_CHECK(h, PTGPU_ERR_INVALID, "null handle");
    PT_CHECK(h->state >= 3, PTGPU_ERR_INVALID, "build_bvh_tree has not run");
    PT_CHECK(treelet_size >= 3 && treelet_size <= kMaxTreelet, PTGPU_ERR_INVALID, "treelet size %d not in [3, %d]",
             treelet_size, kMaxTreelet);
    PT_HIP(hipSetDevice(h->ctx->device));
    PT_HIP(hipEventRecord(h->ev[7], h->ctx->stream));
    PT_TRY(clbvh_optimize(h, treelet_size));
    PT_HIP(hipEventRecord(h->ev[8], h->ctx->stream));
    h->timed[4] = true;
    h->optimized = true;
    return PTGPU_OK;
}
