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
// depth of every node (and, with hist, the number of nodes per depth)
constexpr int kLocalDepths = 64;   // depths counted per block in LDS (deeper ones: global atomics)
__global__ __launch_bounds__(256) void k_opt_depth(OptNode *on, int m, unsigned *max_depth, unsigned *hist) {
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
    if (hist) {
        // block-local counts first: tens of depths, so one global add per depth per block instead of
        // one per node on a handful of contended words
        __shared__ unsigned lh[kLocalDepths];
        if (threadIdx.x < kLocalDepths) lh[threadIdx.x] = 0;
        __syncthreads();
        if (i < m) {
            if (d < (unsigned)kLocalDepths) atomicAdd(lh + d, 1u);
            else if (d < (unsigned)kMaxLevels) atomicAdd(hist + d, 1u);
        }
        __syncthreads();
        if (threadIdx.x < kLocalDepths && lh[threadIdx.x]) atomicAdd(hist + threadIdx.x, lh[threadIdx.x]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned v = __shfl_down(d, o, kWave);
        d = v > d ? v : d;
    }
    if (__lane_id() == 0) atomicMax(max_depth, d);
}

// tree height after optimisation: deepest internal depth + 1 (the deepest leaf), for the lazy readback
__global__ void k_opt_height(const unsigned *max_depth, unsigned *height) { *height = *max_depth + 1u; }

// Per-depth node lists: dstart = exclusive prefix sums of the depth histogram (one thread; a tree has
// tens of depths), then each node takes a slot of its depth's range.  Order within a depth is free:
// the treelets rooted at one depth are disjoint.
__global__ void k_opt_scan(const unsigned *hist, int levels, unsigned *dstart, unsigned *cursor) {
    unsigned run = 0;
    for (int d = 0; d < levels; d++) {
        dstart[d] = run;
        cursor[d] = 0;
        run += hist[d];
    }
    dstart[levels] = run;
}
__global__ __launch_bounds__(256) void k_opt_scatter(const OptNode *on, int m, int levels, const unsigned *dstart,
                                                     unsigned *cursor, int *order) {
    // block-local ranks per depth, then one global reservation per depth per block
    __shared__ unsigned lc[kLocalDepths], lbase[kLocalDepths];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x < kLocalDepths) lc[threadIdx.x] = 0;
    __syncthreads();
    const int d = i < m ? on[i].depth : -1;
    const bool listed = d >= 0 && d < levels;   // depths >= levels: not listed
    unsigned r = 0;
    if (listed && d < kLocalDepths) r = atomicAdd(lc + d, 1u);
    __syncthreads();
    if (threadIdx.x < kLocalDepths && lc[threadIdx.x]) lbase[threadIdx.x] = atomicAdd(cursor + threadIdx.x, lc[threadIdx.x]);
    __syncthreads();
    if (listed) order[dstart[d] + (d < kLocalDepths ? lbase[d] + r : atomicAdd(cursor + d, 1u))] = i;
}

// Treelets of depths d_hi down to d_lo, from the per-depth node lists.  A workgroup of WAVES waves is
// 8 * WAVES groups of 8 lanes, one treelet per group at a time.  The group leader grows the treelet
// (a chain of dependent loads); the group's lanes share the subset unions and the DP by subset size;
// the leader rebuilds.  Barriers stay block-uniform: every group walks the same subset-size loop
// whatever its treelet's size.  A launch over several depths is one workgroup (the top of the tree,
// at most 2^d nodes per depth d): its barrier orders one depth's results before the next.
constexpr int kGroup = 8;
constexpr int kGroups = 64 / kGroup;

template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_treelet(const int *__restrict__ order,
                                                        const unsigned *__restrict__ dstart, int d_hi, int d_lo,
                                                        int tsize, OptNode *on) {
    constexpr int NG = kGroups * WAVES;
    __shared__ float s_area[NG][1 << kMaxTreelet];
    __shared__ float s_cost[NG][1 << kMaxTreelet];
    __shared__ unsigned char s_best[NG][1 << kMaxTreelet];
    __shared__ float s_box[NG][kMaxTreelet][6];
    __shared__ int s_ref[NG][kMaxTreelet];
    __shared__ int s_int[NG][kMaxTreelet];
    __shared__ int s_nl[NG];
    __shared__ float s_cur[NG];
    const int lane = threadIdx.x, grp = lane / kGroup, sub = lane % kGroup;
    // a treelet leaf j: ref, box and subtree cost, taken from the record of its parent
    auto put_leaf = [&](int j, int ref, const float *box, float cost) {
        s_ref[grp][j] = ref;
        for (int f = 0; f < 6; f++) s_box[grp][j][f] = box[f];
        s_cost[grp][1 << j] = cost;
    };

    for (int d = d_hi; d >= d_lo; d--) {
        const int lo = (int)dstart[d], cnt = (int)dstart[d + 1] - lo;
        for (int base = blockIdx.x * NG; base < cnt; base += gridDim.x * NG) {
            const int root = base + grp < cnt ? order[lo + base + grp] : -1;
            int par = -1, pslot = 0;   // the root's parent and the slot of the root in it (leader only)
            if (root >= 0 && sub == 0) {
                // grow: replace the treelet leaf of largest area that is an internal node (one fetch:
                // its record carries its children's boxes and costs)
                int inl[kMaxTreelet], nl = 2, ni = 1;
                float ar[kMaxTreelet];
                inl[0] = root;
                const OptNode R = on[root];
                par = R.parent;
                // the parent sits one depth up and is not rewritten at this depth: its child slot can
                // be read now, beside the growth loads, instead of after the rebuild
                if (par >= 0) pslot = on[par].c[0] == root ? 0 : 1;
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
            // the treelet's leaf boxes in registers (every lane): subset unions and the rebuild read them
            float bx[kMaxTreelet][6];
#pragma unroll
            for (int j = 0; j < kMaxTreelet; j++)
#pragma unroll
                for (int f = 0; f < 6; f++) bx[j][f] = j < nl ? s_box[grp][j][f] : 0.f;
            auto union_of = [&](int X, float *lo, float *hi) {
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    lo[a] = __builtin_inff();
                    hi[a] = -__builtin_inff();
                }
#pragma unroll
                for (int j = 0; j < kMaxTreelet; j++) {
                    if (!((X >> j) & 1)) continue;
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        lo[a] = fminf(lo[a], bx[j][a]);
                        hi[a] = fmaxf(hi[a], bx[j][3 + a]);
                    }
                }
            };
            // areas of every subset's union box
            for (int s = sub + 1; s <= full; s += kGroup) {
                float lo[3], hi[3];
                union_of(s, lo, hi);
                s_area[grp][s] = surface_area(lo, hi);
            }
            __syncthreads();
            // optimal cost by subset size; a split is named by the side holding the lowest member.
            // The splits of s are p = low | q for q a proper subset of the rest, visited in
            // descending order, two at a time so that their cost loads overlap; the first minimum
            // wins, as in a one-at-a-time scan.
            for (int k = 2; k <= tsize; k++) {
                for (int s = sub + 1; s <= full; s += kGroup) {
                    if (__popc((unsigned)s) != k) continue;
                    const int low = s & -s, rest = s ^ low;
                    float best = __builtin_inff();
                    int bp = low;
                    int q = (rest - 1) & rest;   // rest itself would be p == s
                    bool more = rest != 0;
                    while (more) {
                        const int qa = q, qb = (q - 1) & rest;
                        const bool hb = qa != 0;
                        const int pa = low | qa, pb = low | qb;
                        const float ca = s_cost[grp][pa] + s_cost[grp][s ^ pa];
                        const float cb = hb ? s_cost[grp][pb] + s_cost[grp][s ^ pb] : __builtin_inff();
                        if (ca < best) {
                            best = ca;
                            bp = pa;
                        }
                        if (hb && cb < best) {
                            best = cb;
                            bp = pb;
                        }
                        more = hb && qb != 0;
                        q = (qb - 1) & rest;
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
                            float lo[3], hi[3];
                            union_of(X, lo, hi);
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
                if (par >= 0) on[par].cc[pslot] = root_cost;
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
    unsigned *hist, *dstart, *cursor;   // per-depth counts, list starts (levels + 1), fill cursors
    int *order;                         // node indices grouped by depth
};

int opt_layout(ptgpu_clbvh *h, OptLayout *L) {
    const size_t n = (size_t)h->n, m = n - 1;
    const size_t lv = (size_t)kMaxLevels + 1;
    PT_TRY(h->treelet.ensure(m * sizeof(OptNode) + 256 + 3 * lv * 4 + m * 4));
    char *p = static_cast<char *>(h->treelet.p);
    L->nodes = reinterpret_cast<OptNode *>(p);
    p += m * sizeof(OptNode);
    L->maxd = reinterpret_cast<unsigned *>(p);
    p += 256;
    L->hist = reinterpret_cast<unsigned *>(p);
    L->dstart = L->hist + lv;
    L->cursor = L->dstart + lv;
    L->order = reinterpret_cast<int *>(L->cursor + lv);
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
    PT_TRY(fill_u32(st, L.hist, 0u, (size_t)levels));
    hipLaunchKernelGGL(k_opt_depth, g, b, 0, st, L.nodes, m, L.maxd, L.hist);
    hipLaunchKernelGGL(k_opt_scan, dim3(1), dim3(1), 0, st, (const unsigned *)L.hist, levels, L.dstart, L.cursor);
    hipLaunchKernelGGL(k_opt_scatter, g, b, 0, st, (const OptNode *)L.nodes, m, levels, (const unsigned *)L.dstart,
                       L.cursor, L.order);
    PT_HIP(hipGetLastError());
    // deepest first, one launch per depth over that depth's list; depths 0..4 (at most 2^d <= 16
    // nodes each) in one workgroup launch with a barrier between depths, one round per depth.  That
    // workgroup is 4 waves (32 treelets at a time, 44 KB of LDS): small enough to run beside a
    // render's resident workgroups.
    const int dtop = levels < 5 ? levels : 5;
    int grid = h->ctx->cu_count * 16;
    if (grid > (m + kGroups - 1) / kGroups) grid = (m + kGroups - 1) / kGroups;
    for (int d = levels - 1; d >= dtop; d--)
        hipLaunchKernelGGL(k_treelet<1>, dim3(grid), dim3(64), 0, st, (const int *)L.order,
                           (const unsigned *)L.dstart, d, d, tsize, L.nodes);
    hipLaunchKernelGGL(k_treelet<4>, dim3(1), dim3(256), 0, st, (const int *)L.order, (const unsigned *)L.dstart,
                       dtop - 1, 0, tsize, L.nodes);
    PT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_opt_emit, g, b, 0, st, L.nodes, m, h->tnodes.as<TNode>());
    PT_TRY(fill_u32(st, L.maxd, 0u, 1));
    hipLaunchKernelGGL(k_opt_depth, g, b, 0, st, L.nodes, m, L.maxd, (unsigned *)nullptr);
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
