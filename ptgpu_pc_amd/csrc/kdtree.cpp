// kdtree.cpp -- KD tree (class KDTree, KDTree.h:47-90 / KDTree.cpp:89-270) for the gfx950 engine.
//
// The reference's "KD tree" is an object partition: each node splits its shapes at the mean of a
// per-shape "midpoint" along the longest axis of the node box, and each shape lands on exactly one
// side.  The parity builder below reproduces its decisions exactly -- same midpoint definition
// (for a triangle the comma expression (p1.x, p2.x, p3.x) / 3.0f evaluates to p3.x / 3.0f,
// KDTree.cpp:55-57), same mean (float reciprocal of the count, sequential float sum that skips the
// first shape, KDTree.cpp:121-135), same tie rule (midpoint <= mean goes right), same leaf rule.  It
// works on index arrays with a stable partition instead of copying std::vector<Shape*> per node,
// and keeps shapes within a node in the reference's order, which the next level's mean depends on.
//
// Device layout: internal nodes as TNode (depth-first order), leaves as ranges of Prim.  int32
// indices throughout; the reference's int16 KDNodeGPU heap is produced only on request and only
// when it fits (Q13).  A binned-SAH builder with a depth cap is provided as an extension.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.hpp"

using namespace ptgpu;

namespace {

struct KNode {
    int leaf = 0;
    int left = -1, right = -1;   // node indices (depth-first order)
    int start = 0, count = 0;    // leaves: range in the primitive list
    int depth = 0;
    Bound box{};
};

inline float fmin2(float a, float b) { return (a < b) ? a : b; }
inline float fmax2(float a, float b) { return (a > b) ? a : b; }

Bound shape_bound_host(const Shape &s) {   // KDTree::getBound KDTree.cpp:8-44
    Bound b;
    if (s.type == TRIANGLE) {
        b.min_x = fmin2(fmin2(s.t.p1.x, s.t.p2.x), s.t.p3.x);
        b.max_x = fmax2(fmax2(s.t.p1.x, s.t.p2.x), s.t.p3.x);
        b.min_y = fmin2(fmin2(s.t.p1.y, s.t.p2.y), s.t.p3.y);
        b.max_y = fmax2(fmax2(s.t.p1.y, s.t.p2.y), s.t.p3.y);
        b.min_z = fmin2(fmin2(s.t.p1.z, s.t.p2.z), s.t.p3.z);
        b.max_z = fmax2(fmax2(s.t.p1.z, s.t.p2.z), s.t.p3.z);
    } else {
        b.min_x = s.s.p.x - s.s.rad; b.max_x = s.s.p.x + s.s.rad;
        b.min_y = s.s.p.y - s.s.rad; b.max_y = s.s.p.y + s.s.rad;
        b.min_z = s.s.p.z - s.s.rad; b.max_z = s.s.p.z + s.s.rad;
    }
    return b;
}

Vec midpoint_host(const Shape &s) {   // KDTree::getMidpoint KDTree.cpp:46-70
    Vec v;
    if (s.type == TRIANGLE) {
        v.x = s.t.p3.x / 3.0f; v.y = s.t.p3.y / 3.0f; v.z = s.t.p3.z / 3.0f;
    } else {
        v = s.s.p;
    }
    return v;
}

inline void expand(Bound &box, const Bound &b) {   // KDTreeNode::expandBound KDTree.h:35-44
    if (b.min_x < box.min_x) box.min_x = b.min_x;
    if (b.min_y < box.min_y) box.min_y = b.min_y;
    if (b.min_z < box.min_z) box.min_z = b.min_z;
    if (b.max_x > box.max_x) box.max_x = b.max_x;
    if (b.max_y > box.max_y) box.max_y = b.max_y;
    if (b.max_z > box.max_z) box.max_z = b.max_z;
}

inline int longest_axis(const Bound &b) {   // KDTree::getLongestAxis KDTree.cpp:72-86
    float xd = b.max_x - b.min_x, yd = b.max_y - b.min_y, zd = b.max_z - b.min_z;
    if (xd >= yd) return xd >= zd ? 0 : 2;
    return yd >= zd ? 1 : 2;
}

inline float comp(const Vec &v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

}  // namespace

struct ptgpu_kdtree {
    ptgpu_ctx *ctx = nullptr;
    int n = 0;
    int maxdepth = 0;
    long long szbuf = 0;          // the reference's m_szbuf (kn length)
    std::vector<KNode> nodes;     // depth-first pre-order
    std::vector<int> kn;          // shape indices, leaves in depth-first order
    std::vector<Shape> shapes;    // host copy (for the heap export bounds and prims)
    DevBuf d_nodes, d_nodes4, d_prims;
    DevBuf lights, scan;              // light list cache (index order), for lightCnt == lights_for
    int lights_for = -1, nlights = 0;
    DevAccel acc;
    bool sah = false;
};

namespace {

struct Work {
    int parent;      // node index of the parent, -1 for the root
    int side;        // 0 left, 1 right
    int lo, hi;      // range in the work index array
    int depth;
};

// parity builder (KDTree::build)
void build_mean_split(ptgpu_kdtree *T, const std::vector<Bound> &bnd, const std::vector<Vec> &mid) {
    const int n = T->n;
    std::vector<int> idx(n), tmp(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    std::vector<Work> st;
    st.push_back({-1, 0, 0, n, 0});
    while (!st.empty()) {
        Work w = st.back();
        st.pop_back();
        const int me = (int)T->nodes.size();
        T->nodes.emplace_back();
        if (w.parent >= 0) (w.side ? T->nodes[w.parent].right : T->nodes[w.parent].left) = me;
        KNode &nd = T->nodes[me];
        nd.depth = w.depth;
        if (w.depth > T->maxdepth) T->maxdepth = w.depth;
        const int cnt = w.hi - w.lo;
        if (cnt == 0) continue;
        auto make_leaf = [&](KNode &leafnode) {
            leafnode.leaf = 1;
            leafnode.box = bnd[idx[w.lo]];
            for (int i = w.lo + 1; i < w.hi; i++) expand(leafnode.box, bnd[idx[i]]);
            leafnode.start = (int)T->kn.size();
            leafnode.count = cnt;
            for (int i = w.lo; i < w.hi; i++) T->kn.push_back(idx[i]);
            T->szbuf += cnt;
        };
        if (cnt == 1) {
            make_leaf(nd);
            continue;
        }
        const float recp = (float)(1.0 / (double)cnt);
        Vec m = {0.f, 0.f, 0.f};
        Bound box = bnd[idx[w.lo]];
        for (int i = w.lo + 1; i < w.hi; i++) {
            expand(box, bnd[idx[i]]);
            Vec v = mid[idx[i]];
            v.x *= recp; v.y *= recp; v.z *= recp;
            m.x += v.x; m.y += v.y; m.z += v.z;
        }
        nd.box = box;
        const int axis = longest_axis(box);
        const float mv = comp(m, axis);
        int nl = 0, nr = 0;
        for (int i = w.lo; i < w.hi; i++)
            if (!(mv >= comp(mid[idx[i]], axis))) nl++;
        nr = cnt - nl;
        if (nl == 0 || nr == 0) {
            make_leaf(nd);
            continue;
        }
        // stable partition: left part keeps input order, right part keeps input order
        int pl = w.lo, pr = w.lo + nl;
        for (int i = w.lo; i < w.hi; i++) {
            const int s = idx[i];
            if (mv >= comp(mid[s], axis)) tmp[pr++] = s;
            else tmp[pl++] = s;
        }
        std::copy(tmp.begin() + w.lo, tmp.begin() + w.hi, idx.begin() + w.lo);
        st.push_back({me, 1, w.lo + nl, w.hi, w.depth + 1});
        st.push_back({me, 0, w.lo, w.lo + nl, w.depth + 1});
    }
}

// binned SAH object partition with a depth cap (extension)
void build_sah(ptgpu_kdtree *T, const std::vector<Bound> &bnd, int max_depth, int max_leaf) {
    const int n = T->n;
    constexpr int kBins = 32;
    const float Ci = 1.2f, Ct = 1.0f;   // clheader.h:14-15
    std::vector<int> idx(n);
    std::vector<Vec> cen(n);
    for (int i = 0; i < n; i++) {
        idx[i] = i;
        cen[i] = {(bnd[i].min_x + bnd[i].max_x) * 0.5f, (bnd[i].min_y + bnd[i].max_y) * 0.5f,
                  (bnd[i].min_z + bnd[i].max_z) * 0.5f};
    }
    auto area = [](const Bound &b) {
        float dx = b.max_x - b.min_x, dy = b.max_y - b.min_y, dz = b.max_z - b.min_z;
        return 2.f * (dx * dy + dx * dz + dy * dz);
    };
    auto empty_box = []() {
        Bound b;
        b.min_x = b.min_y = b.min_z = INFINITY;
        b.max_x = b.max_y = b.max_z = -INFINITY;
        return b;
    };
    std::vector<Work> st;
    st.push_back({-1, 0, 0, n, 0});
    while (!st.empty()) {
        Work w = st.back();
        st.pop_back();
        const int me = (int)T->nodes.size();
        T->nodes.emplace_back();
        if (w.parent >= 0) (w.side ? T->nodes[w.parent].right : T->nodes[w.parent].left) = me;
        KNode &nd = T->nodes[me];
        nd.depth = w.depth;
        if (w.depth > T->maxdepth) T->maxdepth = w.depth;
        const int cnt = w.hi - w.lo;
        if (cnt == 0) continue;
        Bound box = bnd[idx[w.lo]], cb;
        cb.min_x = cb.max_x = cen[idx[w.lo]].x;
        cb.min_y = cb.max_y = cen[idx[w.lo]].y;
        cb.min_z = cb.max_z = cen[idx[w.lo]].z;
        for (int i = w.lo + 1; i < w.hi; i++) {
            expand(box, bnd[idx[i]]);
            const Vec &c = cen[idx[i]];
            Bound pb = {c.x, c.x, c.y, c.y, c.z, c.z};
            expand(cb, pb);
        }
        nd.box = box;
        int best_axis = -1, best_bin = -1;
        float best_cost = Ct * (float)cnt;
        if (cnt > max_leaf && w.depth < max_depth) {
            const float pa = area(box);
            for (int axis = 0; axis < 3; axis++) {
                const float lo = axis == 0 ? cb.min_x : (axis == 1 ? cb.min_y : cb.min_z);
                const float hi = axis == 0 ? cb.max_x : (axis == 1 ? cb.max_y : cb.max_z);
                if (!(hi > lo)) continue;
                Bound bb[kBins];
                int bc[kBins] = {0};
                for (auto &b : bb) b = empty_box();
                const float scale = kBins / (hi - lo);
                for (int i = w.lo; i < w.hi; i++) {
                    int k = (int)((comp(cen[idx[i]], axis) - lo) * scale);
                    k = k < 0 ? 0 : (k >= kBins ? kBins - 1 : k);
                    bc[k]++;
                    expand(bb[k], bnd[idx[i]]);
                }
                float ra[kBins];
                int rc[kBins];
                Bound acc = empty_box();
                int ac = 0;
                for (int k = kBins - 1; k > 0; k--) {
                    expand(acc, bb[k]);
                    ac += bc[k];
                    ra[k] = ac ? area(acc) : 0.f;
                    rc[k] = ac;
                }
                acc = empty_box();
                ac = 0;
                for (int k = 0; k < kBins - 1; k++) {
                    expand(acc, bb[k]);
                    ac += bc[k];
                    if (ac == 0 || rc[k + 1] == 0) continue;
                    const float cost = Ci + Ct * (area(acc) * (float)ac + ra[k + 1] * (float)rc[k + 1]) / pa;
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = axis;
                        best_bin = k;
                    }
                }
            }
        }
        if (best_axis < 0) {
            if (cnt > max_leaf && w.depth < max_depth) {
                // no useful SAH split (e.g. equal centroids): median split keeps leaves bounded
                const int axis = longest_axis(cb);
                const int midk = w.lo + cnt / 2;
                std::nth_element(idx.begin() + w.lo, idx.begin() + midk, idx.begin() + w.hi,
                                 [&](int a, int b) { return comp(cen[a], axis) < comp(cen[b], axis); });
                st.push_back({me, 1, midk, w.hi, w.depth + 1});
                st.push_back({me, 0, w.lo, midk, w.depth + 1});
                continue;
            }
            nd.leaf = 1;
            nd.start = (int)T->kn.size();
            nd.count = cnt;
            for (int i = w.lo; i < w.hi; i++) T->kn.push_back(idx[i]);
            T->szbuf += cnt;
            continue;
        }
        const float lo = best_axis == 0 ? cb.min_x : (best_axis == 1 ? cb.min_y : cb.min_z);
        const float hi = best_axis == 0 ? cb.max_x : (best_axis == 1 ? cb.max_y : cb.max_z);
        const float scale = kBins / (hi - lo);
        auto mid_it = std::stable_partition(idx.begin() + w.lo, idx.begin() + w.hi, [&](int s) {
            int k = (int)((comp(cen[s], best_axis) - lo) * scale);
            k = k < 0 ? 0 : (k >= kBins ? kBins - 1 : k);
            return k <= best_bin;
        });
        const int split = (int)(mid_it - idx.begin());
        st.push_back({me, 1, split, w.hi, w.depth + 1});
        st.push_back({me, 0, w.lo, split, w.depth + 1});
    }
}

Prim host_prim(const Shape &s, int id, bool last) {
    Prim p;
    int type = (int)s.type | (last ? kLastInLeaf : 0);
    float fid, ftype;
    memcpy(&fid, &id, 4);
    memcpy(&ftype, &type, 4);
    if (s.type == SPHERE) {
        p.p0 = make_float4(s.s.p.x, s.s.p.y, s.s.p.z, fid);
        p.p1 = make_float4(s.s.rad, 0.f, 0.f, ftype);
        p.p2 = make_float4(0.f, 0.f, 0.f, ftype);
    } else {
        p.p0 = make_float4(s.t.p1.x, s.t.p1.y, s.t.p1.z, fid);
        p.p1 = make_float4(s.t.p2.x - s.t.p1.x, s.t.p2.y - s.t.p1.y, s.t.p2.z - s.t.p1.z, ftype);
        p.p2 = make_float4(s.t.p3.x - s.t.p1.x, s.t.p3.y - s.t.p1.y, s.t.p3.z - s.t.p1.z, ftype);
    }
    return p;
}

// depth-first nodes -> TNode (internal nodes only, renumbered) + Prim (kn order), then upload
int upload_device(ptgpu_kdtree *T) {
    const int nn = (int)T->nodes.size();
    std::vector<int> inum(nn, -1);
    int ni = 0;
    for (int k = 0; k < nn; k++)
        if (!T->nodes[k].leaf && T->nodes[k].left >= 0) inum[k] = ni++;
    std::vector<TNode> tn((size_t)(ni > 0 ? ni : 1));
    auto ref_of = [&](int c, int *cnt) {
        const KNode &ch = T->nodes[c];
        if (ch.leaf) { *cnt = ch.count; return ~ch.start; }
        if (ch.left < 0) { *cnt = 0; return ~0; }   // empty node (only for an empty input)
        *cnt = 0;
        return inum[c];
    };
    for (int k = 0; k < nn; k++) {
        if (inum[k] < 0) continue;
        const KNode &nd = T->nodes[k];
        const Bound &L = T->nodes[nd.left].box, &R = T->nodes[nd.right].box;
        int cl, cr;
        const int rl = ref_of(nd.left, &cl), rr = ref_of(nd.right, &cr);
        TNode t;
        t.a = make_float4(L.min_x, L.min_y, L.min_z, L.max_x);
        t.b = make_float4(L.max_y, L.max_z, R.min_x, R.min_y);
        t.c = make_float4(R.min_z, R.max_x, R.max_y, R.max_z);
        t.r = make_int4(rl, rr, cl, cr);
        tn[inum[k]] = t;
    }
    std::vector<Prim> pr(T->kn.size() > 0 ? T->kn.size() : 1);
    // leaves own disjoint runs [start, start + count) of kn: flag the last primitive of each run
    std::vector<char> last(T->kn.size() + 1, 0);
    for (const KNode &nd : T->nodes)
        if (nd.leaf && nd.count > 0) last[(size_t)nd.start + nd.count - 1] = 1;
    for (size_t q = 0; q < T->kn.size(); q++) pr[q] = host_prim(T->shapes[T->kn[q]], T->kn[q], last[q] != 0);
    PT_TRY(T->d_nodes.ensure(tn.size() * sizeof(TNode)));
    PT_TRY(T->d_prims.ensure(pr.size() * sizeof(Prim)));
    hipStream_t st = T->ctx->stream;
    PT_HIP(hipMemcpyAsync(T->d_nodes.p, tn.data(), tn.size() * sizeof(TNode), hipMemcpyHostToDevice, st));
    PT_HIP(hipMemcpyAsync(T->d_prims.p, pr.data(), pr.size() * sizeof(Prim), hipMemcpyHostToDevice, st));
    PT_HIP(hipStreamSynchronize(st));
    DevAccel &A = T->acc;
    A.nodes = T->d_nodes.as<TNode>();
    A.prims = T->d_prims.as<Prim>();
    A.nprims = (int)T->kn.size();
    A.nnodes = ni;
    A.nshapes = T->n;
    A.leaf_runs = 1;
    A.mode = PTGPU_ACCEL_KD;
    const KNode &root = T->nodes[0];
    if (root.leaf) {
        A.root = ~root.start;
        A.root_cnt = root.count;
    } else {
        A.root = 0;
        A.root_cnt = 0;
    }
    A.depth = T->maxdepth;
    A.nodes4 = nullptr;
    A.nodes4h = nullptr;
    PT_TRY(build_wide(T->ctx, &A, T->d_nodes4));
    // the tree is complete when the build returns: a render on any context's stream may use it
    PT_HIP(hipStreamSynchronize(st));
    return PTGPU_OK;
}

int kd_common(ptgpu_ctx *ctx, const Shape *h_shapes, int n, bool sah, int max_depth, int max_leaf,
              ptgpu_kdtree **out) {
    PT_CHECK(h_shapes && out, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(n >= 1, PTGPU_ERR_INVALID, "shape count %d < 1", n);
    *out = nullptr;
    if (ctx) PT_HIP(hipSetDevice(ctx->device));
    ptgpu_kdtree *T = new ptgpu_kdtree();
    T->ctx = ctx;
    T->n = n;
    T->sah = sah;
    T->shapes.assign(h_shapes, h_shapes + n);
    std::vector<Bound> bnd(n);
    std::vector<Vec> mid(n);
    for (int i = 0; i < n; i++) {
        bnd[i] = shape_bound_host(h_shapes[i]);
        mid[i] = midpoint_host(h_shapes[i]);
    }
    if (sah) build_sah(T, bnd, max_depth, max_leaf);
    else build_mean_split(T, bnd, mid);
    int s = ctx ? upload_device(T) : PTGPU_OK;   // ctx == NULL: host tree only (no device layout)
    if (s != PTGPU_OK) {
        ptgpu_kdtree_destroy(T);
        return s;
    }
    *out = T;
    return PTGPU_OK;
}

}  // namespace

namespace ptgpu {
int kdtree_lights(ptgpu_kdtree *kd, const Shape *d_shapes, int n, int light_cnt, const int **lights, int *nlights) {
    PT_CHECK(kd->ctx != nullptr, PTGPU_ERR_INVALID, "KD tree was built without a context (host only)");
    if (kd->lights_for != light_cnt) {
        PT_TRY(build_light_list(kd->ctx, d_shapes, n, light_cnt, kd->lights, kd->scan, &kd->nlights));
        kd->lights_for = light_cnt;
    }
    *lights = kd->lights.as<int>();
    *nlights = kd->nlights;
    return PTGPU_OK;
}
}  // namespace ptgpu

extern "C" {

int ptgpu_kdtree_build(ptgpu_ctx *ctx, const Shape *h_shapes, int n, ptgpu_kdtree **out) {
    return kd_common(ctx, h_shapes, n, false, 0, 0, out);
}

int ptgpu_kdtree_build_sah(ptgpu_ctx *ctx, const Shape *h_shapes, int n, int max_depth, int max_leaf,
                           ptgpu_kdtree **out) {
    PT_CHECK(max_depth >= 1 && max_depth <= 60, PTGPU_ERR_INVALID, "max_depth %d out of [1,60]", max_depth);
    PT_CHECK(max_leaf >= 1, PTGPU_ERR_INVALID, "max_leaf %d < 1", max_leaf);
    return kd_common(ctx, h_shapes, n, true, max_depth, max_leaf, out);
}

int ptgpu_kdtree_stats(ptgpu_kdtree *kd, int *max_depth, int *nodes, int *leaves, int *kn_len) {
    PT_CHECK(kd, PTGPU_ERR_INVALID, "null handle");
    int nl = 0;
    for (auto &nd : kd->nodes) nl += nd.leaf;
    if (max_depth) *max_depth = kd->maxdepth;
    if (nodes) *nodes = (int)kd->nodes.size();
    if (leaves) *leaves = nl;
    if (kn_len) *kn_len = (int)kd->kn.size();
    return PTGPU_OK;
}

// Depth-first export: node5[k] = {leaf, left, right, start, count}, box[k]; kn in leaf order.
int ptgpu_kdtree_export_dfs(ptgpu_kdtree *kd, int *node5, Bound *box, int *kn, int cap_nodes, int cap_kn) {
    PT_CHECK(kd && node5 && box && kn, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(cap_nodes >= (int)kd->nodes.size() && cap_kn >= (int)kd->kn.size(), PTGPU_ERR_INVALID,
             "buffers too small (%d nodes, %d kn needed)", (int)kd->nodes.size(), (int)kd->kn.size());
    for (size_t k = 0; k < kd->nodes.size(); k++) {
        const KNode &nd = kd->nodes[k];
        node5[k * 5 + 0] = nd.leaf;
        node5[k * 5 + 1] = nd.left;
        node5[k * 5 + 2] = nd.right;
        node5[k * 5 + 3] = nd.leaf ? nd.start : 0;
        node5[k * 5 + 4] = nd.leaf ? nd.count : 0;
        box[k] = nd.box;
    }
    memcpy(kn, kd->kn.data(), kd->kn.size() * sizeof(int));
    return PTGPU_OK;
}

// KDTree::getTrees / traverseTree / fillNode, KDTree.cpp:184-270: BFS into a 1-based heap of
// 2^(maxdepth+1) nodes with children at 2i, 2i+1; leaves list their shapes in kn in BFS order.
int ptgpu_kdtree_get_trees(ptgpu_kdtree *kd, KDNodeGPU **h_kng, int *kngCnt, int **h_kn, int *knCnt) {
    PT_CHECK(kd && h_kng && kngCnt && h_kn && knCnt, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(!kd->sah, PTGPU_ERR_UNSUPPORTED, "the KDNodeGPU heap export is defined for the reference builder");
    PT_CHECK(kd->maxdepth <= 13 && kd->szbuf <= 32767, PTGPU_ERR_UNSUPPORTED,
             "tree (max depth %d, %lld primitive refs) does not fit the int16 fields of KDNodeGPU", kd->maxdepth,
             kd->szbuf);
    const int size = 1 << (kd->maxdepth + 1);
    KDNodeGPU *heap = (KDNodeGPU *)calloc((size_t)size, sizeof(KDNodeGPU));
    int *kn = (int *)malloc(sizeof(int) * (size_t)(kd->szbuf > 0 ? kd->szbuf : 1));
    PT_CHECK(heap && kn, PTGPU_ERR_OOM, "host allocation failed");
    // BFS over the depth-first node array; -1 marks an absent child slot
    std::vector<int> q, nq;
    q.push_back(0);
    int loc = 1, shp = 0;
    while (!q.empty() && loc < size) {
        nq.clear();
        for (int k : q) {
            if (loc >= size) break;
            if (k >= 0) {
                const KNode &nd = kd->nodes[k];
                KDNodeGPU &h = heap[loc];
                h.bound = nd.box;
                h.leaf = (short)nd.leaf;
                if (nd.leaf) {
                    h.min = (short)shp;
                    for (int i = 0; i < nd.count; i++) kn[shp++] = kd->kn[nd.start + i];
                    h.max = (short)shp;
                }
                h.nLeft = (short)(2 * loc);
                h.nRight = (short)(2 * loc + 1);
                h.nParent = (short)(loc / 2);
                nq.push_back(nd.left);
                nq.push_back(nd.right);
            } else {
                nq.push_back(-1);
                nq.push_back(-1);
            }
            loc++;
        }
        q.swap(nq);
    }
    *h_kng = heap;
    *kngCnt = size;
    *h_kn = kn;
    *knCnt = (int)kd->szbuf;
    return PTGPU_OK;
}

int ptgpu_kdtree_device_accel(ptgpu_kdtree *kd, ptgpu::DevAccel *acc) {
    PT_CHECK(kd && acc, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(kd->ctx != nullptr, PTGPU_ERR_INVALID, "KD tree was built without a context (host only)");
    *acc = kd->acc;
    return PTGPU_OK;
}

int ptgpu_kdtree_destroy(ptgpu_kdtree *kd) {
    if (!kd) return PTGPU_OK;
    if (kd->ctx && kd->ctx->stream) (void)hipStreamSynchronize(kd->ctx->stream);
    kd->d_nodes.release();
    kd->d_nodes4.release();
    kd->lights.release();
    kd->scan.release();
    kd->d_prims.release();
    delete kd;
    return PTGPU_OK;
}

}  // extern "C"
