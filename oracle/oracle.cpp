/*
 * oracle.cpp -- CPU restatement of the PTGPU_PC ray-acceleration and radiance path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by tests/, by the smoke check
 * and by the CPU-baseline leg of bench.py, never by the product library (ptgpu_pc_amd/), which
 * must fail loudly rather than fall back to anything here.
 *
 * Every function cites the reference file:line it restates (paths relative to the reference
 * repository root).  Two floating-point semantics are available where the reference has two:
 *   sem = 0  "kernel": the OpenCL kernel (rendering_kernel.cl), FP32 throughout, OpenCL builtins
 *            sign/clamp/max;
 *   sem = 1  "host twin": include/geomfunc.h, the host copy used by the reference's CPU path
 *            (FP64 det/inv_det in TriangleIntersect, sign/clamp/max as vec.h macros).
 * Built with -ffp-contract=off and IEEE division/sqrt, so every expression rounds as written.
 *
 * Pinning (what checks each function; the tests say which):
 *   or_get_random            known-answer vectors + independent numpy MWC + oracle/_ref
 *   or_clbvh_construct       hand-computed Morton cases (FLT_MIN max, 1024 wrap, 0/0) + numpy restatement
 *   or_bubblesort            equals std::stable_sort(greater) and numpy lexsort on small inputs
 *   or_radix_tree            equals the closed-form top-down split (or_topology_closed_form) for every
 *                            N <= 3000 and sampled large N; Karras invariants
 *   or_refit                 brute-force union over each node's leaf range
 *   or_kd_*                  hand-built small cases (N=1, equal midpoints, two-shape split, heap slots)
 *   intersections            closed forms (axis rays, det==0, t==EPSILON) + oracle/_ref bit-exact (sem=1)
 *   or_intersect (modes)     BVH/KD traversal equals brute force on random rays
 *   radiance                 oracle/_ref RadiancePathTracing bit-exact (sem=1), analytic furnace/miss cases
 *   or_sincos (sem=0)        within one ulp plus the reduction error of the binary64 libm value, and
 *                            exactly rounded on most arguments (the mathematics, not a retyped formula)
 *   per-pixel wrapper        UpdateCamera closed forms (yaw/pitch 0 and 90 degrees), GenerateCameraRay
 *                            by hand, toInt known values, row flip and running mean on an asymmetric scene
 */
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <queue>
#include <vector>

#include "../include/ptgpu.h"

#define OR_EPSILON 0.01f                       /* geom.h:29 */
#define OR_PI 3.14159265358979323846f          /* geom.h:30 */

extern "C" {

/* ----------------------------------------------------------------------------------------- */
/* RNG -- simplernd.h:35-49, rendering_kernel.cl:4-17                                         */
/* ----------------------------------------------------------------------------------------- */
float or_get_random(unsigned int *seed0, unsigned int *seed1) {
    *seed0 = 36969u * ((*seed0) & 65535u) + ((*seed0) >> 16);
    *seed1 = 18000u * ((*seed1) & 65535u) + ((*seed1) >> 16);
    unsigned int ires = ((*seed0) << 16) + (*seed1);
    unsigned int bits = (ires & 0x007fffffu) | 0x40000000u;
    float f;
    memcpy(&f, &bits, 4);
    return (f - 2.f) / 2.f;
}

/* ----------------------------------------------------------------------------------------- */
/* LBVH host part -- CLBVH.cpp                                                                */
/* ----------------------------------------------------------------------------------------- */
static inline float or_min2(float a, float b) { return (a < b) ? a : b; }   /* AccelCommon.h:24 */
static inline float or_max2(float a, float b) { return (a > b) ? a : b; }   /* AccelCommon.h:29 */
static inline float or_min3(float a, float b, float c) { return or_min2(or_min2(a, b), c); }
static inline float or_max3(float a, float b, float c) { return or_max2(or_max2(a, b), c); }

/* CLBVH::getBound(Sphere) CLBVH.cpp:49-61 and getBound(Triangle) CLBVH.cpp:63-75 */
static Bound or_shape_bound(const Shape *s) {
    Bound b;
    if (s->type == SPHERE) {
        b.min_x = s->s.p.x - s->s.rad; b.max_x = s->s.p.x + s->s.rad;
        b.min_y = s->s.p.y - s->s.rad; b.max_y = s->s.p.y + s->s.rad;
        b.min_z = s->s.p.z - s->s.rad; b.max_z = s->s.p.z + s->s.rad;
    } else {
        b.min_x = or_min3(s->t.p1.x, s->t.p2.x, s->t.p3.x);
        b.max_x = or_max3(s->t.p1.x, s->t.p2.x, s->t.p3.x);
        b.min_y = or_min3(s->t.p1.y, s->t.p2.y, s->t.p3.y);
        b.max_y = or_max3(s->t.p1.y, s->t.p2.y, s->t.p3.y);
        b.min_z = or_min3(s->t.p1.z, s->t.p2.z, s->t.p3.z);
        b.max_z = or_max3(s->t.p1.z, s->t.p2.z, s->t.p3.z);
    }
    return b;
}

/* float -> unsigned as the reference's cast (CLBVH.cpp:153-155); the value is always in
 * [0, 1024] or NaN (0/0 on a degenerate axis), which is defined here as 0. */
static inline unsigned int or_f2u(float v) {
    if (!(v >= 0.f)) return 0u;
    return (unsigned int)v;
}

/* Morton codes: CLBVH::CLBVH, CLBVH.cpp:90-164.  Writes index, b, morton_code into shapes and
 * returns the scene centroid bounds in scene6 = {min_x, max_x, min_y, max_y, min_z, max_z}. */
void or_clbvh_construct(Shape *shapes, int n, float *scene6) {
    float min_x = FLT_MAX, max_x = FLT_MIN;   /* FLT_MIN, not -FLT_MAX: CLBVH.cpp:92-96 */
    float min_y = FLT_MAX, max_y = FLT_MIN;
    float min_z = FLT_MAX, max_z = FLT_MIN;
    for (int i = 0; i < n; i++) {
        Vec p = {0.f, 0.f, 0.f};
        shapes[i].index = i;
        shapes[i].b = or_shape_bound(&shapes[i]);
        if (shapes[i].type == SPHERE) {
            p = shapes[i].s.p;
        } else {
            p.x = (shapes[i].b.min_x + shapes[i].b.max_x) / 2.0f;
            p.y = (shapes[i].b.min_y + shapes[i].b.max_y) / 2.0f;
            p.z = (shapes[i].b.min_z + shapes[i].b.max_z) / 2.0f;
        }
        if (p.x < min_x) min_x = p.x;
        if (p.x > max_x) max_x = p.x;
        if (p.y < min_y) min_y = p.y;
        if (p.y > max_y) max_y = p.y;
        if (p.z < min_z) min_z = p.z;
        if (p.z > max_z) max_z = p.z;
    }
    for (int i = 0; i < n; i++) {
        shapes[i].index = i;
        shapes[i].morton_code = 0;
        Vec p = {0.f, 0.f, 0.f};
        if (shapes[i].type == SPHERE) {
            p = shapes[i].s.p;
        } else {
            const Triangle &t = shapes[i].t;
            float lx = or_min3(t.p1.x, t.p2.x, t.p3.x), ly = or_min3(t.p1.y, t.p2.y, t.p3.y);
            float lz = or_min3(t.p1.z, t.p2.z, t.p3.z);
            float hx = or_max3(t.p1.x, t.p2.x, t.p3.x), hy = or_max3(t.p1.y, t.p2.y, t.p3.y);
            float hz = or_max3(t.p1.z, t.p2.z, t.p3.z);
            p.x = (lx + hx) / 2.0f;
            p.y = (ly + hy) / 2.0f;
            p.z = (lz + hz) / 2.0f;
        }
        unsigned int a = or_f2u(((p.x - min_x) / (max_x - min_x)) * 1024.f);
        unsigned int b = or_f2u(((p.y - min_y) / (max_y - min_y)) * 1024.f);
        unsigned int c = or_f2u(((p.z - min_z) / (max_z - min_z)) * 1024.f);
        unsigned int code = 0;
        for (unsigned int j = 0; j < 10; j++) {   /* CLBVH.cpp:158-163 */
            code |= (((a >> (9 - j)) & 1u) << ((10 - j) * 3 - 1)) |
                    (((b >> (9 - j)) & 1u) << ((10 - j) * 3 - 2)) |
                    (((c >> (9 - j)) & 1u) << ((10 - j) * 3 - 3));
        }
        shapes[i].morton_code = code;
    }
    if (scene6) {
        scene6[0] = min_x; scene6[1] = max_x; scene6[2] = min_y;
        scene6[3] = max_y; scene6[4] = min_z; scene6[5] = max_z;
    }
}

/* CLBVH::bubblesort_Geometry, CLBVH.cpp:31-47, with the key setup of buildBVHTree :249-255.
 * O(N^2): the literal algorithm, for small N. */
void or_bubblesort(const unsigned int *keys, int n, int *sorted) {
    struct G { unsigned int key; int value; };
    std::vector<G> geo((size_t)n);
    for (int i = 0; i < n; i++) { geo[i].key = keys[i]; geo[i].value = i; }
    for (unsigned int level = 0; level + 1 < (unsigned int)n; level++)
        for (unsigned int i = 0; i < (unsigned int)n - level - 1; i++)
            if (geo[i].key < geo[i + 1].key) std::swap(geo[i], geo[i + 1]);
    for (int i = 0; i < n; i++) sorted[i] = geo[i].value;
}

/* The same order for any N: stable sort by descending key (the bubble sort swaps only on a
 * strict <, so equal keys keep their input order). */
void or_stable_sort_desc(const unsigned int *keys, int n, int *sorted) {
    std::vector<int> idx((size_t)n);
    for (int i = 0; i < n; i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return keys[a] > keys[b]; });
    for (int i = 0; i < n; i++) sorted[i] = idx[i];
}

/* fastclz BVH.cl:98-110 (count of leading zeros of a 32-bit value, 32 for 0). */
static int or_fastclz(int iv) {
    unsigned int v = (unsigned int)iv;
    int x = (0 != (v >> 16)) * 16;
    x += (0 != (v >> (x + 8))) * 8;
    x += (0 != (v >> (x + 4))) * 4;
    x += (0 != (v >> (x + 2))) * 2;
    x += (0 != (v >> (x + 1)));
    x += (0 != (v >> x));
    return 32 - x;
}
/* longestCommonPrefix BVH.cl:127-134 -- over leaf INDICES, never Morton codes. */
static int or_lcp(int i, int j, int len) {
    if (0 <= j && j < len) return or_fastclz(i ^ j);
    return -1;
}

/* CLBVH::CLBVH buffer init (CLBVH.cpp:83-88) + kernelConstructRadixTree (BVH.cl:141-208) run
 * for every internal node i in [0, n-2].  The reference's first binary search re-probes its
 * first step and its second one keeps probing t=1 until its int divider overflows; both end at
 * the largest step count whose probe still succeeds, which is what the bounded loops below
 * compute (the t=1 probing is continued until it fails). */
void or_radix_tree(int n, BVHNodeGPU *btn, BVHNodeGPU *btl) {
    if (n > 1) memset(btn, 0, sizeof(BVHNodeGPU) * (size_t)(n - 1));
    memset(btl, 0xFF, sizeof(BVHNodeGPU) * (size_t)n);
    const int len = n - 1;
    for (int i = 0; i < len; i++) {
        int d = (or_lcp(i, i + 1, len + 1) - or_lcp(i, i - 1, len + 1)) > 0 ? 1 : -1;
        int sigMin = or_lcp(i, i - d, len + 1);
        long long lmax = 2;
        while (or_lcp(i, (int)(i + lmax * d), len + 1) > sigMin) lmax *= 2;
        long long l = 0;
        {   /* BVH.cl:161-168: probes lmax/2 twice, then lmax/4, ..., 1 */
            long long divider = 2;
            for (long long t = lmax / divider; t >= 1; divider *= 2) {
                if (or_lcp(i, (int)(i + (l + t) * d), len + 1) > sigMin) l += t;
                t = lmax / divider;
            }
        }
        int j = (int)(i + l * d);
        int sigNode = or_lcp(i, j, len + 1);
        long long s = 0;
        {   /* BVH.cl:174-181: ceil(l/2) twice, ceil(l/4), ..., then t=1 repeatedly */
            long long divider = 2;
            long long t = (l + (divider - 1)) / divider;
            while (t >= 1) {
                if (or_lcp(i, (int)(i + (s + t) * d), len + 1) > sigNode) s += t;
                else if (t == 1 && divider > 2 * l + 2) break;  /* further t=1 probes all fail */
                t = (l + (divider - 1)) / divider;
                divider *= 2;
                if (t < 1) t = 1;
            }
        }
        int gamma = (int)(i + s * d + std::min(d, 0));
        BVHNodeGPU *cur = btn + i;
        int lo = std::min(i, j), hi = std::max(i, j);
        if (lo == gamma) { cur->nLeft = gamma; cur->leaf = 2; btl[gamma].nParent = i; }
        else             { cur->nLeft = gamma; btn[gamma].nParent = i; }
        if (hi == gamma + 1) { cur->nRight = gamma + 1; cur->leaf = 2; btl[gamma + 1].nParent = i; }
        else                 { cur->nRight = gamma + 1; btn[gamma + 1].nParent = i; }
        cur->min = lo;
        cur->max = hi;
    }
}

/* Independent closed form of the same topology: the binary radix tree of the keys 0..n-1.  The
 * node covering [a, b] splits after gamma = ((b >> h) << h) - 1 with h = msb(a ^ b); its left
 * child is internal node gamma (or leaf gamma when a == gamma), its right child internal node
 * gamma+1 (or leaf gamma+1 when gamma+1 == b).  Arrays are indexed by internal node. */
void or_topology_closed_form(int n, int *left, int *right, int *lo, int *hi, int *parent_int,
                             int *parent_leaf) {
    if (n < 2) return;
    struct R { int idx, a, b; };
    std::vector<R> st;
    st.push_back({0, 0, n - 1});
    parent_int[0] = 0;
    while (!st.empty()) {
        R r = st.back(); st.pop_back();
        unsigned int x = (unsigned int)(r.a ^ r.b);
        int h = 31 - __builtin_clz(x);
        int gamma = ((r.b >> h) << h) - 1;
        left[r.idx] = gamma; right[r.idx] = gamma + 1; lo[r.idx] = r.a; hi[r.idx] = r.b;
        if (r.a == gamma) parent_leaf[gamma] = r.idx;
        else { parent_int[gamma] = r.idx; st.push_back({gamma, r.a, gamma}); }
        if (gamma + 1 == r.b) parent_leaf[gamma + 1] = r.idx;
        else { parent_int[gamma + 1] = r.idx; st.push_back({gamma + 1, gamma + 1, r.b}); }
    }
}

/* merge_bounds BVH.cl:112-122 */
static Bound or_merge(const Bound &b1, const Bound &b2) {
    Bound b3;
    b3.min_x = or_min2(b1.min_x, b2.min_x); b3.max_x = or_max2(b1.max_x, b2.max_x);
    b3.min_y = or_min2(b1.min_y, b2.min_y); b3.max_y = or_max2(b1.max_y, b2.max_y);
    b3.min_z = or_min2(b1.min_z, b2.min_z); b3.max_z = or_max2(b1.max_z, b2.max_z);
    return b3;
}

/* kernelConstructBVHTree BVH.cl:214-258 with the intended semantics: leaf k gets the AABB and
 * index of shape sorted[k]; every internal node gets merge(left child, right child), where a
 * child is a leaf iff its index equals the node's range end on that side (nLeft == min,
 * nRight == max).  The reference's grandparent counter (BVH.cl:253-256) is not reproduced. */
void or_refit(const Shape *shapes, const int *sorted, int n, BVHNodeGPU *btn, BVHNodeGPU *btl) {
    for (int k = 0; k < n; k++) {
        btl[k].bound = shapes[sorted[k]].b;
        btl[k].nShape = sorted[k];
    }
    if (n < 2) return;
    /* iterative post-order from the root */
    std::vector<std::pair<int, int>> st;   /* (node, state) */
    st.push_back({0, 0});
    while (!st.empty()) {
        auto &top = st.back();
        int i = top.first;
        bool lleaf = btn[i].nLeft == btn[i].min, rleaf = btn[i].nRight == btn[i].max;
        if (top.second == 0) {
            top.second = 1;
            if (!rleaf) st.push_back({btn[i].nRight, 0});
            if (!lleaf) st.push_back({btn[i].nLeft, 0});
            continue;
        }
        const Bound &L = lleaf ? btl[btn[i].nLeft].bound : btn[btn[i].nLeft].bound;
        const Bound &R = rleaf ? btl[btn[i].nRight].bound : btn[btn[i].nRight].bound;
        btn[i].bound = or_merge(L, R);
        st.pop_back();
    }
}

/* ----------------------------------------------------------------------------------------- */
/* KD tree -- KDTree.cpp:8-270, KDTree.h:35-44                                                 */
/* ----------------------------------------------------------------------------------------- */
struct OrKDNode {
    bool leaf = false;
    OrKDNode *left = nullptr, *right = nullptr;
    Bound box{};
    std::vector<const Shape *> shapes;
};
struct OrKDTree {
    int maxdepth = 0;
    long long szbuf = 0;
    OrKDNode *root = nullptr;
    std::vector<OrKDNode *> all;
};

/* KDTree::getMidpoint KDTree.cpp:46-70.  For a triangle the reference expression
 * (p1.x, p2.x, p3.x) / 3.0f is a comma expression and evaluates to p3.x / 3.0f. */
static Vec or_kd_midpoint(const Shape *s) {
    Vec v;
    if (s->type == TRIANGLE) {
        v.x = s->t.p3.x / 3.0f; v.y = s->t.p3.y / 3.0f; v.z = s->t.p3.z / 3.0f;
    } else {
        v = s->s.p;
    }
    return v;
}
/* KDTreeNode::expandBound KDTree.h:35-44 */
static void or_kd_expand(Bound &box, const Bound &b) {
    if (b.min_x < box.min_x) box.min_x = b.min_x;
    if (b.min_y < box.min_y) box.min_y = b.min_y;
    if (b.min_z < box.min_z) box.min_z = b.min_z;
    if (b.max_x > box.max_x) box.max_x = b.max_x;
    if (b.max_y > box.max_y) box.max_y = b.max_y;
    if (b.max_z > box.max_z) box.max_z = b.max_z;
}
/* KDTree::getLongestAxis KDTree.cpp:72-86 */
static int or_kd_axis(const Bound &b) {
    float xd = b.max_x - b.min_x, yd = b.max_y - b.min_y, zd = b.max_z - b.min_z;
    if (xd >= yd) return (xd >= zd) ? 0 : 2;
    return (yd >= zd) ? 1 : 2;
}
static float or_axis_of(const Vec &v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }

/* KDTree::build KDTree.cpp:89-182.  The reference recurses; an explicit work stack gives the same
 * nodes, depths and leaf contents without native recursion on deep trees. */
static OrKDNode *or_kd_build_iter(OrKDTree *T, std::vector<const Shape *> s0) {
    struct Work { OrKDNode **slot; std::vector<const Shape *> s; int depth; };
    OrKDNode *root = nullptr;
    std::vector<Work> st;
    st.push_back({&root, std::move(s0), 0});
    while (!st.empty()) {
        Work w = std::move(st.back());
        st.pop_back();
        std::vector<const Shape *> &s = w.s;
        OrKDNode *node = new OrKDNode();
        T->all.push_back(node);
        *w.slot = node;
        if (w.depth > T->maxdepth) T->maxdepth = w.depth;
        if (s.empty()) continue;
        if (s.size() <= 1) {
            node->shapes = s; node->leaf = true;
            node->box = or_shape_bound(s[0]);
            T->szbuf += (long long)s.size();
            continue;
        }
        Vec mid = {0.f, 0.f, 0.f};
        float recp = (float)(1.0 / (double)s.size());
        node->box = or_shape_bound(s[0]);
        for (size_t i = 1; i < s.size(); i++) {   /* skips s[0] for the mean, KDTree.cpp:127 */
            or_kd_expand(node->box, or_shape_bound(s[i]));
            Vec m = or_kd_midpoint(s[i]);
            m.x *= recp; m.y *= recp; m.z *= recp;
            mid.x += m.x; mid.y += m.y; mid.z += m.z;
        }
        std::vector<const Shape *> ls, rs;
        int axis = or_kd_axis(node->box);
        float mv = or_axis_of(mid, axis);
        for (size_t i = 0; i < s.size(); i++) {   /* midpoint <= mean goes right, KDTree.cpp:142-155 */
            if (mv >= or_axis_of(or_kd_midpoint(s[i]), axis)) rs.push_back(s[i]);
            else ls.push_back(s[i]);
        }
        if (s.size() == ls.size() || s.size() == rs.size()) {
            node->shapes = s; node->leaf = true;
            node->box = or_shape_bound(s[0]);
            for (size_t i = 1; i < s.size(); i++) or_kd_expand(node->box, or_shape_bound(s[i]));
            T->szbuf += (long long)(ls.size() + rs.size());
            continue;
        }
        st.push_back({&node->right, std::move(rs), w.depth + 1});
        st.push_back({&node->left, std::move(ls), w.depth + 1});
    }
    return root;
}

void *or_kd_build(const Shape *shapes, int n) {
    OrKDTree *T = new OrKDTree();
    std::vector<const Shape *> vs;
    for (int i = 0; i < n; i++) vs.push_back(&shapes[i]);
    T->root = or_kd_build_iter(T, std::move(vs));
    return T;
}
void or_kd_free(void *h) {
    OrKDTree *T = (OrKDTree *)h;
    for (auto *p : T->all) delete p;
    delete T;
}
/* maxdepth, szbuf (knCnt), node count, leaf count */
void or_kd_stats(void *h, int *maxdepth, long long *szbuf, int *nodes, int *leaves) {
    OrKDTree *T = (OrKDTree *)h;
    *maxdepth = T->maxdepth; *szbuf = T->szbuf;
    int nn = 0, nl = 0;
    for (auto *p : T->all) { nn++; if (p->leaf) nl++; }
    *nodes = nn; *leaves = nl;
}

/* Canonical depth-first (pre-order, left before right) flattening, int32 fields:
 * node[k] = {leaf, left, right, start, count}, box[k]; leaves list their shape indices in kn in
 * the order the builder stored them.  This is the comparison format for the library's build. */
int or_kd_flatten_dfs(void *h, int *node5, Bound *box, int *kn, int cap_nodes, long long cap_kn) {
    OrKDTree *T = (OrKDTree *)h;
    std::vector<std::pair<OrKDNode *, int>> st;
    int cnt = 0;
    long long kpos = 0;
    /* pre-order numbering: assign index when popped; children indices patched after */
    std::vector<int> fixups;
    st.push_back({T->root, -1});
    while (!st.empty()) {
        auto e = st.back(); st.pop_back();
        OrKDNode *nd = e.first;
        int me = cnt++;
        if (me >= cap_nodes) return -1;
        if (e.second >= 0) {
            /* e.second encodes parent*2 + side */
            int par = e.second >> 1, side = e.second & 1;
            node5[par * 5 + 1 + side] = me;
        }
        node5[me * 5 + 0] = nd->leaf ? 1 : 0;
        node5[me * 5 + 1] = -1;
        node5[me * 5 + 2] = -1;
        node5[me * 5 + 3] = 0;
        node5[me * 5 + 4] = 0;
        box[me] = nd->box;
        if (nd->leaf) {
            node5[me * 5 + 3] = (int)kpos;
            node5[me * 5 + 4] = (int)nd->shapes.size();
            for (auto *sp : nd->shapes) {
                if (kpos >= cap_kn) return -2;
                kn[kpos++] = sp->index;
            }
        } else if (nd->left || nd->right) {
            if (nd->right) st.push_back({nd->right, me * 2 + 1});
            if (nd->left) st.push_back({nd->left, me * 2 + 0});
        }
    }
    return cnt;
}

/* KDTree::getTrees/traverseTree/fillNode KDTree.cpp:184-270: BFS into a 1-based heap of
 * size 2^(maxdepth+1); int16 fields.  Unfilled slots and the min/max of internal nodes, which
 * the reference leaves uninitialised, are zero here.  Returns the heap size, or -1 when
 * heap_cap is too small. */
int or_kd_export_heap(void *h, KDNodeGPU *heap, int heap_cap, int *kn, long long kn_cap) {
    OrKDTree *T = (OrKDTree *)h;
    int size = (int)pow(2.0, (double)(T->maxdepth + 1));
    if (size > heap_cap) return -1;
    memset(heap, 0, sizeof(KDNodeGPU) * (size_t)size);
    int curnodeloc = 1, curshapeloc = 0;
    std::queue<OrKDNode *> q;
    q.push(T->root);
    while (!q.empty()) {
        OrKDNode *node = q.front(); q.pop();
        if (node) {
            KDNodeGPU &k = heap[curnodeloc];
            k.bound = node->box;
            k.leaf = (short)(node->leaf ? 1 : 0);
            if (node->leaf) {
                k.min = (short)curshapeloc;
                for (auto *sp : node->shapes) {
                    if (curshapeloc >= kn_cap) return -2;
                    kn[curshapeloc++] = sp->index;
                }
                k.max = (short)curshapeloc;
            }
            k.nLeft = (short)(2 * curnodeloc);
            k.nRight = (short)(2 * curnodeloc + 1);
            k.nParent = (short)(curnodeloc / 2);
            curnodeloc++;
            q.push(node->left);
            q.push(node->right);
        } else {
            q.push(nullptr);
            q.push(nullptr);
            curnodeloc++;
        }
        if (curnodeloc >= size) break;
    }
    return size;
}

/* ----------------------------------------------------------------------------------------- */
/* Rendering -- rendering_kernel.cl (sem 0) / geomfunc.h (sem 1)                              */
/* ----------------------------------------------------------------------------------------- */
static inline Vec v3(float x, float y, float z) { Vec v; v.x = x; v.y = y; v.z = z; return v; }
static inline Vec vsub(const Vec &a, const Vec &b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline Vec vadd(const Vec &a, const Vec &b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline Vec vmul(const Vec &a, const Vec &b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline Vec vsmul(float k, const Vec &b) { return v3(k * b.x, k * b.y, k * b.z); }
static inline float vdot(const Vec &a, const Vec &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline Vec vxcross(const Vec &a, const Vec &b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline Vec vnorm(const Vec &v) { float l = 1.f / sqrtf(vdot(v, v)); return vsmul(l, v); }
static inline bool viszero(const Vec &v) { return v.x == 0.f && v.y == 0.f && v.z == 0.f; }

/* SphereIntersect rendering_kernel.cl:26-56 (= geomfunc.h:50-75) */
float or_sphere_intersect(const Sphere *s, const Ray *r) {
    Vec op = vsub(s->p, r->o);
    float b = vdot(op, r->d);
    float det = b * b - vdot(op, op) + s->rad * s->rad;
    if (det < 0.f) return 0.f;
    det = sqrtf(det);
    float t = b - det;
    if (t > OR_EPSILON) return t;
    t = b + det;
    if (t > OR_EPSILON) return t;
    return 0.f;
}

/* TriangleIntersect rendering_kernel.cl:59-125 (sem 0, FP32; 1.0/det in double rounds to the
 * same float as 1.f/det) and geomfunc.h:79-130 (sem 1, double det/inv_det). */
float or_triangle_intersect(const Triangle *tr, const Ray *r, int sem) {
    Vec e1 = vsub(tr->p2, tr->p1), e2 = vsub(tr->p3, tr->p1);
    Vec pvec = vxcross(r->d, e2);
    if (sem == 0) {
        float det = vdot(e1, pvec);
        if (det == 0) return 0.0f;
        float inv_det = 1.0f / det;
        Vec tvec = vsub(r->o, tr->p1);
        float u = vdot(tvec, pvec) * inv_det;
        if (u < 0.0f || u > 1.0f) return 0.0f;
        Vec qvec = vxcross(tvec, e1);
        float v = vdot(r->d, qvec) * inv_det;
        if (v < 0.0f || u + v > 1.0f) return 0.0f;
        float t = vdot(e2, qvec) * inv_det;
        if (t > OR_EPSILON) return t;
        return 0.0f;
    } else {
        double det = vdot(e1, pvec);
        if (det == 0) return 0.0f;
        double inv_det = 1.0 / det;
        Vec tvec = vsub(r->o, tr->p1);
        float u = (float)(vdot(tvec, pvec) * inv_det);
        if (u < 0.0 || u > 1.0) return 0.0f;
        Vec qvec = vxcross(tvec, e1);
        float v = (float)(vdot(r->d, qvec) * inv_det);
        if (v < 0.0 || u + v > 1.0) return 0.0f;
        float t = (float)(vdot(e2, qvec) * inv_det);
        if (t > OR_EPSILON) return t;
        return 0.0f;
    }
}

static inline float or_shape_intersect(const Shape *s, const Ray *r, int sem) {
    if (s->type == SPHERE) return or_sphere_intersect(&s->s, r);
    if (s->type == TRIANGLE) return or_triangle_intersect(&s->t, r, sem);
    return 0.f;
}

/* intersection_bound_test rendering_kernel.cl:140-192: slab test, no clip to [0, t]. */
int or_bound_test(const Ray *r, const Bound *bound) {
    float x_a = 1.0f / r->d.x, y_a = 1.0f / r->d.y, z_a = 1.0f / r->d.z;
    float x_e = r->o.x, y_e = r->o.y, z_e = r->o.z;
    float t_xmin, t_xmax, t_ymin, t_ymax, t_zmin, t_zmax;
    if (x_a >= 0) { t_xmin = (bound->min_x - x_e) * x_a; t_xmax = (bound->max_x - x_e) * x_a; }
    else          { t_xmin = (bound->max_x - x_e) * x_a; t_xmax = (bound->min_x - x_e) * x_a; }
    if (y_a >= 0) { t_ymin = (bound->min_y - y_e) * y_a; t_ymax = (bound->max_y - y_e) * y_a; }
    else          { t_ymin = (bound->max_y - y_e) * y_a; t_ymax = (bound->min_y - y_e) * y_a; }
    if (z_a >= 0) { t_zmin = (bound->min_z - z_e) * z_a; t_zmax = (bound->max_z - z_e) * z_a; }
    else          { t_zmin = (bound->max_z - z_e) * z_a; t_zmax = (bound->min_z - z_e) * z_a; }
    float t_min = or_max2(t_xmin, or_max2(t_ymin, t_zmin));
    float t_max = or_min2(t_xmax, or_min2(t_ymax, t_zmax));
    return t_min <= t_max;
}

/* Acceleration description for the traversal restatements. */
typedef struct {
    int mode;                  /* 0 none, 1 BVH (reference layout), 2 KD heap (reference layout),
                                  3 KD depth-first int32 layout (or_kd_flatten_dfs) */
    const BVHNodeGPU *btn, *btl;
    const KDNodeGPU *kng;
    const int *kn;
    const int *kd_node5;
    const Bound *kd_box;
    int quirks;                /* 1: reproduce the reference literally (Q2 both-leaf read, 18-entry
                                  stack that gives up on overflow, first-found wins ties) */
    int shadow_accel;          /* 1 (mode 1, quirks 0): answer IntersectP through the BVH instead of
                                  the reference's brute-force scan.  The answer is the same predicate
                                  over the same shapes (0..shapeCnt-2), so only the time differs. */
} OrAccel;

static inline void or_take(float d, unsigned int id, float *t, unsigned int *best, int quirks) {
    if (d != 0.0f) {
        if (d < *t || (!quirks && d == *t && id > *best)) { *t = d; *best = id; }
    }
}

/* Intersect rendering_kernel.cl:194-363 (geomfunc.h:184-335).  Returns hit flag; t/id out.
 * Without quirks, every mode returns the brute-force result with ties going to the higher
 * shape index, the tie rule of mode 0. */
int or_intersect(const Shape *shapes, unsigned int shapeCnt, const OrAccel *acc, const Ray *r,
                 float *t, unsigned int *id, int sem) {
    const float inf = 1e20f;
    *t = inf;
    unsigned int best = 0xFFFFFFFFu;
    int hit = 0;
    if (acc == nullptr || acc->mode == 0) {
        for (unsigned int i = shapeCnt; i--;) {
            float d = or_shape_intersect(&shapes[i], r, sem);
            if (d != 0.f && d < *t) { *t = d; best = i; }
        }
        if (*t < inf) { *id = best; return 1; }
        return 0;
    }
    const int quirks = acc->quirks;
    if (acc->mode == 1) {
        if (shapeCnt == 1) {
            float d = or_shape_intersect(&shapes[acc->btl[0].nShape], r, sem);
            or_take(d, (unsigned)acc->btl[0].nShape, t, &best, quirks);
            if (best != 0xFFFFFFFFu) { *id = best; return 1; }
            return 0;
        }
        std::vector<int> stack;
        stack.reserve(64);
        stack.push_back(0);
        while (!stack.empty()) {
            int n = stack.back(); stack.pop_back();
            const BVHNodeGPU &nd = acc->btn[n];
            if (!or_bound_test(r, &nd.bound)) continue;
            if (quirks) {
                if (nd.leaf == 0) {
                    stack.push_back(nd.nRight);
                    stack.push_back(nd.nLeft);
                    if (stack.size() > 18) return 0;   /* topIndex < 0: clheader.h:11 */
                } else if (nd.leaf == 2) {
                    for (int c = 0; c < 2; c++) {
                        int li = c == 0 ? nd.nLeft : nd.nRight;
                        unsigned int sid = (unsigned)acc->btl[li].nShape;
                        float d = or_shape_intersect(&shapes[sid], r, sem);
                        if (d != 0.0f) { if (d < *t) { *t = d; best = sid; } hit = 1; }
                    }
                }
                continue;
            }
            bool lleaf = nd.nLeft == nd.min, rleaf = nd.nRight == nd.max;
            if (rleaf) {
                unsigned int sid = (unsigned)acc->btl[nd.nRight].nShape;
                or_take(or_shape_intersect(&shapes[sid], r, sem), sid, t, &best, 0);
            } else {
                stack.push_back(nd.nRight);
            }
            if (lleaf) {
                unsigned int sid = (unsigned)acc->btl[nd.nLeft].nShape;
                or_take(or_shape_intersect(&shapes[sid], r, sem), sid, t, &best, 0);
            } else {
                stack.push_back(nd.nLeft);
            }
        }
        if (quirks) { if (hit) *id = best; return hit; }
        if (best != 0xFFFFFFFFu) { *id = best; return 1; }
        return 0;
    }
    if (acc->mode == 2) {
        std::vector<int> stack;
        stack.push_back(1);
        while (!stack.empty()) {
            int n = stack.back(); stack.pop_back();
            const KDNodeGPU &nd = acc->kng[n];
            if (!or_bound_test(r, &nd.bound)) continue;
            if (nd.leaf == 0) {
                stack.push_back(nd.nRight);
                stack.push_back(nd.nLeft);
                if (quirks && stack.size() > 18) return 0;
            } else if (nd.leaf == 1) {
                for (int i = nd.min; i < nd.max; i++) {
                    unsigned int sid = (unsigned)acc->kn[i];
                    float d = or_shape_intersect(&shapes[sid], r, sem);
                    if (quirks) { if (d != 0.0f) { if (d < *t) { *t = d; best = sid; } hit = 1; } }
                    else or_take(d, sid, t, &best, 0);
                }
            }
        }
        if (quirks) { if (hit) *id = best; return hit; }
        if (best != 0xFFFFFFFFu) { *id = best; return 1; }
        return 0;
    }
    /* mode 3: depth-first int32 KD layout */
    std::vector<int> stack;
    stack.push_back(0);
    while (!stack.empty()) {
        int n = stack.back(); stack.pop_back();
        const int *nd = acc->kd_node5 + n * 5;
        if (!or_bound_test(r, &acc->kd_box[n])) continue;
        if (nd[0] == 0) {
            if (nd[2] >= 0) stack.push_back(nd[2]);
            if (nd[1] >= 0) stack.push_back(nd[1]);
        } else {
            for (int i = nd[3]; i < nd[3] + nd[4]; i++) {
                unsigned int sid = (unsigned)acc->kn[i];
                or_take(or_shape_intersect(&shapes[sid], r, sem), sid, t, &best, 0);
            }
        }
    }
    if (best != 0xFFFFFFFFu) { *id = best; return 1; }
    return 0;
}

/* IntersectP rendering_kernel.cl:375-395 (geomfunc.h:347-365): brute force over shapes
 * shapeCnt-2 .. 0 (the last shape is never tested) in every acceleration mode. */
int or_intersect_p(const Shape *shapes, unsigned int shapeCnt, const Ray *r, float maxt, int sem) {
    if (shapeCnt == 0) return 0;
    for (unsigned int i = shapeCnt - 1; i--;) {
        float d = or_shape_intersect(&shapes[i], r, sem);
        if (d != 0.f && d < maxt) return 1;
    }
    return 0;
}

/* IntersectP's predicate (a shape i < shapeCnt-1 hit at 0 != t < maxt) answered by a walk of the
 * reference-layout BVH with the closest-hit walk's box test and child rule.  Every shape the ray hits
 * lies in a box the walk visits, so the answer equals or_intersect_p's (checked against it on random
 * rays in the tests). */
static int or_intersect_p_bvh(const Shape *shapes, unsigned int shapeCnt, const OrAccel *acc, const Ray *r,
                              float maxt, int sem) {
    if (shapeCnt <= 1) return 0;
    auto test = [&](int leaf) {
        const unsigned int sid = (unsigned)acc->btl[leaf].nShape;
        if (sid == shapeCnt - 1) return 0;
        const float d = or_shape_intersect(&shapes[sid], r, sem);
        return (d != 0.f && d < maxt) ? 1 : 0;
    };
    std::vector<int> stack;
    stack.reserve(64);
    stack.push_back(0);
    while (!stack.empty()) {
        const int n = stack.back();
        stack.pop_back();
        const BVHNodeGPU &nd = acc->btn[n];
        if (!or_bound_test(r, &nd.bound)) continue;
        if (nd.nRight == nd.max) { if (test(nd.nRight)) return 1; } else stack.push_back(nd.nRight);
        if (nd.nLeft == nd.min) { if (test(nd.nLeft)) return 1; } else stack.push_back(nd.nLeft);
    }
    return 0;
}

/* sin and cos of a sampling angle.  Kernel semantics (sem 0): OpenCL leaves sin/cos to the
 * implementation within 4 ulp; the kernel semantics fix one evaluation in binary32 -- Cody-Waite
 * reduction x - k (P1 + P2 + P3), k = rint(x 2/pi), with pi/2 split so that k P1 and k P2 are exact,
 * then sin r = r - r^3/3! + ... + r^9/9! and cos r = 1 - r^2/2! + ... - r^10/10! in Horner form on
 * [-pi/4, pi/4].  Host-twin semantics (sem 1): the C library's sinf/cosf, as the reference's host
 * build calls them (geomfunc.h). */
void or_sincos(float xf, float *s, float *c, int sem) {
    if (sem != 0) {
        *s = sinf(xf);
        *c = cosf(xf);
        return;
    }
    const float two_over_pi = 0.6366197466850281f;
    const float p1 = 1.5703125f, p2 = 4.837512969970703e-4f, p3 = 7.549790126404332e-8f;
    const float k = rintf(xf * two_over_pi);
    const float r = ((xf - k * p1) - k * p2) - k * p3;
    const float r2 = r * r;
    const float sin_r =
        r + r * r2 *
                (-0.1666666716337204f +
                 r2 * (0.008333333767950535f + r2 * (-0.00019841270113829523f + r2 * 2.7557318844628753e-06f)));
    const float cos_r =
        1.0f + r2 * (-0.5f + r2 * (0.0416666679084301f +
                                   r2 * (-0.0013888889225199819f +
                                         r2 * (2.4801587642286904e-05f + r2 * -2.755731998149713e-07f))));
    switch ((int)k & 3) {             /* x = k pi/2 + r */
        case 0: *s = sin_r; *c = cos_r; break;
        case 1: *s = cos_r; *c = -sin_r; break;
        case 2: *s = -sin_r; *c = -cos_r; break;
        default: *s = -cos_r; *c = sin_r; break;
    }
}

/* UniformSampleSphere rendering_kernel.cl:127-135 */
static Vec or_uniform_sample_sphere(float u1, float u2, int sem) {
    const float zz = 1.f - 2.f * u1;
    float q = 1.f - zz * zz;
    const float r = sqrtf(0.f < q ? q : 0.f);   /* OpenCL max(0,q) / vec.h max macro agree off NaN */
    const float phi = 2.f * OR_PI * u2;
    float sphi, cphi;
    or_sincos(phi, &sphi, &cphi, sem);
    return v3(r * cphi, r * sphi, zz);
}

static inline float or_clampf(float x, float a, float b, int sem) {
    if (sem == 0) return fminf(fmaxf(x, a), b);          /* OpenCL clamp builtin */
    return (x < a) ? a : ((x > b) ? b : x);            /* vec.h:43 macro */
}
static inline float or_sign(float x, int sem) {
    if (sem == 0) { if (x > 0.f) return 1.f; if (x < 0.f) return -1.f; return x == x ? x : 0.f; }
    return (x > 0) ? 1.f : -1.f;                        /* vec.h:44 macro */
}

/* Statistics of a radiance run: closest-hit queries, any-hit (shadow) queries. */
typedef struct { unsigned long long closest, anyhit; } OrStats;

/* The two GetRandom calls that feed UniformSampleTriangle / UniformSampleSphere are arguments of
 * one call (rendering_kernel.cl:439,459; geomfunc.h:398,418), so their order is unspecified C.
 * Clang-based OpenCL compilers evaluate arguments left to right (u1 drawn first: sem 0); GCC and
 * MSVC x64 host builds evaluate them right to left (u2 drawn first: sem 1). */
static inline void or_draw_pair(unsigned int *s0, unsigned int *s1, float *u1, float *u2, int sem) {
    if (sem == 0) {
        *u1 = or_get_random(s0, s1);
        *u2 = or_get_random(s0, s1);
    } else {
        *u2 = or_get_random(s0, s1);
        *u1 = or_get_random(s0, s1);
    }
}

/* SampleLights rendering_kernel.cl:397-493 */
static Vec or_sample_lights(const Shape *shapes, unsigned int shapeCnt, unsigned int lightCnt,
                            unsigned int *s0, unsigned int *s1, const Vec &hitPoint,
                            const Vec &normal, int sem, OrStats *st, const OrAccel *acc) {
    const bool bvh_shadow = acc && acc->shadow_accel && acc->mode == 1 && !acc->quirks && shapeCnt > 1;
    Vec result = v3(0.f, 0.f, 0.f);
    unsigned int i, visited;
    for (i = 0, visited = 0; i < shapeCnt && visited < lightCnt; i++) {
        const Shape *light = &shapes[i];
        if (viszero(light->e)) continue;
        visited++;
        Ray shadowRay;
        shadowRay.o = hitPoint;
        Vec lightPoint, lightN;
        if (light->type == TRIANGLE) {
            float u1, u2;
            or_draw_pair(s0, s1, &u1, &u2, sem);
            float su1 = sqrtf(u1);                       /* UniformSampleTriangle :369-373 */
            float bu = 1.0f - su1, bv = u2 * su1;
            Vec e1 = vsub(light->t.p2, light->t.p1), e2 = vsub(light->t.p3, light->t.p1);
            e1 = vsmul(bu, e1);
            e2 = vsmul(bv, e2);
            lightPoint = vadd(vadd(light->t.p1, e1), e2);
            lightN = vnorm(vxcross(e1, e2));
        } else {
            float u1, u2;
            or_draw_pair(s0, s1, &u1, &u2, sem);
            Vec usp = or_uniform_sample_sphere(u1, u2, sem);
            lightPoint = vadd(vsmul(light->s.rad, usp), light->s.p);
            lightN = usp;
        }
        shadowRay.d = vsub(lightPoint, hitPoint);
        const float len = sqrtf(vdot(shadowRay.d, shadowRay.d));
        shadowRay.d = vsmul(1.f / len, shadowRay.d);
        float wo = vdot(shadowRay.d, lightN);
        if (wo > 0.f) continue;
        wo = -wo;
        wo = or_clampf(wo, 0.f, 1.f, sem);
        const float wi = vdot(shadowRay.d, normal);
        if (wi > 0.f) {
            if (st) st->anyhit++;
            const int occluded = bvh_shadow ? or_intersect_p_bvh(shapes, shapeCnt, acc, &shadowRay, len - OR_EPSILON, sem)
                                            : or_intersect_p(shapes, shapeCnt, &shadowRay, len - OR_EPSILON, sem);
            if (!occluded) {
                const float s = light->area * wi * wo / (len * len);
                result = vadd(result, vsmul(s, light->e));
            }
        }
    }
    return result;
}

/* RadiancePathTracing + RadianceOnePathTracing rendering_kernel.cl:495-820
 * (geomfunc.h:461-729).  max_depth is MAX_DEPTH (geom.h:32): max_depth+1 segments. */
static Vec or_radiance(const Shape *shapes, unsigned int shapeCnt, unsigned int lightCnt,
                       const OrAccel *acc, const Ray *startRay, unsigned int *s0, unsigned int *s1,
                       int max_depth, int sem, OrStats *st) {
    Ray cur = *startRay;
    Vec rad = v3(0.f, 0.f, 0.f), thr = v3(1.f, 1.f, 1.f);
    int specularBounce = 1;
    for (unsigned int depth = 0;; ++depth) {
        if (depth > (unsigned int)max_depth) return rad;
        float t;
        unsigned int id = 0;
        if (st) st->closest++;
        if (!or_intersect(shapes, shapeCnt, acc, &cur, &t, &id, sem)) return rad;
        Vec hitPoint = vadd(cur.o, vsmul(t, cur.d));
        const Shape &s = shapes[id];
        Vec normal;
        if (s.type == SPHERE) normal = vsub(hitPoint, s.s.p);
        else normal = vxcross(vsub(s.t.p2, s.t.p1), vsub(s.t.p3, s.t.p1));
        normal = vnorm(normal);
        const float dp = vdot(normal, cur.d);
        const float invSignDP = -1.f * or_sign(dp, sem);
        Vec nl = vsmul(invSignDP, normal);
        if (!viszero(s.e)) {
            if (specularBounce) {
                Vec eCol = vsmul(fabsf(dp), s.e);
                eCol = vmul(thr, eCol);
                rad = vadd(rad, eCol);
            }
            return rad;
        }
        if (s.refl == DIFF) {
            specularBounce = 0;
            thr = vmul(thr, s.c);
            Vec Ld = or_sample_lights(shapes, shapeCnt, lightCnt, s0, s1, hitPoint, nl, sem, st, acc);
            Ld = vmul(thr, Ld);
            rad = vadd(rad, Ld);
            float r1 = 2.f * OR_PI * or_get_random(s0, s1);
            float r2 = or_get_random(s0, s1);
            float r2s = sqrtf(r2);
            Vec w = nl;
            Vec a = (fabsf(w.x) > .1f) ? v3(0.f, 1.f, 0.f) : v3(1.f, 0.f, 0.f);
            Vec u = vnorm(vxcross(a, w));
            Vec v = vxcross(w, u);
            float sr1, cr1;
            or_sincos(r1, &sr1, &cr1, sem);
            u = vsmul(cr1 * r2s, u);
            v = vsmul(sr1 * r2s, v);
            Vec newDir = vadd(u, v);
            w = vsmul(sqrtf(1 - r2), w);
            newDir = vadd(newDir, w);
            cur.o = hitPoint; cur.d = newDir;
            continue;
        } else if (s.refl == SPEC) {
            specularBounce = 1;
            Vec newDir = vsmul(2.f * vdot(normal, cur.d), normal);
            newDir = vsub(cur.d, newDir);
            thr = vmul(thr, s.c);
            cur.o = hitPoint; cur.d = newDir;
            continue;
        } else {
            specularBounce = 1;
            Vec newDir = vsmul(2.f * vdot(normal, cur.d), normal);
            newDir = vsub(cur.d, newDir);
            Ray reflRay; reflRay.o = hitPoint; reflRay.d = newDir;
            int into = (vdot(normal, nl) > 0);
            float nc = 1.f, nt = 1.5f;
            float nnt = into ? nc / nt : nt / nc;
            float ddn = vdot(cur.d, nl);
            float cos2t = 1.f - nnt * nnt * (1.f - ddn * ddn);
            if (cos2t < 0.f) {
                thr = vmul(thr, s.c);
                cur = reflRay;
                continue;
            }
            float kk = (float)(into ? 1 : -1) * (ddn * nnt + sqrtf(cos2t));
            Vec nkk = vsmul(kk, normal);
            Vec transDir = vsmul(nnt, cur.d);
            transDir = vsub(transDir, nkk);
            transDir = vnorm(transDir);
            float a = nt - nc, b = nt + nc;
            float R0 = a * a / (b * b);
            float c = 1 - (into ? -ddn : vdot(transDir, normal));
            float Re = R0 + (1 - R0) * c * c * c * c * c;
            float Tr = 1.f - Re;
            float P = .25f + .5f * Re;
            float RP = Re / P;
            float TP = Tr / (1.f - P);
            if (or_get_random(s0, s1) < P) {
                thr = vsmul(RP, thr);
                thr = vmul(thr, s.c);
                cur = reflRay;
            } else {
                thr = vsmul(TP, thr);
                thr = vmul(thr, s.c);
                cur.o = hitPoint; cur.d = transDir;
            }
            continue;
        }
    }
}

/* GenerateCameraRay rendering_kernel.cl:1042-1070 */
static Ray or_camera_ray(const Camera *cam, unsigned int *s0, unsigned int *s1, int width,
                         int height, int x, int y) {
    const float invWidth = 1.f / (float)width;
    const float invHeight = 1.f / (float)height;
    const float r1 = or_get_random(s0, s1) - .5f;
    const float r2 = or_get_random(s0, s1) - .5f;
    const float kcx = ((float)x + r1) * invWidth - .5f;
    const float kcy = ((float)y + r2) * invHeight - .5f;
    Vec rdir = v3(cam->x.x * kcx + cam->y.x * kcy + cam->dir.x,
                  cam->x.y * kcx + cam->y.y * kcy + cam->dir.y,
                  cam->x.z * kcx + cam->y.z * kcy + cam->dir.z);
    Vec rorig = vadd(vsmul(0.1f, rdir), cam->orig);
    rdir = vnorm(rdir);
    Ray ray; ray.o = rorig; ray.d = rdir;
    return ray;
}

/* toInt vec.h:60 with the OpenCL clamp builtin (sem 0) or the vec.h clamp macro (sem 1). */
static inline int or_to_int(float x, int sem) {
    return (int)(powf(or_clampf(x, 0.f, 1.f, sem), 1.f / 2.2f) * 255.f + .5f);
}

/* GenerateCameraRay for each listed pixel id (x = gid % width, y = gid / width), drawing from and
 * advancing that pixel's seeds[2 gid], seeds[2 gid + 1] as RadianceGPU does (rendering_kernel.cl:1100-1110). */
void or_camera_rays(const Camera *cam, unsigned int *seeds, int width, int height, const int *pix, int n,
                    Ray *out) {
    for (int q = 0; q < n; q++) {
        const int gid = pix[q], x = gid % width, y = gid / width;
        unsigned int s0 = seeds[2 * gid], s1 = seeds[2 * gid + 1];
        out[q] = or_camera_ray(cam, &s0, &s1, width, height, x, y);
        seeds[2 * gid] = s0;
        seeds[2 * gid + 1] = s1;
    }
}

/* One pixel of RadianceGPU rendering_kernel.cl:1100-1149, for `spp` successive launches starting
 * at currentSample.  colors/seeds/pixels are the full frame buffers. */
static void or_pixel(int gid, Vec *colors, unsigned int *seeds, const Camera *cam,
                     const Shape *shapes, unsigned int shapeCnt, unsigned int lightCnt,
                     const OrAccel *acc, int width, int height, int currentSample, int spp,
                     unsigned int *pixels, int max_depth, int sem, OrStats *st) {
    const int x = gid % width, y = gid / width;
    unsigned int s0 = seeds[2 * gid], s1 = seeds[2 * gid + 1];
    const int i = (height - y - 1) * width + x;
    for (int k = 0; k < spp; k++) {
        const int cs = currentSample + k;
        Ray ray = or_camera_ray(cam, &s0, &s1, width, height, x, y);
        Vec r = or_radiance(shapes, shapeCnt, lightCnt, acc, &ray, &s0, &s1, max_depth, sem, st);
        if (cs == 0) {
            colors[i] = r;
        } else {
            const float k1 = (float)cs;
            const float k2 = 1.f / ((float)cs + 1.f);
            colors[i].x = (colors[i].x * k1 + r.x) * k2;
            colors[i].y = (colors[i].y * k1 + r.y) * k2;
            colors[i].z = (colors[i].z * k1 + r.z) * k2;
        }
    }
    if (pixels)
        pixels[y * width + x] = (unsigned int)or_to_int(colors[i].x, sem) |
                                ((unsigned int)or_to_int(colors[i].y, sem) << 8) |
                                ((unsigned int)or_to_int(colors[i].z, sem) << 16) | 0xff000000u;
    seeds[2 * gid] = s0;
    seeds[2 * gid + 1] = s1;
}

/* `spp` RadianceGPU launches over the whole frame (or over the `npix` pixel ids in pix_list when
 * pix_list != NULL).  OpenMP over pixels; every pixel is independent, so the result does not
 * depend on the thread count.  stats2 (may be NULL) receives {closest, any-hit} query counts. */
void or_radiance_frame(Vec *colors, unsigned int *seeds, const Camera *cam, const Shape *shapes,
                       unsigned int shapeCnt, unsigned int lightCnt, const OrAccel *acc, int width,
                       int height, int currentSample, int spp, unsigned int *pixels, int max_depth,
                       int sem, const int *pix_list, int npix, int nthreads,
                       unsigned long long *stats2) {
    const int total = pix_list ? npix : width * height;
    unsigned long long c0 = 0, c1 = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : c0, c1)
    for (int q = 0; q < total; q++) {
        int gid = pix_list ? pix_list[q] : q;
        OrStats st = {0, 0};
        or_pixel(gid, colors, seeds, cam, shapes, shapeCnt, lightCnt, acc, width, height,
                 currentSample, spp, pixels, max_depth, sem, &st);
        c0 += st.closest; c1 += st.anyhit;
    }
    if (stats2) { stats2[0] = c0; stats2[1] = c1; }
}

/* Batch closest-hit / any-hit queries (rays as Ray[n]). */
void or_trace_closest(const Shape *shapes, unsigned int shapeCnt, const OrAccel *acc,
                      const Ray *rays, int n, float *t_out, int *id_out, int sem, int nthreads) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
    for (int i = 0; i < n; i++) {
        float t; unsigned int id = 0;
        if (or_intersect(shapes, shapeCnt, acc, &rays[i], &t, &id, sem)) { t_out[i] = t; id_out[i] = (int)id; }
        else { t_out[i] = 0.f; id_out[i] = -1; }
    }
}
void or_trace_any_bvh(const Shape *shapes, unsigned int shapeCnt, const OrAccel *acc, const Ray *rays,
                      const float *maxt, int n, int *hit_out, int sem, int nthreads) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
    for (int i = 0; i < n; i++) hit_out[i] = or_intersect_p_bvh(shapes, shapeCnt, acc, &rays[i], maxt[i], sem);
}
void or_trace_any(const Shape *shapes, unsigned int shapeCnt, const Ray *rays, const float *maxt,
                  int n, int *hit_out, int sem, int nthreads) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
    for (int i = 0; i < n; i++) hit_out[i] = or_intersect_p(shapes, shapeCnt, &rays[i], maxt[i], sem);
}

/* UpdateCamera displayfunc.cpp:732-756 */
void or_update_camera(Camera *cam) {
    float xd = sinf(cam->yaw) * cosf(cam->pitch);
    float yd = sinf(cam->pitch);
    float zd = cosf(cam->yaw) * cosf(cam->pitch);
    Vec dtc = v3(xd, yd, zd);
    cam->dir = vsmul(-1.f, dtc);
    cam->dir = vnorm(cam->dir);
    const Vec up = v3(0.f, 1.f, 0.f);
    const float fov = (OR_PI / 180.f) * 45.f;
    cam->x = vnorm(vxcross(cam->dir, up));
    cam->x = vsmul(cam->width * fov / cam->height, cam->x);
    cam->y = vnorm(vxcross(cam->x, cam->dir));
    cam->y = vsmul(fov, cam->y);
}

/* Accel descriptor helpers for ctypes callers */
size_t or_accel_size(void) { return sizeof(OrAccel); }

} /* extern "C" */

extern "C" {
/* One path (RadiancePathTracing) from a given ray and seed pair; result3 = radiance. */
void or_radiance_path(const Shape *shapes, unsigned int shapeCnt, unsigned int lightCnt, const OrAccel *acc,
                      const Ray *ray, unsigned int *s0, unsigned int *s1, int max_depth, int sem, float *result3) {
    Vec r = or_radiance(shapes, shapeCnt, lightCnt, acc, ray, s0, s1, max_depth, sem, nullptr);
    result3[0] = r.x; result3[1] = r.y; result3[2] = r.z;
}
}

extern "C" {
/* SampleLights for one hit point (result3 = Ld); exposed for unit tests. */
void or_sample_lights_at(const Shape *shapes, unsigned int shapeCnt, unsigned int lightCnt, unsigned int *s0,
                         unsigned int *s1, const Vec *hitPoint, const Vec *normal, int sem, float *result3) {
    Vec r = or_sample_lights(shapes, shapeCnt, lightCnt, s0, s1, *hitPoint, *normal, sem, nullptr, nullptr);
    result3[0] = r.x; result3[1] = r.y; result3[2] = r.z;
}
}
