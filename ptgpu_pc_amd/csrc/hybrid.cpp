// hybrid.cpp -- the reference's CPU_PARTRENDERING mode (smallptGPU.cpp:636-884): a frame split into
// tiles (160 x 120 in the reference's DrawFrame), some rendered on the GPU (RadianceBoxGPU,
// rendering_kernel.cl:1152-1236) and the rest on the host's cores (DrawBox), with the split between the
// two sides following the measured time per tile of each (DrawAllBoxes).
//
// Here the GPU tiles of a frame go out in one launch and the host renders its tiles while that launch
// runs, instead of one blocking round trip per tile.  The host runs the kernels' own per-path math
// (render_device.hpp, compiled for the host as well) over the same prepared tree, with the GPU's seed
// slot and pixel pack, so a frame comes out with the bits of a GPU-only frame whichever side renders a
// tile.  The reference's host tiles differ there (geomfunc.h's FP64 triangle test, the flipped seed
// slot of smallptGPU.cpp:668-669, no alpha in the pack); those differences are not reproduced.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

#include "render_device.hpp"

namespace ptgpu {
namespace {

// A prepared traversal structure in host memory: the same records as on the device.
struct HostAccel {
    const TNode *nodes = nullptr;
    int nnodes = 0;
    const Prim *prims = nullptr;
    int nprims = 0;
    int root = ~0, root_cnt = 0;
    int nshapes = 0;
    int mode = PTGPU_ACCEL_NONE;
};

struct HostFrame {
    const Shape *shapes;
    const int *lights;
    int nlights;
    const Camera *camera;
    Vec *colors;
    unsigned *seeds;
    unsigned *pixels;   // may be NULL
    int width, height, currentSample, spp, max_depth;
};

constexpr int kHostStack = 128;   // binary walk: one push per level, trees are at most 62 levels high
constexpr float kInfF = __builtin_inff();   // the slab test's miss

// Closest hit: the kernels' binary walk (trace_closest) or brute force (brute_closest); the answer is the
// brute-force one either way (smaller t, then the higher shape index).
int host_closest(const HostAccel &A, f3 o, f3 d, float *tout) {
    float tb = kInf;
    int bid = -1;
    if (A.mode == PTGPU_ACCEL_NONE) {
        for (int i = A.nprims - 1; i >= 0; i--) {
            int id;
            const float t = prim_isect(A.prims, i, o, d, &id);
            if (t != 0.f && t < tb) { tb = t; bid = id; }
        }
        *tout = tb;
        return bid;
    }
    auto test_leaf = [&](int first, int cnt) {
        if (first < 0 || cnt < 0 || cnt > A.nprims - first) return;
        for (int q = 0; q < cnt; q++) {
            int id;
            const float t = prim_isect(A.prims, first + q, o, d, &id);
            if (t != 0.f && (t < tb || (t == tb && id > bid))) { tb = t; bid = id; }
        }
    };
    if (A.root < 0) {
        test_leaf(~A.root, A.root_cnt);
        *tout = tb;
        return bid;
    }
    const f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    int stk[kHostStack];
    int node = A.root, sp = 0;
    for (unsigned guard = 0; guard < kMaxVisits; guard++) {
        if ((unsigned)node >= (unsigned)A.nnodes) break;
        const TNode &n = A.nodes[node];
        float t0 = slab(n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, o, inv, tb);
        float t1 = slab(n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, o, inv, tb);
        if (n.r.x < 0 && t0 != kInfF) { test_leaf(~n.r.x, n.r.z); t0 = kInfF; }
        if (n.r.y < 0 && t1 != kInfF) { test_leaf(~n.r.y, n.r.w); t1 = kInfF; }
        const bool h0 = t0 != kInfF, h1 = t1 != kInfF;
        if (h0 && h1) {
            const bool first0 = t0 <= t1;
            if (sp < kHostStack) stk[sp++] = first0 ? n.r.y : n.r.x;
            node = first0 ? n.r.x : n.r.y;
        } else if (h0) {
            node = n.r.x;
        } else if (h1) {
            node = n.r.y;
        } else {
            if (sp == 0) break;
            node = stk[--sp];
        }
    }
    *tout = tb;
    return bid;
}

// IntersectP (rendering_kernel.cl:375-395): a shape other than the last hit at 0 != t < maxt.
bool host_any(const HostAccel &A, f3 o, f3 d, float maxt) {
    if (A.mode == PTGPU_ACCEL_NONE) {
        for (int i = A.nprims - 2; i >= 0; i--) {
            int id;
            const float t = prim_isect(A.prims, i, o, d, &id);
            if (t != 0.f && t < maxt) return true;
        }
        return false;
    }
    const int skip = A.nshapes - 1;
    auto test_leaf = [&](int first, int cnt) -> bool {
        if (first < 0 || cnt < 0 || cnt > A.nprims - first) return false;
        for (int q = 0; q < cnt; q++) {
            int id;
            const float t = prim_isect(A.prims, first + q, o, d, &id);
            if (t != 0.f && t < maxt && id != skip) return true;
        }
        return false;
    };
    if (A.root < 0) return test_leaf(~A.root, A.root_cnt);
    const f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    int stk[kHostStack];
    int node = A.root, sp = 0;
    for (unsigned guard = 0; guard < kMaxVisits; guard++) {
        if ((unsigned)node >= (unsigned)A.nnodes) break;
        const TNode &n = A.nodes[node];
        float t0 = slab(n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, o, inv, maxt);
        float t1 = slab(n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, o, inv, maxt);
        if (n.r.x < 0 && t0 != kInfF) { if (test_leaf(~n.r.x, n.r.z)) return true; t0 = kInfF; }
        if (n.r.y < 0 && t1 != kInfF) { if (test_leaf(~n.r.y, n.r.w)) return true; t1 = kInfF; }
        const bool h0 = t0 != kInfF, h1 = t1 != kInfF;
        if (h0 && h1) {
            if (sp < kHostStack) stk[sp++] = n.r.y;
            node = n.r.x;
        } else if (h0) {
            node = n.r.x;
        } else if (h1) {
            node = n.r.y;
        } else {
            if (sp == 0) break;
            node = stk[--sp];
        }
    }
    return false;
}

// One pixel of RadianceGPU (rendering_kernel.cl:1100-1149) for spp consecutive samples: camera ray,
// RadiancePathTracing (rendering_kernel.cl:495-820) with the kernels' functions, running mean, pack.
void host_pixel(const HostAccel &A, const HostFrame &F, int gid) {
    const int W = F.width, H = F.height;
    const int x = gid % W, y = gid / W;
    const int ci = (H - y - 1) * W + x;
    unsigned s0 = F.seeds[2 * gid], s1 = F.seeds[2 * gid + 1];
    f3 col = F.currentSample > 0 ? ldvec(F.colors + ci) : mk(0.f, 0.f, 0.f);
    for (int k = 0; k < F.spp; k++) {
        f3 ro, rd;
        camera_ray(F.camera, s0, s1, W, H, x, y, ro, rd);
        f3 rad = mk(0.f, 0.f, 0.f), thr = mk(1.f, 1.f, 1.f);
        bool spec = true;
        for (int depth = 0; depth <= F.max_depth; depth++) {
            float t;
            const int id = host_closest(A, ro, rd, &t);
            if (id < 0) break;
            const f3 hp = add3(ro, smul(t, rd));
            const Surface S = surface_at(F.shapes + id, hp, rd);
            if (!iszero3(S.e)) {
                if (spec) rad = add3(rad, mul3(thr, smul(fabsf(S.dp), S.e)));
                break;
            }
            if (S.refl == DIFF) {
                spec = false;
                thr = mul3(thr, S.c);
                f3 Ld = mk(0.f, 0.f, 0.f);
                for (int li = 0; li < F.nlights; li++) {
                    f3 sd, contrib;
                    float len;
                    if (!sample_light(F.shapes + F.lights[li], s0, s1, hp, S.nl, &sd, &len, &contrib)) continue;
                    if (!host_any(A, hp, sd, len - kEpsilon)) Ld = add3(Ld, contrib);
                }
                Ld = mul3(thr, Ld);
                rad = add3(rad, Ld);
                rd = diffuse_dir(S.nl, s0, s1);
                ro = hp;
            } else {
                spec = true;
                specular_or_refract(S, rd, thr, s0, s1);
                ro = hp;
            }
        }
        col = running_mean(col, rad, F.currentSample + k);
    }
    float *cp = reinterpret_cast<float *>(F.colors + ci);
    cp[0] = col.x; cp[1] = col.y; cp[2] = col.z;
    if (F.pixels) F.pixels[y * W + x] = pack_rgba(col);
    F.seeds[2 * gid] = s0;
    F.seeds[2 * gid + 1] = s1;
}

// Renders the listed tiles (origins, tw x th each, clipped to the frame) on `threads` host threads.
void host_tiles(const HostAccel &A, const HostFrame &F, const std::vector<int2> &tiles, int tw, int th, int threads) {
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int k = next++; k < (int)tiles.size(); k = next++) {
            const int x0 = tiles[(size_t)k].x, y0 = tiles[(size_t)k].y;
            const int x1 = std::min(x0 + tw, F.width), y1 = std::min(y0 + th, F.height);
            for (int y = y0; y < y1; y++)
                for (int x = x0; x < x1; x++) host_pixel(A, F, y * F.width + x);
        }
    };
    const int nt = std::max(1, std::min(threads, (int)tiles.size()));
    std::vector<std::thread> pool;
    for (int i = 1; i < nt; i++) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
}

// Reference layouts -> the traversal records, on the host (the kernels k_convert_bvh / k_convert_kd /
// k_index_prims do the same on the device).  Links outside the arrays are refused.
int host_accel_from_ref(const ptgpu_radiance_args *a, int accel, std::vector<TNode> &nodes, std::vector<Prim> &prims,
                        HostAccel *A) {
    const int n = (int)a->shapeCnt;
    A->nshapes = n;
    A->mode = accel;
    if (accel == PTGPU_ACCEL_NONE) {
        prims.resize((size_t)n);
        for (int i = 0; i < n; i++) prims[(size_t)i] = prim_of(a->shapes + i, i, kLastInLeaf);
        A->prims = prims.data();
        A->nprims = n;
        return PTGPU_OK;
    }
    if (accel == PTGPU_ACCEL_BVH) {
        PT_CHECK(a->btn && a->btl, PTGPU_ERR_INVALID, "BVH mode needs btn and btl");
        prims.resize((size_t)n);
        for (int i = 0; i < n; i++) {
            const int sid = a->btl[i].nShape;
            PT_CHECK((unsigned)sid < (unsigned)n, PTGPU_ERR_INVALID, "btl: a leaf's nShape is out of range");
            prims[(size_t)i] = prim_of(a->shapes + sid, sid, kLastInLeaf);
        }
        nodes.resize((size_t)(n > 1 ? n - 1 : 0));
        for (int i = 0; i < n - 1; i++) {
            const BVHNodeGPU &b = a->btn[i];
            const bool ll = b.nLeft == b.min, rl = b.nRight == b.max;
            PT_CHECK((unsigned)b.nLeft < (unsigned)(ll ? n : n - 1) && (unsigned)b.nRight < (unsigned)(rl ? n : n - 1),
                     PTGPU_ERR_INVALID, "btn/btl: a child index is out of range");
            const Bound &L = ll ? a->btl[b.nLeft].bound : a->btn[b.nLeft].bound;
            const Bound &R = rl ? a->btl[b.nRight].bound : a->btn[b.nRight].bound;
            TNode &t = nodes[(size_t)i];
            t.a = make_float4(L.min_x, L.min_y, L.min_z, L.max_x);
            t.b = make_float4(L.max_y, L.max_z, R.min_x, R.min_y);
            t.c = make_float4(R.min_z, R.max_x, R.max_y, R.max_z);
            t.r = make_int4(ll ? ~b.nLeft : b.nLeft, rl ? ~b.nRight : b.nRight, ll ? 1 : 0, rl ? 1 : 0);
        }
        A->nodes = nodes.data();
        A->nnodes = n - 1;
        A->prims = prims.data();
        A->nprims = n;
        A->root = n >= 2 ? 0 : ~0;
        A->root_cnt = n >= 2 ? 0 : 1;
        return PTGPU_OK;
    }
    PT_CHECK(accel == PTGPU_ACCEL_KD, PTGPU_ERR_INVALID, "unknown accel mode %d", accel);
    PT_CHECK(a->kng && a->kn && a->kngCnt >= 2 && a->knCnt >= 0, PTGPU_ERR_INVALID, "KD mode needs kng and kn");
    const int kc = a->kngCnt, kn = a->knCnt;
    prims.resize((size_t)std::max(kn, 1));
    for (int h = 0; h < kn; h++) {
        PT_CHECK((unsigned)a->kn[h] < (unsigned)n, PTGPU_ERR_INVALID, "kn: a shape index is out of range");
        prims[(size_t)h] = prim_of(a->shapes + a->kn[h], a->kn[h], 0);
    }
    nodes.assign((size_t)kc, TNode{});
    auto leaf_ok = [&](const KDNodeGPU &k) { return k.min >= 0 && k.max >= k.min && k.max <= kn; };
    for (int h = 1; h < kc; h++) {
        const KDNodeGPU &k = a->kng[h];
        if (k.leaf != 0 || 2 * h + 1 >= kc) continue;
        const KDNodeGPU &cl = a->kng[2 * h], &cr = a->kng[2 * h + 1];
        const bool ll = cl.leaf == 1, rl = cr.leaf == 1;
        PT_CHECK((!ll || leaf_ok(cl)) && (!rl || leaf_ok(cr)), PTGPU_ERR_INVALID, "kng: a leaf range is out of range");
        TNode &t = nodes[(size_t)h];
        t.a = make_float4(cl.bound.min_x, cl.bound.min_y, cl.bound.min_z, cl.bound.max_x);
        t.b = make_float4(cl.bound.max_y, cl.bound.max_z, cr.bound.min_x, cr.bound.min_y);
        t.c = make_float4(cr.bound.min_z, cr.bound.max_x, cr.bound.max_y, cr.bound.max_z);
        t.r = make_int4(ll ? ~(int)cl.min : 2 * h, rl ? ~(int)cr.min : 2 * h + 1, ll ? (int)cl.max - (int)cl.min : 0,
                        rl ? (int)cr.max - (int)cr.min : 0);
    }
    const KDNodeGPU &root = a->kng[1];
    if (root.leaf == 1) {
        PT_CHECK(leaf_ok(root), PTGPU_ERR_INVALID, "kng: a leaf range is out of range");
        A->root = ~(int)root.min;
        A->root_cnt = (int)root.max - (int)root.min;
    } else {
        A->root = 1;
    }
    A->nodes = nodes.data();
    A->nnodes = kc;
    A->prims = prims.data();
    A->nprims = kn;
    return PTGPU_OK;
}

// the emitting shapes SampleLights visits: the first light_cnt with non-zero emission, in index order
std::vector<int> host_lights(const Shape *shapes, int n, int light_cnt) {
    std::vector<int> out;
    for (int i = 0; i < n && (int)out.size() < light_cnt; i++)
        if (shapes[i].e.x != 0.f || shapes[i].e.y != 0.f || shapes[i].e.z != 0.f) out.push_back(i);
    return out;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace
}  // namespace ptgpu

using namespace ptgpu;

struct ptgpu_hybrid {
    ptgpu_ctx *ctx = nullptr;
    int threads = 1;
    hipStream_t stream = nullptr;      // host-side copies, beside the GPU tiles' launch
    hipEvent_t start = nullptr, gpu_begin = nullptr, gpu_end = nullptr;
    DevBuf d_tiles;
    std::vector<Shape> shapes;
    std::vector<TNode> nodes;
    std::vector<Prim> prims;
    std::vector<int> lights;
    std::vector<Vec> colors;
    std::vector<unsigned> seeds, pixels;
};

extern "C" {

int ptgpu_host_radiance_box(const ptgpu_radiance_args *a, int accel, int max_depth, int x, int y, int bwidth,
                            int bheight, int threads) {
    PT_CHECK(a && a->colors && a->seeds && a->camera && a->shapes, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(a->shapeCnt >= 1 && a->width > 0 && a->height > 0 && a->currentSample >= 0, PTGPU_ERR_INVALID,
             "bad frame");
    PT_CHECK((unsigned long long)a->width * (unsigned long long)a->height < (1ull << 30), PTGPU_ERR_UNSUPPORTED,
             "frame of 2^30 pixels or more");
    PT_CHECK(x >= 0 && y >= 0 && bwidth >= 1 && bheight >= 1 && max_depth >= 0 && max_depth < 1024, PTGPU_ERR_INVALID,
             "bad box or depth");
    std::vector<TNode> nodes;
    std::vector<Prim> prims;
    HostAccel A;
    PT_TRY(host_accel_from_ref(a, accel, nodes, prims, &A));
    const std::vector<int> lights = host_lights(a->shapes, (int)a->shapeCnt, (int)a->lightCnt);
    HostFrame F{a->shapes, lights.data(), (int)lights.size(), a->camera, a->colors, a->seeds, a->pixels,
                a->width, a->height, a->currentSample, 1, max_depth};
    host_tiles(A, F, std::vector<int2>{make_int2(x, y)}, bwidth, bheight, threads < 1 ? 1 : threads);
    return PTGPU_OK;
}

int ptgpu_hybrid_create(ptgpu_ctx *ctx, int threads, ptgpu_hybrid **out) {
    PT_CHECK(ctx && out, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(threads >= 1 && threads <= 1024, PTGPU_ERR_INVALID, "threads %d out of range", threads);
    PT_HIP(hipSetDevice(ctx->device));
    ptgpu_hybrid *h = new ptgpu_hybrid();
    h->ctx = ctx;
    h->threads = threads;
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&h->start);
    if (e == hipSuccess) e = hipEventCreate(&h->gpu_begin);
    if (e == hipSuccess) e = hipEventCreate(&h->gpu_end);
    if (e != hipSuccess) {
        ptgpu_hybrid_destroy(h);
        return hip_fail(e, "ptgpu_hybrid_create", __FILE__, __LINE__);
    }
    *out = h;
    return PTGPU_OK;
}

int ptgpu_hybrid_frame(ptgpu_hybrid *h, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const ptgpu_render_args *a, int bwidth,
                       int bheight, float *rcpu, ptgpu_hybrid_stats *stats) {
    PT_CHECK(h && a && rcpu, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(bwidth >= 1 && bheight >= 1, PTGPU_ERR_INVALID, "bad tile size %dx%d", bwidth, bheight);
    PT_CHECK(a->width > 0 && a->height > 0, PTGPU_ERR_INVALID, "bad frame size %dx%d", a->width, a->height);
    PT_CHECK(!a->ray_counter && !a->trav_stats, PTGPU_ERR_UNSUPPORTED, "hybrid frames take no counters");
    ptgpu_ctx *ctx = h->ctx;
    PT_NO_CAPTURE(ctx, "a hybrid CPU/GPU frame");
    const double t0 = now_ms();
    // tiles in the reference's order and proportion (DrawAllBoxes): the first to the GPU, then the
    // host whenever GPU tiles / host tiles (counted from 1) has reached rcpu
    std::vector<int2> gpu, cpu;
    int n_gpu = 0, n_cpu = 1;
    bool cpu_turn = false;
    for (int y = 0; y < a->height; y += bheight)
        for (int x = 0; x < a->width; x += bwidth) {
            if (cpu_turn) {
                cpu.push_back(make_int2(x, y));
                n_cpu++;
            } else {
                gpu.push_back(make_int2(x, y));
                n_gpu++;
            }
            cpu_turn = (float)n_gpu / (float)n_cpu >= *rcpu;
        }
    // the traversal structure and light list, then the GPU tiles in one launch on the context's stream
    DevAccel acc;
    RenderLaunch L;
    PT_TRY(prepare_render(ctx, bvh, kd, a, &acc, &L));
    PT_HIP(hipEventRecord(h->start, ctx->stream));   // the caller's work and the preparation come first
    PT_TRY(h->d_tiles.ensure(std::max<size_t>(gpu.size(), 1) * sizeof(int2)));
    PT_HIP(hipEventRecord(h->gpu_begin, ctx->stream));
    if (!gpu.empty()) {
        PT_HIP(hipMemcpyAsync(h->d_tiles.p, gpu.data(), gpu.size() * sizeof(int2), hipMemcpyHostToDevice, ctx->stream));
        L.tiles = h->d_tiles.as<int2>();
        L.ntiles = (int)gpu.size();
        L.tile_w = bwidth;
        L.tile_h = bheight;
        PT_TRY(launch_render(ctx, acc, L));
    }
    PT_HIP(hipEventRecord(h->gpu_end, ctx->stream));
    // host tiles: what they read comes over on the hybrid's own stream while the GPU tiles render (the
    // GPU touches none of the host tiles' pixels, and only reads the scene)
    double cpu_ms = 0.0;
    if (!cpu.empty()) {
        const size_t px = (size_t)a->width * (size_t)a->height;
        hipStream_t st = h->stream;
        PT_HIP(hipStreamWaitEvent(st, h->start, 0));
        h->shapes.resize(a->shapeCnt);
        h->prims.resize((size_t)std::max(acc.nprims, 1));
        h->nodes.resize((size_t)std::max(acc.nnodes, 1));
        h->lights.resize((size_t)std::max(L.nlights, 1));
        h->seeds.resize(2 * px);
        h->colors.resize(px);
        h->pixels.resize(px);
        Camera cam;
        PT_HIP(hipMemcpyAsync(h->shapes.data(), a->shapes, a->shapeCnt * sizeof(Shape), hipMemcpyDeviceToHost, st));
        PT_HIP(hipMemcpyAsync(h->prims.data(), acc.prims, (size_t)acc.nprims * sizeof(Prim), hipMemcpyDeviceToHost, st));
        if (acc.nnodes > 0 && acc.nodes)
            PT_HIP(hipMemcpyAsync(h->nodes.data(), acc.nodes, (size_t)acc.nnodes * sizeof(TNode), hipMemcpyDeviceToHost,
                                  st));
        if (L.nlights > 0)
            PT_HIP(hipMemcpyAsync(h->lights.data(), L.lights, (size_t)L.nlights * 4, hipMemcpyDeviceToHost, st));
        PT_HIP(hipMemcpyAsync(&cam, a->camera, sizeof(Camera), hipMemcpyDeviceToHost, st));
        PT_HIP(hipMemcpyAsync(h->seeds.data(), a->seeds, px * 8, hipMemcpyDeviceToHost, st));
        if (a->currentSample > 0)
            PT_HIP(hipMemcpyAsync(h->colors.data(), a->colors, px * 12, hipMemcpyDeviceToHost, st));
        PT_HIP(hipStreamSynchronize(st));
        HostAccel A;
        A.nodes = h->nodes.data();
        A.nnodes = acc.nodes ? acc.nnodes : 0;
        A.prims = h->prims.data();
        A.nprims = acc.nprims;
        A.root = acc.root;
        A.root_cnt = acc.root_cnt;
        A.nshapes = acc.nshapes;
        A.mode = acc.mode;
        HostFrame F{h->shapes.data(), h->lights.data(), L.nlights, &cam, h->colors.data(), h->seeds.data(),
                    a->pixels ? h->pixels.data() : nullptr, a->width, a->height, a->currentSample, a->spp,
                    a->max_depth};
        const double c0 = now_ms();
        host_tiles(A, F, cpu, bwidth, bheight, h->threads);
        cpu_ms = now_ms() - c0;
        // the host tiles' rows go back: colours at the flipped row, seeds and pixels at the pixel row
        const size_t W = (size_t)a->width, H = (size_t)a->height;
        for (const int2 &t : cpu) {
            const size_t tw = std::min<size_t>((size_t)bwidth, W - (size_t)t.x);
            const size_t th = std::min<size_t>((size_t)bheight, H - (size_t)t.y);
            const size_t crow = H - (size_t)t.y - th;   // first colour row of the tile
            PT_HIP(hipMemcpy2DAsync(a->colors + crow * W + (size_t)t.x, W * 12, h->colors.data() + crow * W + (size_t)t.x,
                                    W * 12, tw * 12, th, hipMemcpyHostToDevice, st));
            PT_HIP(hipMemcpy2DAsync(a->seeds + 2 * ((size_t)t.y * W + (size_t)t.x), W * 8,
                                    h->seeds.data() + 2 * ((size_t)t.y * W + (size_t)t.x), W * 8, tw * 8, th,
                                    hipMemcpyHostToDevice, st));
            if (a->pixels)
                PT_HIP(hipMemcpy2DAsync(a->pixels + (size_t)t.y * W + (size_t)t.x, W * 4,
                                        h->pixels.data() + (size_t)t.y * W + (size_t)t.x, W * 4, tw * 4, th,
                                        hipMemcpyHostToDevice, st));
        }
        PT_HIP(hipStreamSynchronize(st));
    }
    PT_HIP(hipStreamSynchronize(ctx->stream));
    float gpu_ms = 0.f;
    PT_HIP(hipEventElapsedTime(&gpu_ms, h->gpu_begin, h->gpu_end));
    // next frame's proportion: host time per host tile over GPU time per GPU tile (DrawAllBoxes)
    if (!cpu.empty() && !gpu.empty() && gpu_ms > 0.f)
        *rcpu = (float)((cpu_ms / (double)cpu.size()) / ((double)gpu_ms / (double)gpu.size()));
    if (stats) {
        stats->gpu_tiles = (int)gpu.size();
        stats->cpu_tiles = (int)cpu.size();
        stats->gpu_ms = gpu_ms;
        stats->cpu_ms = (float)cpu_ms;
        stats->total_ms = (float)(now_ms() - t0);
    }
    return PTGPU_OK;
}

int ptgpu_hybrid_destroy(ptgpu_hybrid *h) {
    if (!h) return PTGPU_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->d_tiles.release();
    if (h->start) (void)hipEventDestroy(h->start);
    if (h->gpu_begin) (void)hipEventDestroy(h->gpu_begin);
    if (h->gpu_end) (void)hipEventDestroy(h->gpu_end);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PTGPU_OK;
}

}  // extern "C"
