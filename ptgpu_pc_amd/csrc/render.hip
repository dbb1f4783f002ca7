// render.hip -- traversal engine and path-tracing megakernel for gfx950.
//
// Semantics follow rendering_kernel.cl (RadianceGPU :1072-1150, RadiancePathTracing :755-820,
// SampleLights :397-493, Intersect :194-363, IntersectP :375-395) in FP32 with every expression
// rounded as written (the library is built with -ffp-contract=off and IEEE division/sqrt).
//
// Differences from the reference that do not change results:
//   * Closest hit traverses a two-children-per-node layout (TNode) near-child first with the slab
//     test clipped to [0, t_best]; ties in t go to the higher shape index, which is exactly what
//     the brute-force loop (i = N-1 .. 0 with a strict <) returns.
//   * IntersectP is an any-hit traversal of the same tree that ignores shape N-1, the shape the
//     reference's brute-force loop never tests (Q10).
//   * SampleLights walks a precomputed list of the first lightCnt emitters in index order, drawing
//     the same random numbers in the same order as the reference's scan over all shapes.
//   * Persistent waves: each lane owns one pixel at a time and runs its spp samples back to back
//     (the running mean and RNG state stay in registers); a lane whose pixel is finished takes
//     the next pixel from a global counter, wave-wide through one atomic (ballot + popcount).
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "lbvh.hpp"
#include "render_device.hpp"

namespace ptgpu {

// One frame's work in a batch: a prepared tree, its scene and the frame's buffers (the reference's
// RadianceGPU arguments).
struct RenderJob {
    DevAccel acc;
    Vec *colors;
    unsigned *seeds;
    const Camera *camera;
    const Shape *shapes;
    const int *lights;
    int nlights;
    int width, height, currentSample, spp, max_depth;
    unsigned *pixels;
};

struct RenderParams {
    DevAccel acc;
    Vec *colors;
    unsigned *seeds;
    const Camera *camera;
    const Shape *shapes;
    const int *lights;
    int nlights;
    int width, height, currentSample, spp, max_depth;
    unsigned *pixels;
    unsigned long long *ray_counter;
    unsigned long long *trav_stats;
    unsigned *work_counter;
    int total;
    int regroup;
    int quad;                     // traversing lanes at or below which four lanes share each query
    const RenderJob *jobs;        // batch launches: job table, pixel prefix sums (njobs + 1)
    const unsigned *job_begin;
    int njobs;
    const int2 *tiles;            // tiled launches (RadianceBoxGPU): tile origins; NULL renders the whole frame
    int tile_w, tile_h;
    float one;                    // 1.0f, opaque to the compiler (slab_hn's mixed-precision differences)
    int depth_cap;                // deepest tree the launch's stack holds (checked when acc.height_check is set)
    unsigned *status;             // the context's check words: [0] flags, [1] deepest height seen
};

// The short traversal-stack class of the binary16 walk: 36 entries per lane, so a workgroup's stack
// takes 36 KB of LDS and four workgroups fit on a CU (three with the 48-entry class).  A walk that
// would push past it marks kOverflow in its visit count and is finished against every primitive.
constexpr int kStackShort = 36;
constexpr int kStackTiny = 8;   // a test class: nearly every walk outgrows it (tuning short_stack 2)
constexpr unsigned kOverflow = 1u << 31;

// A pointer read from device memory (a batch's job table) or moved between lanes is a generic
// pointer to the compiler, and loads through it become flat loads, which count against both memory
// wait counters and make every wait cover all outstanding LDS and scalar accesses too.  Every buffer
// the render kernels read or write lives in device (global) memory, so the per-lane pointers of
// the render loop are held as global-address-space pointers (gmem): loads and stores through them
// are global ones, also after a generic-pointer helper is inlined on from_gmem's result.  to_gmem
// goes through the address bits: a cast would fold with from_gmem's back into the generic pointer.
template <class T>
using gmem = __attribute__((address_space(1))) T;
template <class T>
__device__ __forceinline__ gmem<T> *to_gmem(T *p) {
    return reinterpret_cast<gmem<T> *>(reinterpret_cast<unsigned long long>(p));
}
template <class T>
__device__ __forceinline__ T *from_gmem(gmem<T> *p) { return (T *)p; }

// A launch recorded right after a build sized its stack for a height bound (acc.height_check set).
// The built tree is checked before any work: a build error (flag 1) or a tree deeper than the stack
// holds (flag 2) is recorded in the context's status words and the launch renders nothing.  Every
// thread reads the same words, so whole workgroups return together, before any barrier.
__device__ __forceinline__ bool recorded_tree_unfit(const RenderParams &P) {
    if (!P.acc.height_check) return false;
    const unsigned err = P.acc.height_check[0], h = P.acc.height_check[2];
    if (err == 0u && h <= (unsigned)P.depth_cap) return false;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicOr(P.status, err ? 1u : 2u);
        atomicMax(P.status + 1, h);
    }
    return true;
}

// The same check over a batch's jobs (P.status is set when some job's tree was recorded right after
// its build): the launch renders nothing if any job's tree does not fit.  Block-uniform: every block
// reads the same words and decides after one barrier.
__device__ bool recorded_jobs_unfit(const RenderParams &P) {
    __shared__ int s_unfit;
    if (threadIdx.x == 0) s_unfit = 0;
    __syncthreads();
    for (int j = threadIdx.x; j < P.njobs; j += blockDim.x) {
        const gmem<const unsigned> *hc = to_gmem(P.jobs[j].acc.height_check);
        if (!hc) continue;
        const unsigned err = hc[0], h = hc[2];
        if (err == 0u && h <= (unsigned)P.depth_cap) continue;
        atomicOr(&s_unfit, 1);
        if (blockIdx.x == 0) {
            atomicOr(P.status, err ? 1u : 2u);
            atomicMax(P.status + 1, h);
        }
    }
    __syncthreads();
    return s_unfit != 0;
}

// The frame pixel of claimed work item m: pixel m of the frame, or, in a tiled launch, pixel
// m % (tile_w * tile_h) of tile m / (tile_w * tile_h) in row-major order within the tile; -1 when a
// partial tile at the frame's right or bottom edge has no pixel there.
__device__ __forceinline__ int claimed_pixel(const RenderParams &P, unsigned m) {
    if (!P.tiles) return (int)m;
    const unsigned per = (unsigned)(P.tile_w * P.tile_h);
    const int2 o = P.tiles[m / per];
    const unsigned t = m % per;
    const int x = o.x + (int)(t % (unsigned)P.tile_w), y = o.y + (int)(t / (unsigned)P.tile_w);
    return (x < P.width && y < P.height) ? y * P.width + x : -1;
}

// The path-tracing megakernel.  Lane state: one pixel (gid), its remaining samples, the RNG state
// and running mean; path state: ray, radiance, throughput, specularBounce, depth.
template <int STACK, bool STATS>
__global__ __launch_bounds__(256) void k_render(RenderParams P) {
    if (recorded_tree_unfit(P)) return;
    __shared__ int s_stack[STACK * 256];
    int *stk = s_stack + threadIdx.x;
    const DevAccel &A = P.acc;
    const unsigned long long lt = lanemask_lt();

    bool active = false, drained = false;
    int gid = 0, ci = 0, k = 0;
    unsigned s0 = 0, s1 = 0;
    f3 col = mk(0.f, 0.f, 0.f);
    f3 ro = mk(0.f, 0.f, 0.f), rd = mk(0.f, 0.f, 1.f), rad = mk(0.f, 0.f, 0.f), thr = mk(1.f, 1.f, 1.f);
    int spec = 1, depth = 0;
    unsigned long long n_closest = 0, n_any = 0;
    unsigned st_nodes = 0, st_prims = 0;
    unsigned *const pnv = STATS ? &st_nodes : nullptr;
    unsigned *const ppt = STATS ? &st_prims : nullptr;

    for (;;) {
        if (!drained) {
            const unsigned long long need = __ballot(!active);
            if (need) {
                const unsigned cnt = (unsigned)__popcll(need);
                const int leader = __ffsll((long long)need) - 1;
                unsigned base = 0;
                if ((int)__lane_id() == leader) base = atomicAdd(P.work_counter, cnt);
                base = __shfl(base, leader, kWave);
                if (base + cnt >= (unsigned)P.total) drained = true;
                if (!active) {
                    const unsigned my = base + (unsigned)__popcll(need & lt);
                    const int g = my < (unsigned)P.total ? claimed_pixel(P, my) : -1;
                    if (g >= 0) {   // (a tile pixel past the frame's edge: the lane claims again)
                        active = true;
                        gid = g;
                        const int x = gid % P.width, y = gid / P.width;
                        ci = (P.height - y - 1) * P.width + x;
                        s0 = P.seeds[2 * gid];
                        s1 = P.seeds[2 * gid + 1];
                        k = 0;
                        if (P.currentSample > 0) col = ldvec(P.colors + ci);
                        camera_ray(P.camera, s0, s1, P.width, P.height, x, y, ro, rd);
                        rad = mk(0.f, 0.f, 0.f); thr = mk(1.f, 1.f, 1.f); spec = 1; depth = 0;
                    }
                }
            }
        }
        if (__ballot(active) == 0) break;
        if (!active) continue;

        // ---- one path segment (RadianceOnePathTracing, rendering_kernel.cl:495-753) ----
        bool ended = false;
        if (depth > P.max_depth) {
            ended = true;
        } else {
            n_closest++;
            float t;
            const int id = closest<STACK>(A, ro, rd, &t, stk, pnv, ppt);
            if (id < 0) {
                ended = true;
            } else {
                const f3 hp = add3(ro, smul(t, rd));
                const Surface S = surface_at(P.shapes + id, hp, rd);
                if (!iszero3(S.e)) {
                    if (spec) rad = add3(rad, mul3(thr, smul(fabsf(S.dp), S.e)));
                    ended = true;
                } else if (S.refl == DIFF) {
                    spec = 0;
                    thr = mul3(thr, S.c);
                    f3 Ld = mk(0.f, 0.f, 0.f);
                    for (int li = 0; li < P.nlights; li++) {
                        f3 sd, contrib;
                        float len;
                        if (!sample_light(P.shapes + P.lights[li], s0, s1, hp, S.nl, &sd, &len, &contrib)) continue;
                        n_any++;
                        if (!anyhit<STACK>(A, hp, sd, len - kEpsilon, stk, pnv, ppt)) Ld = add3(Ld, contrib);
                    }
                    Ld = mul3(thr, Ld);
                    rad = add3(rad, Ld);
                    rd = diffuse_dir(S.nl, s0, s1);
                    ro = hp;
                } else {
                    spec = 1;
                    specular_or_refract(S, rd, thr, s0, s1);
                    ro = hp;
                }
                depth++;
            }
        }
        if (ended) {
            col = running_mean(col, rad, P.currentSample + k);
            k++;
            if (k < P.spp) {
                const int x = gid % P.width, y = gid / P.width;
                camera_ray(P.camera, s0, s1, P.width, P.height, x, y, ro, rd);
                rad = mk(0.f, 0.f, 0.f); thr = mk(1.f, 1.f, 1.f); spec = 1; depth = 0;
            } else {
                float *cp = reinterpret_cast<float *>(P.colors + ci);
                cp[0] = col.x; cp[1] = col.y; cp[2] = col.z;
                if (P.pixels) {
                    const int x = gid % P.width, y = gid / P.width;
                    P.pixels[y * P.width + x] = pack_rgba(col);
                }
                P.seeds[2 * gid] = s0;
                P.seeds[2 * gid + 1] = s1;
                active = false;
            }
        }
    }
    if (P.ray_counter) {
        // wave reduction, one atomic per wave
        unsigned long long a = n_closest, b = n_any;
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_down(a, o, kWave);
            b += __shfl_down(b, o, kWave);
        }
        if (__lane_id() == 0) {
            atomicAdd(P.ray_counter, a);
            atomicAdd(P.ray_counter + 1, b);
        }
    }
    if (STATS && P.trav_stats) {
        unsigned long long a = st_nodes, b = st_prims;
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_down(a, o, kWave);
            b += __shfl_down(b, o, kWave);
        }
        if (__lane_id() == 0) {
            atomicAdd(P.trav_stats, a);
            atomicAdd(P.trav_stats + 1, b);
        }
    }
}

// ---- megakernel with per-lane states ----
//
// The same paths as k_render, scheduled differently: every lane carries its own query (closest or
// any hit) and the wave advances all traversing lanes by one node per loop iteration.  A lane whose
// query ends does not hold the wave: it waits in a ready state until `regroup` lanes are ready (or
// none is traversing), then the wave runs one shading pass for exactly those lanes, each of which
// leaves it with its next query (or its pixel done).  The random numbers are drawn in the same
// order as in k_render (shadow-ray outcomes never draw), so the bits are the same.  Needs a tree
// whose root is an internal node (brute force and single-leaf trees run in k_render).
enum : int {
    kNeed = 0,    // wants a pixel
    kTrav = 1,    // traversing
    kCDone = 2,   // closest-hit query finished: tb / bid hold the result
    kADone = 3,   // any-hit query finished: bid >= 0 means occluded
    kLight = 4,   // next light of SampleLights (index li)
    kSeg = 5,     // start of a path segment (depth checked here)
    kEnded = 6,   // sample finished
    kIdle = 7,    // no work left
};

template <int STACK, bool STATS, bool WIDE, bool MULTI, bool HALF>
__global__ __launch_bounds__(256, STACK == kStackShort || STACK == kStackTiny ? 4 : 1) void k_render_lanes(RenderParams P) {
    if (!MULTI && recorded_tree_unfit(P)) return;
    if (MULTI && P.status && recorded_jobs_unfit(P)) return;
    __shared__ int s_stack[STACK * 256];
    int *stk = s_stack + threadIdx.x;
    // The short class (binary16 walk only) holds fewer entries than the tree's height could need: a
    // push that does not fit sets kOverflow in the lane's visit count, which ends the walk at its next
    // step, and the query is then finished against every primitive (see kTrav below).
    constexpr bool kShort = STACK == kStackShort || STACK == kStackTiny;
    static_assert(!kShort || (WIDE && HALF), "the short stack class serves the binary16 walk");
    const unsigned long long lt = lanemask_lt();
    // A single frame's first kLdsLights lights, 16 words each (words 0-12 of the light's Shape, then
    // its area), copied once per workgroup: SampleLights then reads LDS instead of two dependent
    // global loads (the light's index, then its Shape) per light in every shading pass.
    // The camera likewise: every new sample reads it.
    constexpr int kLdsLights = 32;
    __shared__ float s_lights[MULTI ? 16 : kLdsLights * 16];
    __shared__ Camera s_cam;
    int n_lds_lights = 0;
    if (!MULTI) {
        n_lds_lights = P.nlights < kLdsLights ? P.nlights : kLdsLights;
        for (int i = threadIdx.x; i < n_lds_lights * 16; i += blockDim.x) {
            const int l = i >> 4, w = i & 15;
            const float *sw = reinterpret_cast<const float *>(P.shapes + P.lights[l]);
            s_lights[i] = w <= 12 ? sw[w] : (w == 13 ? sw[25] : 0.f);
        }
        constexpr int kCamWords = (int)(sizeof(Camera) / sizeof(float));
        if ((int)threadIdx.x < kCamWords)
            reinterpret_cast<float *>(&s_cam)[threadIdx.x] = reinterpret_cast<const float *>(P.camera)[threadIdx.x];
        __syncthreads();
    }
    // A batch's first pixel per job (prefix sums of the jobs' pixel counts) for up to kLdsJobs jobs:
    // the search for a claimed pixel's job then reads LDS instead of one dependent global load per
    // halving.
    constexpr int kLdsJobs = 256;
    __shared__ unsigned s_job_begin[MULTI ? kLdsJobs : 1];
    if (MULTI) {
        const int nj = P.njobs < kLdsJobs ? P.njobs : kLdsJobs;
        for (int i = threadIdx.x; i < nj; i += blockDim.x) s_job_begin[i] = P.job_begin[i];
        __syncthreads();
    }
    // frame fields: the launch's own, or (batch) the lane's job's
#define JF(field) to_gmem(MULTI ? P.jobs[jb].field : P.field)
    int jb = 0;
    // batch: the lane's job's fields that every shading pass reads, kept in registers from the pixel
    // claim on (a field read through the job table is one more dependent load)
    const gmem<const Shape> *j_shapes = to_gmem(P.shapes);
    const gmem<const int> *j_lights = to_gmem(P.lights);
    int j_nlights = P.nlights, j_max_depth = P.max_depth;
    // and those every new sample reads
    const gmem<const Camera> *j_camera = to_gmem(P.camera);
    int j_width = P.width, j_height = P.height, j_spp = P.spp, j_currentSample = P.currentSample;
#define JSHAPES (MULTI ? from_gmem(j_shapes) : P.shapes)
#define JFR(field) (MULTI ? j_##field : P.field)
    // the tree of the lane's current query (uniform for a single frame)
    const gmem<const TNode> *t_nodes = to_gmem(P.acc.nodes);
    const gmem<const TNode4> *t_nodes4 = to_gmem(P.acc.nodes4);
    const gmem<const TNode4H> *t_nodes4h = to_gmem(P.acc.nodes4h);
    const gmem<const Prim> *t_prims = to_gmem(P.acc.prims);
    int t_nnodes = P.acc.nnodes, t_nprims = P.acc.nprims, t_nshapes = P.acc.nshapes, t_root = P.acc.root;

    int st = kNeed;
    bool drained = false;
    int gid = 0, k = 0, depth = 0, li = 0;
    bool spec = true, qany = false;
    unsigned s0 = 0, s1 = 0;
    f3 col = mk(0.f, 0.f, 0.f), ro = col, rd = mk(0.f, 0.f, 1.f), rad = col, thr = mk(1.f, 1.f, 1.f);
    f3 nl = col, Ld = col, contrib = col, inv = col;
    RaySlab rsl = ray_slab(mk(1.f, 1.f, 1.f));
    float tb = kInf;
    int bid = -1, node = 0, sp = 0;
    int pl = -1;   // binary16 walk: first primitive of a leaf run waiting to be tested (-1 none)
    unsigned nv = 0;
    unsigned long long n_closest = 0, n_any = 0;
    unsigned st_nodes = 0, st_prims = 0;

    // the primitives of one leaf for the lane's current query; true when an any-hit query is decided
    auto leaf = [&](int first, int cnt) -> bool {
        if (first < 0 || cnt > t_nprims - first) return false;   // malformed caller tree
        for (int q = 0; q < cnt; q++) {
            int id;
            const float t = prim_isect(from_gmem(t_prims), first + q, ro, rd, &id);
            if (STATS) st_prims++;
            if (qany) {
                if (t != 0.f && t < tb && id != t_nshapes - 1) { bid = 1; return true; }
            } else if (t != 0.f && (t < tb || (t == tb && id > bid))) {
                tb = t;
                bid = id;
            }
        }
        return false;
    };
    // the primitive run of one leaf (binary16 nodes carry no counts: a run ends at kLastInLeaf)
    auto leaf_run = [&](int first) -> bool {
        if (first < 0 || first >= t_nprims) return false;   // malformed caller tree
        for (int q = first; q < t_nprims; q++) {
            int id;
            bool last;
            const float t = prim_isect_run(from_gmem(t_prims), q, ro, rd, &id, &last);
            if (STATS) st_prims++;
            if (qany) {
                if (t != 0.f && t < tb && id != t_nshapes - 1) { bid = 1; return true; }
            } else if (t != 0.f && (t < tb || (t == tb && id > bid))) {
                tb = t;
                bid = id;
            }
            if (last) break;
        }
        return false;
    };
    // A walk that outgrew the short stack is finished against every primitive, from the best hit found
    // so far.  The closest hit is the least (t, then the higher id) whatever the order of the tests,
    // and an occluder ends an any-hit query, so the result is the walk's.
    auto finish_overflowed = [&]() {
        for (int q = 0; q < t_nprims && !(qany && bid >= 0); q++) {
            const Prim *tq = from_gmem(t_prims) + q;
            int id;
            bool last;
            const float t = prim_isect_vals(tq->p0, tq->p1, tq->p2, ro, rd, &id, &last);
            if (STATS) st_prims++;
            if (qany) {
                if (t != 0.f && t < tb && id != t_nshapes - 1) bid = 1;
            } else if (t != 0.f && (t < tb || (t == tb && id > bid))) {
                tb = t;
                bid = id;
            }
        }
    };
    // start the lane's query along (ro, rd); tb = maxt for any-hit
    auto start = [&](bool any, float maxt) {
        qany = any;
        tb = any ? maxt : kInf;
        bid = -1;
        inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
        if (HALF) rsl = ray_slab(inv);
        node = t_root;
        pl = -1;
        sp = 0;
        nv = 0;
        st = kTrav;
    };

    for (;;) {
        // ---- traversal: one node of every traversing lane per iteration, until `regroup` lanes
        // (or all that are left) have finished their queries ----
        const unsigned long long live = __ballot(st != kIdle);
        if (!live) break;
        // a sparse wave (the end of the frame) shades sooner, so no lane's path waits long on others
        const unsigned nlive = (unsigned)__popcll(live);
        const unsigned quarter = (nlive + 3u) >> 2;
        const unsigned regroup = (unsigned)P.regroup < quarter ? (unsigned)P.regroup : quarter;
        for (;;) {
            const unsigned long long trav = __ballot(st == kTrav);
            if (!trav || (unsigned)__popcll(live & ~trav) >= regroup) break;
            if (HALF && (unsigned)__popcll(trav) <= (unsigned)P.quad) {
                // ---- quad phase: a sparse wave lends its idle lanes to the queries in flight.  The
                // k-th traversing lane's query is walked by lanes 4k..4k+3, lane c of the quad taking
                // child c of each node: one 16-byte load, one box test and (for a leaf child) one
                // primitive run per lane, then the quad agrees on the hit, the next node and the
                // pushes through DPP.  Same nodes, same hits, shorter steps. ----
                const int qk = (int)(__lane_id() >> 2), qc = (int)(__lane_id() & 3);
                const int nq = __popcll(trav);
                int owner = 0;
                {
                    unsigned long long m = trav;
                    for (int j = 0; j < nq; j++) {
                        const int b = __ffsll((long long)m) - 1;
                        if (j == qk) owner = b;
                        m &= m - 1;
                    }
                }
                const bool inq = qk < nq;
                // the owner's query (every lane runs the shuffles; lanes beyond nq read lane 0)
                f3 q_o, q_d, q_inv;
                q_o.x = __shfl(ro.x, owner); q_o.y = __shfl(ro.y, owner); q_o.z = __shfl(ro.z, owner);
                q_d.x = __shfl(rd.x, owner); q_d.y = __shfl(rd.y, owner); q_d.z = __shfl(rd.z, owner);
                q_inv.x = __shfl(inv.x, owner); q_inv.y = __shfl(inv.y, owner); q_inv.z = __shfl(inv.z, owner);
                const RaySlab q_rs = ray_slab(q_inv);
                float q_tb = __shfl(tb, owner);
                int q_bid = __shfl(bid, owner), q_node = __shfl(node, owner), q_sp = __shfl(sp, owner);
                int q_pl = __shfl(pl, owner);
                unsigned q_nv = (unsigned)__shfl((int)nv, owner);
                const bool q_any = __shfl((int)qany, owner) != 0;
                const gmem<const TNode4H> *q_nodes = t_nodes4h;
                const gmem<const Prim> *q_prims = t_prims;
                int q_nnodes = t_nnodes, q_nprims = t_nprims, q_nshapes = t_nshapes;
                if (MULTI) {
                    auto shp = [&](const void *ptr) -> const void * {
                        const unsigned long long v = reinterpret_cast<unsigned long long>(ptr);
                        const unsigned lo = (unsigned)__shfl((int)(unsigned)v, owner);
                        const unsigned hi = (unsigned)__shfl((int)(unsigned)(v >> 32), owner);
                        return reinterpret_cast<const void *>(((unsigned long long)hi << 32) | lo);
                    };
                    q_nodes = to_gmem(static_cast<const TNode4H *>(shp(from_gmem(t_nodes4h))));
                    q_prims = to_gmem(static_cast<const Prim *>(shp(from_gmem(t_prims))));
                    q_nnodes = __shfl(t_nnodes, owner);
                    q_nprims = __shfl(t_nprims, owner);
                    q_nshapes = __shfl(t_nshapes, owner);
                }
                int *q_stk = s_stack + (threadIdx.x & ~63u) + owner;   // the owner's stack column
                bool q_done = !inq;
                bool q_end = false;   // the query finished (as opposed to the phase ending)
                constexpr float inf = __builtin_inff();
                const unsigned ready0 = (unsigned)__popcll(live & ~trav);
                for (;;) {
                    const unsigned long long fin = __ballot(inq && q_end && qc == 0);
                    if (ready0 + (unsigned)__popcll(fin) >= regroup || (unsigned)__popcll(fin) >= (unsigned)nq) break;
                    if (!q_done) {
                        const bool qhn = q_node >= 0;
                        if ((qhn && q_node >= q_nnodes) || ++q_nv > kMaxVisits) {
                            q_done = q_end = true;   // malformed caller tree
                        } else {
                            // On lane 0 the owner's pending run, then on every lane its leaf child, against
                            // the step's tb.  The run's first primitive is loaded before the child (every
                            // lane issues both loads; the others read primitive 0 and ignore it), so its
                            // test overlaps the child's latency.
                            const bool qhl = qc == 0 && q_pl >= 0 && q_pl < q_nprims;
                            const Prim *qpp = from_gmem(q_prims) + (qhl ? q_pl : 0);
                            const float4 m0 = qpp->p0, m1 = qpp->p1, m2 = qpp->p2;
                            asm volatile("" ::: "memory");   // compiler-only: keeps the order above
                            const uint4 w = from_gmem(q_nodes)[qhn ? q_node : 0].c[qc];   // ignored with no node
                            float bt = q_tb;
                            int bb = q_bid;
                            bool occl = false;
                            auto take = [&](float t, int id) -> bool {   // true: occluded, stop
                                if (STATS) st_prims++;
                                if (q_any) {
                                    if (t != 0.f && t < q_tb && id != q_nshapes - 1) { occl = true; return true; }
                                } else if (t != 0.f && (t < bt || (t == bt && id > bb))) {
                                    bt = t;
                                    bb = id;
                                }
                                return false;
                            };
                            if (qhl) {
                                int id;
                                bool last;
                                const float t = prim_isect_vals(m0, m1, m2, q_o, q_d, &id, &last);
                                if (!take(t, id) && !last) {
                                    for (int q = q_pl + 1; q < q_nprims; q++) {
                                        const float t2 = prim_isect_run(from_gmem(q_prims), q, q_o, q_d, &id, &last);
                                        if (take(t2, id) || last) break;
                                    }
                                }
                            }
                            const int ref = qhn ? (int)w.w : kEmptyRef;
                            float tc = slab_hn(w.x, w.y, w.z, q_o, q_inv, q_rs, q_tb * kSlack, P.one);
                            if (!qhn) tc = inf;
                            if (STATS) st_nodes += ref != kEmptyRef;
                            if (!occl && ref < 0 && tc != inf && ~ref < q_nprims) {
                                for (int q = ~ref; q < q_nprims; q++) {
                                    int id;
                                    bool last;
                                    const float t = prim_isect_run(from_gmem(q_prims), q, q_o, q_d, &id, &last);
                                    if (take(t, id) || last) break;
                                }
                            }
                            q_pl = -1;
                            if (ref < 0) tc = inf;   // only internal children from here on
                            // the quad's hit: occlusion by any lane; closest by (t, then higher id)
                            occl = (qxor<1>((int)occl) | qxor<2>((int)occl) | qxor<1>(qxor<2>((int)occl)) |
                                    (int)occl) != 0;
                            {
                                float ot = qxorf<1>(bt);
                                int oi = qxor<1>(bb);
                                if (ot < bt || (ot == bt && oi > bb)) { bt = ot; bb = oi; }
                                ot = qxorf<2>(bt);
                                oi = qxor<2>(bb);
                                if (ot < bt || (ot == bt && oi > bb)) { bt = ot; bb = oi; }
                            }
                            if (q_any) {
                                if (occl) {
                                    q_bid = 1;
                                    q_done = q_end = true;
                                }
                            } else {
                                q_tb = bt;
                                q_bid = bb;
                            }
                            if (!q_done) {
                                // internal children: nearest next, the rest pushed farthest first; order
                                // by (t, slot), which every lane of the quad evaluates alike
                                const float t1 = qrotf<1>(tc), t2 = qrotf<2>(tc), t3 = qrotf<3>(tc);
                                const int r1 = qrot<1>(ref), r2 = qrot<2>(ref), r3 = qrot<3>(ref);
                                const int s1 = (qc + 1) & 3, s2 = (qc + 2) & 3, s3 = (qc + 3) & 3;
                                const int nh = (tc != inf) + (t1 != inf) + (t2 != inf) + (t3 != inf);
                                auto before = [](float ta, int sa, float tb_, int sb) {
                                    return ta < tb_ || (ta == tb_ && sa < sb);
                                };
                                const int rank = before(t1, s1, tc, qc) + before(t2, s2, tc, qc) + before(t3, s3, tc, qc);
                                float nt = tc;
                                int nr = ref, ns = qc;
                                if (before(t1, s1, nt, ns)) { nt = t1; nr = r1; ns = s1; }
                                if (before(t2, s2, nt, ns)) { nt = t2; nr = r2; ns = s2; }
                                if (before(t3, s3, nt, ns)) { nt = t3; nr = r3; ns = s3; }
                                if (nh > 0) {
                                    if (tc != inf && rank > 0) {
                                        const int pos = q_sp + nh - 1 - rank;
                                        if (pos < STACK) q_stk[pos * 256] = ref;
                                    }
                                    q_sp += nh - 1;
                                    // a push past the stack (the same test on every lane of the quad)
                                    if (kShort && q_sp > STACK) q_nv |= kOverflow;
                                    if (q_sp > STACK) q_sp = STACK;
                                    q_node = nr;
                                } else if (q_sp == 0) {
                                    q_done = q_end = true;
                                } else {
                                    const int r = q_stk[(--q_sp) * 256];   // a node, or a leaf to test next
                                    q_node = r >= 0 ? r : -1;
                                    q_pl = r >= 0 ? -1 : ~r;
                                }
                            }
                        }
                    }
                }
                // hand each query back to its owner: the owner of quad k reads lane 4k
                const int src = 4 * (int)__popcll(trav & lt);
                const int b_node = __shfl(q_node, src), b_sp = __shfl(q_sp, src), b_bid = __shfl(q_bid, src);
                const float b_tb = __shfl(q_tb, src);
                const unsigned b_nv = (unsigned)__shfl((int)q_nv, src);
                const bool b_end = __shfl((int)q_end, src) != 0;
                const int b_pl = __shfl(q_pl, src);
                if (st == kTrav) {
                    node = b_node;
                    pl = b_pl;
                    sp = b_sp;
                    bid = b_bid;
                    tb = b_tb;
                    nv = b_nv;
                    if (b_end) {
                        // a walk that overflowed ends at its next quad step and is finished here
                        if (kShort && (b_nv & kOverflow)) finish_overflowed();
                        st = qany ? kADone : kCDone;
                    }
                }
                continue;
            }
            if (!WIDE && st == kTrav) {
                bool done = false;
                if ((unsigned)node >= (unsigned)t_nnodes || ++nv > kMaxVisits) {
                    done = true;   // malformed caller tree
                } else {
                    const TNode *np = from_gmem(t_nodes) + node;
                    const float4 a = np->a, b = np->b, c = np->c;
                    const int4 r = np->r;
                    if (STATS) st_nodes += 2;
                    float t0 = slab(a.x, a.y, a.z, a.w, b.x, b.y, ro, inv, tb);
                    float t1 = slab(b.z, b.w, c.x, c.y, c.z, c.w, ro, inv, tb);
                    if (r.x < 0 && t0 != __builtin_inff()) {
                        done = leaf(~r.x, r.z);
                        t0 = __builtin_inff();
                    }
                    if (!done && r.y < 0 && t1 != __builtin_inff()) {
                        done = leaf(~r.y, r.w);
                        t1 = __builtin_inff();
                    }
                    if (!done) {
                        const bool h0 = t0 != __builtin_inff(), h1 = t1 != __builtin_inff();
                        if (h0 && h1) {
                            const bool first0 = qany || t0 <= t1;
                            if (sp < STACK) stk[(sp++) * 256] = first0 ? r.y : r.x;
                            node = first0 ? r.x : r.y;
                        } else if (h0) {
                            node = r.x;
                        } else if (h1) {
                            node = r.y;
                        } else if (sp == 0) {
                            done = true;
                        } else {
                            node = stk[(--sp) * 256];
                        }
                    }
                }
                if (done) st = qany ? kADone : kCDone;
            }
            if (HALF && st == kTrav) {
                // One step = the current node (if any) and the pending leaf run (if any), loaded
                // together: a hit leaf is not tested in the step that finds it but in the next one,
                // beside the next node's loads, so a step costs one memory round trip, not two.
                bool done = false;
                constexpr float inf = __builtin_inff();
                const bool hn = node >= 0;
                if ((hn && node >= t_nnodes) || ++nv > kMaxVisits) {
                    done = true;   // malformed caller tree
                } else {
                    // Every lane issues both loads, the run's primitive first: a lane with no run reads
                    // primitive 0 and one with no node reads node 0 (both exist in a tree with binary16
                    // nodes), and ignores them.  With no branch around the loads the wait for the
                    // primitive leaves the node's four loads in flight, so the triangle test overlaps
                    // the node's latency.
                    const bool hl = pl >= 0 && pl < t_nprims;
                    const Prim *pp = from_gmem(t_prims) + (hl ? pl : 0);
                    float4 a0 = pp->p0, a1 = pp->p1, a2 = pp->p2;
                    // compiler-only barrier (no operands, no instruction): the primitive's loads stay
                    // here, ahead of the node's, instead of sinking into the branch that tests them
                    asm volatile("" ::: "memory");
                    const TNode4H *np = from_gmem(t_nodes4h) + (hn ? node : 0);
                    const uint4 q0 = np->c[0], q1 = np->c[1], q2 = np->c[2], q3 = np->c[3];
                    // the pending run (one primitive per LBVH leaf), before the boxes so tb is tight
                    if (hl) {
                        // true: the run ends here (an occluder found, or its last primitive)
                        auto test_prim = [&](const float4 &b0, const float4 &b1, const float4 &b2, int q) -> bool {
                            const int id = __float_as_int(b0.w);
                            const float t = (__float_as_int(b2.w) & 0xff) == SPHERE ? sph_isect(b0, b1.x, ro, rd)
                                                                                     : tri_isect(b0, b1, b2, ro, rd);
                            if (STATS) st_prims++;
                            if (qany) {
                                if (t != 0.f && t < tb && id != t_nshapes - 1) { bid = 1; done = true; return true; }
                            } else if (t != 0.f && (t < tb || (t == tb && id > bid))) {
                                tb = t;
                                bid = id;
                            }
                            return (__float_as_int(b1.w) & kLastInLeaf) || q + 1 >= t_nprims;
                        };
                        // the first primitive outside the loop: the loop's own loads issue in another
                        // order, and a shared loop head would make this test wait for the node too
                        if (!test_prim(a0, a1, a2, pl)) {
                            for (int q = pl + 1;; q++) {
                                const Prim *tq = from_gmem(t_prims) + q;
                                const float4 b0 = tq->p0, b1 = tq->p1, b2 = tq->p2;
                                if (test_prim(b0, b1, b2, q)) break;
                            }
                        }
                    }
                    pl = -1;
                    int nn = -1;
                    if (!done && hn) {
                        int r0 = (int)q0.w, r1 = (int)q1.w, r2 = (int)q2.w, r3 = (int)q3.w;
                        const float tbs = tb * kSlack;
                        float t0 = slab_hn(q0.x, q0.y, q0.z, ro, inv, rsl, tbs, P.one);
                        float t1 = slab_hn(q1.x, q1.y, q1.z, ro, inv, rsl, tbs, P.one);
                        float t2 = slab_hn(q2.x, q2.y, q2.z, ro, inv, rsl, tbs, P.one);
                        float t3 = slab_hn(q3.x, q3.y, q3.z, ro, inv, rsl, tbs, P.one);
                        if (STATS) st_nodes += (r0 != kEmptyRef) + (r1 != kEmptyRef) + (r2 != kEmptyRef) + (r3 != kEmptyRef);
                        auto cswap = [](float &ta, int &ra, float &tb_, int &rb) {
                            const bool sw = tb_ < ta;
                            const float tt = sw ? tb_ : ta;
                            tb_ = sw ? ta : tb_;
                            ta = tt;
                            const int rr = sw ? rb : ra;
                            rb = sw ? ra : rb;
                            ra = rr;
                        };
                        cswap(t0, r0, t1, r1);
                        cswap(t2, r2, t3, r3);
                        cswap(t0, r0, t2, r2);
                        cswap(t1, r1, t3, r3);
                        cswap(t1, r1, t2, r2);
                        // nearest internal child next, nearest leaf pending, the rest pushed farthest first
                        bool c0 = false, c1 = false, c2 = false, c3 = false;
                        auto choose = [&](float t, int r, bool &c) {
                            if (t == inf) return;
                            if (r >= 0) {
                                if (nn < 0) { nn = r; c = true; }
                            } else if (pl < 0) {
                                pl = ~r;
                                c = true;
                            }
                        };
                        choose(t0, r0, c0);
                        choose(t1, r1, c1);
                        choose(t2, r2, c2);
                        choose(t3, r3, c3);
                        auto push = [&](float t, bool c, int r) {
                            if (t == inf || c) return;
                            if (sp < STACK) stk[(sp++) * 256] = r;
                            else if (kShort) nv |= kOverflow;
                        };
                        push(t3, c3, r3);
                        push(t2, c2, r2);
                        push(t1, c1, r1);
                        push(t0, c0, r0);
                    }
                    node = nn;
                    if (!done) {
                        // refill from the stack: a node for the next step, and a leaf to load beside it.
                        // The top two entries are read together (one LDS round trip, not two).
                        const int e1 = sp > 0 ? stk[(sp - 1) * 256] : 0;
                        const int e2 = sp > 1 ? stk[(sp - 2) * 256] : 0;
                        int top = e1;
                        if (node < 0 && sp > 0 && e1 >= 0) {
                            node = e1;
                            --sp;
                            top = e2;
                        }
                        if (pl < 0 && sp > 0 && top < 0) {
                            pl = ~top;
                            --sp;
                        }
                        if (node < 0 && pl < 0) done = true;   // the stack is empty
                    }
                }
                if (kShort && done && (nv & kOverflow)) finish_overflowed();
                if (done) st = qany ? kADone : kCDone;
            }
            if (WIDE && !HALF && st == kTrav) {
                bool done = false;
                if ((unsigned)node >= (unsigned)t_nnodes || ++nv > kMaxVisits) {
                    done = true;   // malformed caller tree
                } else {
                    float4 lx, ly, lz, hx, hy, hz;
                    int4 rf, cn;
                    if (HALF) {
                        const TNode4H *np = from_gmem(t_nodes4h) + node;
                        const uint4 q0 = np->c[0], q1 = np->c[1], q2 = np->c[2], q3 = np->c[3];
                        rf = make_int4((int)q0.w, (int)q1.w, (int)q2.w, (int)q3.w);
                        cn = make_int4(0, 0, 0, 0);
                        lx = make_float4(half_lo(q0.x), half_lo(q1.x), half_lo(q2.x), half_lo(q3.x));
                        hx = make_float4(half_hi(q0.x), half_hi(q1.x), half_hi(q2.x), half_hi(q3.x));
                        ly = make_float4(half_lo(q0.y), half_lo(q1.y), half_lo(q2.y), half_lo(q3.y));
                        hy = make_float4(half_hi(q0.y), half_hi(q1.y), half_hi(q2.y), half_hi(q3.y));
                        lz = make_float4(half_lo(q0.z), half_lo(q1.z), half_lo(q2.z), half_lo(q3.z));
                        hz = make_float4(half_hi(q0.z), half_hi(q1.z), half_hi(q2.z), half_hi(q3.z));
                    } else {
                        const TNode4 *np = from_gmem(t_nodes4) + node;
                        lx = np->lox; ly = np->loy; lz = np->loz;
                        hx = np->hix; hy = np->hiy; hz = np->hiz;
                        rf = np->ref;
                        cn = np->cnt;
                    }
                    constexpr float inf = __builtin_inff();
                    // branch-free: unused slots hold a box beyond every query's clip distance
                    float t0 = slab(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, ro, inv, tb);
                    float t1 = slab(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, ro, inv, tb);
                    float t2 = slab(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, ro, inv, tb);
                    float t3 = slab(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, ro, inv, tb);
                    if (STATS) st_nodes += (rf.x != kEmptyRef) + (rf.y != kEmptyRef) + (rf.z != kEmptyRef) +
                                           (rf.w != kEmptyRef);
                    // leaf children that are hit: their primitives now, one leaf per loop turn
                    unsigned lm = (rf.x < 0 && t0 != inf ? 1u : 0u) | (rf.y < 0 && t1 != inf ? 2u : 0u) |
                                  (rf.z < 0 && t2 != inf ? 4u : 0u) | (rf.w < 0 && t3 != inf ? 8u : 0u);
                    if (lm & 1u) t0 = inf;
                    if (lm & 2u) t1 = inf;
                    if (lm & 4u) t2 = inf;
                    if (lm & 8u) t3 = inf;
                    while (lm && !done) {
                        const unsigned k = (unsigned)__builtin_ctz(lm);
                        lm &= lm - 1u;
                        const int lr = k == 0 ? rf.x : k == 1 ? rf.y : k == 2 ? rf.z : rf.w;
                        if (HALF) {
                            done = leaf_run(~lr);
                        } else {
                            const int lc = k == 0 ? cn.x : k == 1 ? cn.y : k == 2 ? cn.z : cn.w;
                            done = leaf(~lr, lc);
                        }
                    }
                    if (!done) {
                        // inner children: nearest first (order only changes the visit count)
                        int r0 = rf.x, r1 = rf.y, r2 = rf.z, r3 = rf.w;
                        auto cswap = [](float &ta, int &ra, float &tb_, int &rb) {
                            const bool sw = tb_ < ta;
                            const float tt = sw ? tb_ : ta;
                            tb_ = sw ? ta : tb_;
                            ta = tt;
                            const int rr = sw ? rb : ra;
                            rb = sw ? ra : rb;
                            ra = rr;
                        };
                        cswap(t0, r0, t1, r1);
                        cswap(t2, r2, t3, r3);
                        cswap(t0, r0, t2, r2);
                        cswap(t1, r1, t3, r3);
                        cswap(t1, r1, t2, r2);
                        if (t3 != inf && sp < STACK) stk[(sp++) * 256] = r3;
                        if (t2 != inf && sp < STACK) stk[(sp++) * 256] = r2;
                        if (t1 != inf && sp < STACK) stk[(sp++) * 256] = r1;
                        if (t0 != inf) node = r0;
                        else if (sp == 0) done = true;
                        else node = stk[(--sp) * 256];
                    }
                }
                if (done) st = qany ? kADone : kCDone;
            }
        }
        // ---- shading pass: pixels for lanes that need one (one atomic per wave) ----
        const unsigned long long need = __ballot(st == kNeed);
        if (need) {
            const unsigned cnt = (unsigned)__popcll(need);
            const int leader = __ffsll((long long)need) - 1;
            unsigned base = 0;
            if ((int)__lane_id() == leader) base = atomicAdd(P.work_counter, cnt);
            base = __shfl(base, leader, kWave);
            if (base + cnt >= (unsigned)P.total) drained = true;
            if (st == kNeed) {
                const unsigned my = base + (unsigned)__popcll(need & lt);
                if (my < (unsigned)P.total) {
                    if (MULTI) {   // the job holding work item `my` (prefix sums of the jobs' pixel counts)
                        // two loops rather than one over a selected pointer, so the LDS search
                        // compiles to LDS reads (a pointer that may be either would be a flat one)
                        int lo = 0, hi = P.njobs - 1;
                        unsigned first;
                        if (P.njobs <= kLdsJobs) {
                            while (lo < hi) {
                                const int mid = (lo + hi + 1) >> 1;
                                if (s_job_begin[mid] <= my) lo = mid; else hi = mid - 1;
                            }
                            first = s_job_begin[lo];
                        } else {
                            const gmem<const unsigned> *jbg = to_gmem(P.job_begin);
                            while (lo < hi) {
                                const int mid = (lo + hi + 1) >> 1;
                                if (jbg[mid] <= my) lo = mid; else hi = mid - 1;
                            }
                            first = jbg[lo];
                        }
                        jb = lo;
                        gid = (int)(my - first);
                        // the job's tree, for every query of this pixel
                        const DevAccel &A = P.jobs[jb].acc;
                        t_nodes = to_gmem(A.nodes);
                        t_nodes4 = to_gmem(A.nodes4);
                        t_nodes4h = to_gmem(A.nodes4h);
                        t_prims = to_gmem(A.prims);
                        t_nnodes = A.nnodes;
                        t_nprims = A.nprims;
                        t_nshapes = A.nshapes;
                        t_root = A.root;
                        j_shapes = to_gmem(P.jobs[jb].shapes);
                        j_lights = to_gmem(P.jobs[jb].lights);
                        j_nlights = P.jobs[jb].nlights;
                        j_max_depth = P.jobs[jb].max_depth;
                        j_camera = to_gmem(P.jobs[jb].camera);
                        j_width = P.jobs[jb].width;
                        j_height = P.jobs[jb].height;
                        j_spp = P.jobs[jb].spp;
                        j_currentSample = P.jobs[jb].currentSample;
                    } else {
                        gid = claimed_pixel(P, my);
                    }
                    if (gid >= 0) {
                        const int W = JFR(width), H = JFR(height);
                        const int x = gid % W, y = gid / W;
                        s0 = JF(seeds)[2 * gid];
                        s1 = JF(seeds)[2 * gid + 1];
                        k = 0;
                        if (JFR(currentSample) > 0) {
                            const gmem<const float> *cv =
                                reinterpret_cast<const gmem<const float> *>(JF(colors) + (H - y - 1) * W + x);
                            col = mk(cv[0], cv[1], cv[2]);
                        }
                        camera_ray(MULTI ? from_gmem(j_camera) : &s_cam, s0, s1, W, H, x, y, ro, rd);
                        rad = mk(0.f, 0.f, 0.f); thr = mk(1.f, 1.f, 1.f); spec = true; depth = 0;
                        st = kSeg;
                    } else {
                        st = drained ? kIdle : kNeed;   // a tile pixel past the frame's edge: claim again
                    }
                } else {
                    st = kIdle;
                }
            }
        }
        // ---- per-lane state machine until the lane holds a query or its pixel is done ----
        while (st >= kCDone && st <= kEnded) {
            if (st == kCDone) {
                if (bid < 0) {
                    st = kEnded;
                } else {
                    const f3 hp = add3(ro, smul(tb, rd));
                    const Surface S = surface_at(JSHAPES + bid, hp, rd);
                    if (!iszero3(S.e)) {
                        if (spec) rad = add3(rad, mul3(thr, smul(fabsf(S.dp), S.e)));
                        st = kEnded;
                    } else if (S.refl == DIFF) {
                        spec = false;
                        thr = mul3(thr, S.c);
                        Ld = mk(0.f, 0.f, 0.f);
                        nl = S.nl;
                        ro = hp;
                        li = 0;
                        st = kLight;
                    } else {
                        spec = true;
                        specular_or_refract(S, rd, thr, s0, s1);
                        ro = hp;
                        depth++;
                        st = kSeg;
                    }
                }
            }
            if (st == kADone) {
                if (bid < 0) Ld = add3(Ld, contrib);
                li++;
                st = kLight;
            }
            if (st == kLight) {
                // lights whose sample needs no shadow ray only draw their random numbers
                float len = 0.f;
                bool shoot = false;
                const int nlights = MULTI ? j_nlights : P.nlights;
                for (; li < nlights; li++) {
                    const bool sh = !MULTI && li < n_lds_lights
                                        ? sample_light_words(s_lights + 16 * li, s_lights[16 * li + 13], s0, s1, ro,
                                                             nl, &rd, &len, &contrib)
                                        : sample_light(JSHAPES + (MULTI ? from_gmem(j_lights) : P.lights)[li], s0, s1, ro, nl, &rd, &len,
                                                       &contrib);
                    if (sh) {
                        shoot = true;
                        break;
                    }
                }
                if (shoot) {
                    n_any++;
                    start(true, len - kEpsilon);
                    continue;
                }
                Ld = mul3(thr, Ld);
                rad = add3(rad, Ld);
                rd = diffuse_dir(nl, s0, s1);
                depth++;
                st = kSeg;
            }
            if (st == kSeg && depth > (MULTI ? j_max_depth : P.max_depth)) st = kEnded;
            if (st == kEnded) {
                col = running_mean(col, rad, JFR(currentSample) + k);
                k++;
                const int W = JFR(width), H = JFR(height);
                const int x = gid % W, y = gid / W;
                if (k < JFR(spp)) {
                    camera_ray(MULTI ? from_gmem(j_camera) : &s_cam, s0, s1, W, H, x, y, ro, rd);
                    rad = mk(0.f, 0.f, 0.f); thr = mk(1.f, 1.f, 1.f); spec = true; depth = 0;
                    st = kSeg;
                } else {
                    gmem<float> *cp = reinterpret_cast<gmem<float> *>(JF(colors) + (H - y - 1) * W + x);
                    cp[0] = col.x; cp[1] = col.y; cp[2] = col.z;
                    gmem<unsigned> *px = JF(pixels);
                    if (px) px[y * W + x] = pack_rgba(col);
                    gmem<unsigned> *sd = JF(seeds);
                    sd[2 * gid] = s0;
                    sd[2 * gid + 1] = s1;
                    st = drained ? kIdle : kNeed;
                }
            }
            if (st == kSeg) {
                n_closest++;
                start(false, 0.f);
            }
        }
    }
    if (P.ray_counter) {
        unsigned long long a = n_closest, b = n_any;
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_down(a, o, kWave);
            b += __shfl_down(b, o, kWave);
        }
        if (__lane_id() == 0) {
            atomicAdd(P.ray_counter, a);
            atomicAdd(P.ray_counter + 1, b);
        }
    }
    if (STATS && P.trav_stats) {
        unsigned long long a = st_nodes, b = st_prims;
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_down(a, o, kWave);
            b += __shfl_down(b, o, kWave);
        }
        if (__lane_id() == 0) {
            atomicAdd(P.trav_stats, a);
            atomicAdd(P.trav_stats + 1, b);
        }
    }
#undef JF
#undef JSHAPES
#undef JFR
}

// Batch queries (Intersect / IntersectP on arbitrary rays).
template <int STACK>
__global__ __launch_bounds__(256) void k_trace(DevAccel A, const Ray *rays, const float *maxt, int n, float *tout,
                                               int *idout, int *hitout, int any, unsigned *stats) {
    __shared__ int s_stack[STACK * 256];
    int *stk = s_stack + threadIdx.x;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *rf = reinterpret_cast<const float *>(rays + i);
    const f3 o = mk(rf[0], rf[1], rf[2]), d = mk(rf[3], rf[4], rf[5]);
    unsigned nv = 0, pt = 0;
    if (any) {
        hitout[i] = anyhit<STACK>(A, o, d, maxt[i], stk, &nv, &pt) ? 1 : 0;
    } else {
        float t;
        const int id = closest<STACK>(A, o, d, &t, stk, &nv, &pt);
        tout[i] = id >= 0 ? t : 0.f;
        idout[i] = id;
    }
    if (stats) {
        atomicAdd(stats, nv);
        atomicAdd(stats + 1, pt);
    }
}

// ---- scene preparation kernels ----
__global__ void k_index_prims(const Shape *shapes, int n, Prim *prims) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) prims[i] = prim_of(shapes + i, i, kLastInLeaf);
}

__device__ __forceinline__ bool emits(const Shape *s) {
    const float *w = reinterpret_cast<const float *>(s);
    return !(w[10] == 0.f && w[11] == 0.f && w[12] == 0.f);
}

constexpr int kLightChunk = 4096;
__global__ __launch_bounds__(256) void k_light_count(const Shape *shapes, int n, unsigned *counts) {
    __shared__ unsigned s_c;
    if (threadIdx.x == 0) s_c = 0;
    __syncthreads();
    unsigned c = 0;
    const int base = blockIdx.x * kLightChunk;
    for (int q = threadIdx.x; q < kLightChunk; q += 256) {
        const int i = base + q;
        if (i < n && emits(shapes + i)) c++;
    }
    if (c) atomicAdd(&s_c, c);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = s_c;
}
__global__ void k_light_scan(unsigned *counts, int nb, unsigned *total) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned run = 0;
    for (int b = 0; b < nb; b++) {
        unsigned c = counts[b];
        counts[b] = run;
        run += c;
    }
    *total = run;
}
__global__ __launch_bounds__(256) void k_light_write(const Shape *shapes, int n, const unsigned *offs, int cap,
                                                     int *out) {
    __shared__ unsigned s_w[4];
    const int base = blockIdx.x * kLightChunk;
    unsigned run = offs[blockIdx.x];
    for (int q0 = 0; q0 < kLightChunk; q0 += 256) {
        const int i = base + q0 + threadIdx.x;
        const unsigned f = (i < n && emits(shapes + i)) ? 1u : 0u;
        // block exclusive scan of f
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        unsigned x = f;
        for (int o = 1; o < 64; o <<= 1) {
            unsigned y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) s_w[w] = x;
        __syncthreads();
        unsigned pre = 0, tot = 0;
        for (int k2 = 0; k2 < 4; k2++) { if (k2 < w) pre += s_w[k2]; tot += s_w[k2]; }
        __syncthreads();
        const unsigned pos = run + pre + x - f;
        if (f && (int)pos < cap) out[pos] = i;
        run += tot;
    }
}

// TNode -> TNode4: the four grandchildren of binary node i (a leaf child is kept as it is).  Refs
// keep the binary numbering, so every internal node has a four-wide twin and a walk from the root
// that always takes two binary levels per step stays inside the array.
__global__ __launch_bounds__(256) void k_collapse4(const TNode *__restrict__ bn, int nnodes, TNode4 *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnodes) return;
    float lo[3][4], hi[3][4];
    int ref[4], cnt[4], k = 0;
    auto put = [&](float x0, float y0, float z0, float x1, float y1, float z1, int r, int c) {
        lo[0][k] = x0; lo[1][k] = y0; lo[2][k] = z0;
        hi[0][k] = x1; hi[1][k] = y1; hi[2][k] = z1;
        ref[k] = r; cnt[k] = c;
        k++;
    };
    auto put_child = [&](const TNode &t, int side) {
        if (side == 0) put(t.a.x, t.a.y, t.a.z, t.a.w, t.b.x, t.b.y, t.r.x, t.r.z);
        else put(t.b.z, t.b.w, t.c.x, t.c.y, t.c.z, t.c.w, t.r.y, t.r.w);
    };
    const TNode t = bn[i];
    for (int side = 0; side < 2; side++) {
        const int r = side == 0 ? t.r.x : t.r.y;
        if (r < 0 || r >= nnodes) {
            put_child(t, side);   // leaf (or a ref the traversal rejects)
        } else {
            const TNode g = bn[r];
            put_child(g, 0);
            put_child(g, 1);
        }
    }
    // an unused slot is a point at 1e30: every query clips at t <= 1e20, so its slab test misses
    for (; k < 4;) put(kFarBox, kFarBox, kFarBox, kFarBox, kFarBox, kFarBox, kEmptyRef, 0);
    TNode4 o;
    o.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
    o.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
    o.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
    o.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
    o.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
    o.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
    o.ref = make_int4(ref[0], ref[1], ref[2], ref[3]);
    o.cnt = make_int4(cnt[0], cnt[1], cnt[2], cnt[3]);
    out[i] = o;
}

// binary16 bounds rounded outward: the nearest binary16 (round to nearest even), stepped one
// binary16 ulp down (lo) or up (hi) when it lies on the wrong side of x.  A NaN bound becomes the
// infinity on its own side; values beyond the binary16 range round to +-65504 or the infinity.
__device__ __forceinline__ unsigned short half_down(float x) {
    if (x != x) return 0xFC00u;   // -inf
    const _Float16 h = (_Float16)x;
    unsigned short b = __builtin_bit_cast(unsigned short, h);
    if ((float)h > x) {
        if (b == 0x0000u || b == 0x8000u) b = 0x8001u;   // +-0 -> the largest negative subnormal
        else if (b & 0x8000u) b = (unsigned short)(b + 1u);   // negative: away from zero
        else b = (unsigned short)(b - 1u);                    // positive: toward zero
    }
    return b;
}
__device__ __forceinline__ unsigned short half_up(float x) {
    if (x != x) return 0x7C00u;   // +inf
    const _Float16 h = (_Float16)x;
    unsigned short b = __builtin_bit_cast(unsigned short, h);
    if ((float)h < x) {
        if (b == 0x0000u || b == 0x8000u) b = 0x0001u;   // +-0 -> the smallest positive subnormal
        else if (b & 0x8000u) b = (unsigned short)(b - 1u);   // negative: toward zero
        else b = (unsigned short)(b + 1u);                    // positive: away from zero
    }
    return b;
}

// TNode4 -> TNode4H (leaf counts dropped: leaves are runs ending in kLastInLeaf)
__global__ __launch_bounds__(256) void k_half4(const TNode4 *__restrict__ in, int nnodes, TNode4H *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnodes) return;
    const TNode4 t = in[i];
    const float lo[3][4] = {{t.lox.x, t.lox.y, t.lox.z, t.lox.w}, {t.loy.x, t.loy.y, t.loy.z, t.loy.w},
                            {t.loz.x, t.loz.y, t.loz.z, t.loz.w}};
    const float hi[3][4] = {{t.hix.x, t.hix.y, t.hix.z, t.hix.w}, {t.hiy.x, t.hiy.y, t.hiy.z, t.hiy.w},
                            {t.hiz.x, t.hiz.y, t.hiz.z, t.hiz.w}};
    const int ref[4] = {t.ref.x, t.ref.y, t.ref.z, t.ref.w};
    unsigned w[3][4];
    for (int a = 0; a < 3; a++)
        for (int k = 0; k < 4; k++)
            w[a][k] = ref[k] == kEmptyRef ? 0x7C007C00u
                                          : (unsigned)half_down(lo[a][k]) | ((unsigned)half_up(hi[a][k]) << 16);
    TNode4H o;
    for (int k = 0; k < 4; k++) o.c[k] = make_uint4(w[0][k], w[1][k], w[2][k], (unsigned)ref[k]);
    out[i] = o;
}

// reference BVH layout -> TNode/Prim (child is a leaf iff nLeft == min / nRight == max)
__global__ void k_convert_bvh(const Shape *shapes, int n, const BVHNodeGPU *btn, const BVHNodeGPU *btl,
                              TNode *nodes, Prim *prims) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) prims[i] = prim_of(shapes + btl[i].nShape, btl[i].nShape, kLastInLeaf);   // one per leaf
    if (i < n - 1) {
        const int nl = btn[i].nLeft, nr = btn[i].nRight;
        const bool lleaf = nl == btn[i].min, rleaf = nr == btn[i].max;
        const Bound L = load_bound(lleaf ? &btl[nl].bound : &btn[nl].bound);
        const Bound R = load_bound(rleaf ? &btl[nr].bound : &btn[nr].bound);
        TNode t;
        t.a = make_float4(L.min_x, L.min_y, L.min_z, L.max_x);
        t.b = make_float4(L.max_y, L.max_z, R.min_x, R.min_y);
        t.c = make_float4(R.min_z, R.max_x, R.max_y, R.max_z);
        t.r = make_int4(lleaf ? ~nl : nl, rleaf ? ~nr : nr, lleaf ? 1 : 0, rleaf ? 1 : 0);
        nodes[i] = t;
    }
}

// Parents of a caller's reference-layout tree, taken from the child links the walk follows (the
// nParent fields are not trusted): par[i] for internal node i, par[n-1+k] for leaf k, -1 where no
// node points.  err[0] collects malformed links: 1 a child index out of range, 2 a node with two
// parents or a link back to the root.
__global__ void k_ref_parents(const BVHNodeGPU *btn, int n, int *par, unsigned *err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int nl = btn[i].nLeft, nr = btn[i].nRight;
    const bool lleaf = nl == btn[i].min, rleaf = nr == btn[i].max;
    if ((unsigned)nl >= (unsigned)(lleaf ? n : n - 1) || (unsigned)nr >= (unsigned)(rleaf ? n : n - 1)) {
        atomicOr(err, 1u);
        return;
    }
    const int sl = lleaf ? n - 1 + nl : nl, sr = rleaf ? n - 1 + nr : nr;
    if (sl == 0 || sr == 0 || sl == sr) atomicOr(err, 2u);
    if (atomicExch(par + sl, i) != -1) atomicOr(err, 2u);
    if (sl != sr && atomicExch(par + sr, i) != -1) atomicOr(err, 2u);
}

// Height of that tree: the longest walk from a leaf up to the root (node 0), in edges, into err[1].
// A leaf no node points at sets err[0] bit 4; a walk longer than kMaxRefHeight (too deep for the
// traversal stack, or a cycle) sets bit 8; a leaf whose nShape is not a shape index sets bit 16.
constexpr int kMaxRefHeight = 62;
__global__ void k_ref_height(const BVHNodeGPU *btl, const int *par, int n, unsigned *err) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if ((unsigned)btl[k].nShape >= (unsigned)n) atomicOr(err, 16u);
    int p = par[n - 1 + k], d = 1;
    while (p > 0 && d <= kMaxRefHeight) {
        p = par[p];
        d++;
    }
    if (p < 0) atomicOr(err, 4u);
    else if (p != 0) atomicOr(err, 8u);
    else atomicMax(err + 1, (unsigned)d);
}

// reference KD heap layout (1-based, children 2i/2i+1, leaves own kn[min..max)) -> TNode/Prim
__global__ void k_convert_kd(const Shape *shapes, const KDNodeGPU *kng, int kngCnt, const int *kn, int knCnt,
                             TNode *nodes, Prim *prims) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h < knCnt) prims[h] = prim_of(shapes + kn[h], kn[h], 0);   // caller ranges may overlap: no flags
    if (h >= 1 && h < kngCnt && kng[h].leaf == 0 && 2 * h + 1 < kngCnt) {
        const int cl = 2 * h, cr = 2 * h + 1;
        const bool lleaf = kng[cl].leaf == 1, rleaf = kng[cr].leaf == 1;
        const Bound L = load_bound(&kng[cl].bound), R = load_bound(&kng[cr].bound);
        TNode t;
        t.a = make_float4(L.min_x, L.min_y, L.min_z, L.max_x);
        t.b = make_float4(L.max_y, L.max_z, R.min_x, R.min_y);
        t.c = make_float4(R.min_z, R.max_x, R.max_y, R.max_z);
        t.r = make_int4(lleaf ? ~(int)kng[cl].min : cl, rleaf ? ~(int)kng[cr].min : cr,
                        lleaf ? (int)kng[cl].max - (int)kng[cl].min : 0,
                        rleaf ? (int)kng[cr].max - (int)kng[cr].min : 0);
        nodes[h] = t;
    }
}

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------
int build_index_prims_into(ptgpu_ctx *ctx, const Shape *d_shapes, int n, Prim *prims) {
    hipLaunchKernelGGL(k_index_prims, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, d_shapes, n, prims);
    PT_HIP(hipGetLastError());
    return PTGPU_OK;
}

int build_index_prims(ptgpu_ctx *ctx, const Shape *d_shapes, int n, DevBuf &out_prims) {
    PT_TRY(out_prims.ensure((size_t)n * sizeof(Prim)));
    PT_TRY(build_index_prims_into(ctx, d_shapes, n, out_prims.as<Prim>()));
    return PTGPU_OK;
}

// one buffer: the fp32 four-wide nodes, then their binary16 copy
size_t wide_bytes(const DevAccel &acc) { return (size_t)(acc.nnodes > 0 ? acc.nnodes : 0) * (sizeof(TNode4) + sizeof(TNode4H)); }

int build_wide(ptgpu_ctx *ctx, DevAccel *acc, DevBuf &out_nodes4) {
    acc->nodes4 = nullptr;
    acc->nodes4h = nullptr;
    if (acc->mode == PTGPU_ACCEL_NONE || acc->root < 0 || acc->nnodes <= 0 || !acc->nodes) return PTGPU_OK;
    PT_TRY(out_nodes4.ensure(wide_bytes(*acc)));
    return build_wide_into(ctx, acc, out_nodes4.as<TNode4>());
}

int build_wide_into(ptgpu_ctx *ctx, DevAccel *acc, TNode4 *w4) {
    acc->nodes4 = nullptr;
    acc->nodes4h = nullptr;
    if (acc->mode == PTGPU_ACCEL_NONE || acc->root < 0 || acc->nnodes <= 0 || !acc->nodes) return PTGPU_OK;
    const size_t n4 = (size_t)acc->nnodes;
    const dim3 grid((unsigned)((n4 + 255) / 256));
    hipLaunchKernelGGL(k_collapse4, grid, dim3(256), 0, ctx->stream, acc->nodes, acc->nnodes, w4);
    PT_HIP(hipGetLastError());
    acc->nodes4 = w4;
    if (acc->leaf_runs) {
        TNode4H *w4h = reinterpret_cast<TNode4H *>(w4 + n4);
        hipLaunchKernelGGL(k_half4, grid, dim3(256), 0, ctx->stream, w4, acc->nnodes, w4h);
        PT_HIP(hipGetLastError());
        acc->nodes4h = w4h;
    }
    return PTGPU_OK;
}

size_t light_scan_bytes(int n) { return ((size_t)((n + kLightChunk - 1) / kLightChunk) + 4) * 4; }

int build_light_list(ptgpu_ctx *ctx, const Shape *d_shapes, int n, int light_cnt, DevBuf &out, DevBuf &scan,
                     int *found) {
    *found = 0;
    if (light_cnt <= 0 || n <= 0) {
        PT_TRY(out.ensure(4));
        return PTGPU_OK;
    }
    PT_NO_CAPTURE(ctx, "building a light list");
    PT_TRY(scan.ensure(light_scan_bytes(n)));
    PT_TRY(out.ensure((size_t)light_cnt * 4));
    return build_light_list_into(ctx, d_shapes, n, light_cnt, out.as<int>(), scan.as<unsigned>(), found);
}

int build_light_list_into(ptgpu_ctx *ctx, const Shape *d_shapes, int n, int light_cnt, int *out, unsigned *scan,
                          int *found) {
    *found = 0;
    if (light_cnt <= 0 || n <= 0) return PTGPU_OK;
    PT_NO_CAPTURE(ctx, "building a light list");
    const int nb = (n + kLightChunk - 1) / kLightChunk;
    unsigned *counts = scan;
    unsigned *total = counts + nb;
    hipLaunchKernelGGL(k_light_count, dim3(nb), dim3(256), 0, ctx->stream, d_shapes, n, counts);
    PT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_light_scan, dim3(1), dim3(64), 0, ctx->stream, counts, nb, total);
    PT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_light_write, dim3(nb), dim3(256), 0, ctx->stream, d_shapes, n, counts, light_cnt, out);
    PT_HIP(hipGetLastError());
    unsigned h_total = 0;
    PT_HIP(hipMemcpyAsync(&h_total, total, 4, hipMemcpyDeviceToHost, ctx->stream));
    PT_HIP(hipStreamSynchronize(ctx->stream));
    *found = (int)h_total < light_cnt ? (int)h_total : light_cnt;
    return PTGPU_OK;
}

size_t ref_parents_bytes(int n) { return (size_t)(n > 1 ? 2 * n - 1 : 1) * 4 + 64; }

int convert_bvh_ref(ptgpu_ctx *ctx, const Shape *d_shapes, int n, const BVHNodeGPU *d_btn, const BVHNodeGPU *d_btl,
                    DevBuf &nodes, DevBuf &prims, DevAccel *acc) {
    PT_TRY(ctx->tmp_parents.ensure(ref_parents_bytes(n)));
    PT_TRY(nodes.ensure((size_t)(n > 1 ? n - 1 : 1) * sizeof(TNode)));
    PT_TRY(prims.ensure((size_t)n * sizeof(Prim)));
    return convert_bvh_ref_into(ctx, d_shapes, n, d_btn, d_btl, nodes.as<TNode>(), prims.as<Prim>(),
                                ctx->tmp_parents.as<int>(), acc);
}

int convert_bvh_ref_into(ptgpu_ctx *ctx, const Shape *d_shapes, int n, const BVHNodeGPU *d_btn,
                         const BVHNodeGPU *d_btl, TNode *nodes, Prim *prims, int *parents, DevAccel *acc) {
    acc->depth = 0;
    if (n >= 2) {
        // The caller's links are checked before any kernel follows them, and the stack is sized from
        // the real height: a caller's tree (getTrees after optimize, or any other builder's) can be
        // much deeper than ceil(log2 n).
        PT_NO_CAPTURE(ctx, "measuring a caller's BVH");
        int *par = parents;
        unsigned *err = reinterpret_cast<unsigned *>(par + (2 * n - 1));
        PT_TRY(fill_u32(ctx->stream, par, 0xFFFFFFFFu, (size_t)(2 * n - 1)));
        PT_TRY(fill_u32(ctx->stream, err, 0u, 2));
        hipLaunchKernelGGL(k_ref_parents, dim3((n + 254) / 256), dim3(256), 0, ctx->stream, d_btn, n, par, err);
        PT_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_ref_height, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, d_btl, par, n, err);
        PT_HIP(hipGetLastError());
        unsigned h_err[2] = {0, 0};
        PT_HIP(hipMemcpyAsync(h_err, err, 8, hipMemcpyDeviceToHost, ctx->stream));
        PT_HIP(hipStreamSynchronize(ctx->stream));
        PT_CHECK(!(h_err[0] & 3u), PTGPU_ERR_INVALID, "btn/btl: %s",
                 (h_err[0] & 1u) ? "a child index is out of range" : "a node has two parents (not a tree)");
        PT_CHECK(!(h_err[0] & 4u), PTGPU_ERR_INVALID, "btn/btl: a leaf is not reachable from the root");
        PT_CHECK(!(h_err[0] & 16u), PTGPU_ERR_INVALID, "btl: a leaf's nShape is out of range");
        PT_CHECK(!(h_err[0] & 8u), PTGPU_ERR_UNSUPPORTED, "btn/btl: tree deeper than %d levels (or cyclic)",
                 kMaxRefHeight);
        acc->depth = (int)h_err[1];
    } else {
        PT_CHECK(n == 1, PTGPU_ERR_INVALID, "shapeCnt %d < 1", n);
        int ns = -1;
        PT_HIP(hipMemcpyAsync(&ns, &d_btl[0].nShape, 4, hipMemcpyDeviceToHost, ctx->stream));
        PT_HIP(hipStreamSynchronize(ctx->stream));
        PT_CHECK(ns == 0, PTGPU_ERR_INVALID, "btl: a leaf's nShape is out of range");
    }
    hipLaunchKernelGGL(k_convert_bvh, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, d_shapes, n, d_btn, d_btl,
                       nodes, prims);
    PT_HIP(hipGetLastError());
    acc->nodes = nodes;
    acc->prims = prims;
    acc->nprims = n;
    acc->nnodes = n - 1;
    acc->nshapes = n;
    acc->mode = PTGPU_ACCEL_BVH;
    acc->leaf_runs = 1;
    acc->root = n >= 2 ? 0 : ~0;
    acc->root_cnt = n >= 2 ? 0 : 1;
    return PTGPU_OK;
}

int convert_kd_ref(ptgpu_ctx *ctx, const Shape *d_shapes, int n, const KDNodeGPU *d_kng, int kngCnt, const int *d_kn,
                   int knCnt, DevBuf &nodes, DevBuf &prims, DevAccel *acc) {
    PT_CHECK(kngCnt >= 2, PTGPU_ERR_INVALID, "kngCnt %d < 2", kngCnt);
    PT_TRY(nodes.ensure((size_t)kngCnt * sizeof(TNode)));
    PT_TRY(prims.ensure((size_t)(knCnt > 0 ? knCnt : 1) * sizeof(Prim)));
    return convert_kd_ref_into(ctx, d_shapes, n, d_kng, kngCnt, d_kn, knCnt, nodes.as<TNode>(), prims.as<Prim>(), acc);
}

int convert_kd_ref_into(ptgpu_ctx *ctx, const Shape *d_shapes, int n, const KDNodeGPU *d_kng, int kngCnt,
                        const int *d_kn, int knCnt, TNode *nodes, Prim *prims, DevAccel *acc) {
    PT_CHECK(kngCnt >= 2, PTGPU_ERR_INVALID, "kngCnt %d < 2", kngCnt);
    PT_NO_CAPTURE(ctx, "converting a KD heap");
    const int m = kngCnt > knCnt ? kngCnt : knCnt;
    hipLaunchKernelGGL(k_convert_kd, dim3((m + 255) / 256), dim3(256), 0, ctx->stream, d_shapes, d_kng, kngCnt, d_kn,
                       knCnt, nodes, prims);
    PT_HIP(hipGetLastError());
    KDNodeGPU root;
    PT_HIP(hipMemcpyAsync(&root, d_kng + 1, sizeof(KDNodeGPU), hipMemcpyDeviceToHost, ctx->stream));
    PT_HIP(hipStreamSynchronize(ctx->stream));
    acc->nodes = nodes;
    acc->prims = prims;
    acc->nprims = knCnt;
    acc->nnodes = kngCnt;
    acc->nshapes = n;
    acc->mode = PTGPU_ACCEL_KD;
    if (root.leaf == 1) {
        acc->root = ~(int)root.min;
        acc->root_cnt = (int)root.max - (int)root.min;
    } else {
        acc->root = 1;
        acc->root_cnt = 0;
    }
    int d = 0;
    while ((1ll << (d + 1)) < kngCnt) d++;
    acc->depth = d;
    return PTGPU_OK;
}

// four-wide walk: at most three pushes per visit and one visit per two binary levels; `least` raises
// the class for tests that run every stack instance on a shallow tree
static int stack_for_wide(const DevAccel &acc, int least) {
    int need = 3 * ((acc.depth + 2) / 2) + 1;
    if (need < least) need = least;
    if (need <= 32) return 32;
    if (need <= 48) return 48;   // 48 KB of LDS per workgroup: three workgroups per CU
    if (need <= 64) return 64;
    return -1;
}

// the deepest tree each stack holds (inverse of the two functions above and below)
static int wide_depth_cap(int stack) { return 2 * ((stack - 1) / 3) - 1; }
static int binary_depth_cap(int stack) { return stack - 2; }

static int stack_for(const DevAccel &acc, int least) {
    const int need = acc.mode == PTGPU_ACCEL_NONE ? 0 : acc.depth + 2;
    if (need <= 24 && least <= 24) return 24;
    if (need <= 64) return 64;
    return -1;
}

int launch_render(ptgpu_ctx *ctx, const DevAccel &acc, const RenderLaunch &L) {
    // work items: every pixel, or every pixel slot of the listed tiles (slots past the frame's edge are skipped)
    const int total = L.tiles ? L.ntiles * L.tile_w * L.tile_h : L.width * L.height;
    if (total <= 0) return PTGPU_OK;
    const bool wf_ok = acc.mode != PTGPU_ACCEL_NONE && wf_supported(L.nlights) && !L.trav_stats && !L.tiles;
    if (ctx->engine == 2)
        PT_CHECK(wf_ok, PTGPU_ERR_UNSUPPORTED, "wavefront engine: needs a tree, <= 8 lights, no stats, no tiles");
    // the wavefront schedule only on request: one frame keeps at most W*H paths in flight, too few
    // to amortise its per-bounce barriers (measured 2x slower than the megakernels at 512x512)
    PT_CHECK(!(acc.height_check && ctx->engine == 2), PTGPU_ERR_UNSUPPORTED,
             "wavefront engine: a traversal recorded before the tree height was read back");
    if (wf_ok && ctx->engine == 2) return launch_render_wavefront(ctx, acc, L);
    const int st = stack_for(acc, ctx->stack_min);
    PT_CHECK(st > 0, PTGPU_ERR_UNSUPPORTED, "tree depth %d exceeds the 62-level traversal stack", acc.depth);
    unsigned *work_counter = L.work_counter;
    if (!work_counter) {
        PT_TRY(ctx->tmp_counter.ensure(64));
        work_counter = ctx->tmp_counter.as<unsigned>();
    }
    PT_TRY(fill_u32(ctx->stream, work_counter, 0u, 1));
    RenderParams P;
    P.acc = acc;
    P.colors = L.colors;
    P.seeds = L.seeds;
    P.camera = L.camera;
    P.shapes = L.shapes;
    P.lights = L.lights;
    P.nlights = L.nlights;
    P.width = L.width;
    P.height = L.height;
    P.currentSample = L.currentSample;
    P.spp = L.spp;
    P.max_depth = L.max_depth;
    P.pixels = L.pixels;
    P.ray_counter = L.ray_counter;
    P.trav_stats = L.trav_stats;
    P.work_counter = work_counter;
    P.total = total;
    P.regroup = ctx->regroup;
    P.quad = ctx->quad;
    P.jobs = nullptr;
    P.job_begin = nullptr;
    P.njobs = 0;
    P.tiles = L.tiles;
    P.tile_w = L.tile_w;
    P.tile_h = L.tile_h;
    P.one = 1.0f;
    P.depth_cap = binary_depth_cap(st);
    P.status = nullptr;
    if (acc.height_check) {
        PT_TRY(ctx->rec_status.ensure(16));   // allocated at context creation; kept here for safety
        P.status = ctx->rec_status.as<unsigned>();
    }
    // persistent waves: enough workgroups to fill every CU several times, never more than needed
    int blocks = ctx->cu_count * ctx->blocks_per_cu;
    const int need = (total + 255) / 256;
    if (blocks > need) blocks = need;
    if (blocks < 1) blocks = 1;
    // auto and engine 3: per-lane states whenever there is a tree to walk (measured fastest at every
    // size tried); brute force and single-leaf scenes keep the lockstep megakernel
    if ((ctx->engine == 0 || ctx->engine == 3) && acc.mode != PTGPU_ACCEL_NONE && acc.root >= 0) {
        const int sw = acc.nodes4 && ctx->wide ? stack_for_wide(acc, ctx->stack_min) : -1;
        const bool half = ctx->wide == 2 && acc.nodes4h;
        if (sw > 0) P.depth_cap = wide_depth_cap(sw);
        // the binary16 walk of a tree in the 48-entry class runs on the short stack, four workgroups
        // per CU (short_stack 2, a test setting, runs every binary16 walk on the 8-entry class).  Held
        // to 128 registers, the short class keeps one in scratch memory, which the allocation-free
        // entry does not ask the runtime for.
        const bool tiny = half && sw > 0 && ctx->short_stack == 2 && !L.no_alloc;
        const bool shortst = half && sw == 48 && ctx->short_stack == 1 && !L.no_alloc;
#define PT_LANES(ST, STATS_, WIDE_, HALF_)                                                                      \
    hipLaunchKernelGGL((k_render_lanes<ST, STATS_, WIDE_, false, HALF_>), dim3(blocks), dim3(256), 0, ctx->stream, P)
#define PT_LANES_W(ST, STATS_) \
    do { if (half) PT_LANES(ST, STATS_, true, true); else PT_LANES(ST, STATS_, true, false); } while (0)
        if (tiny) {
            if (L.trav_stats) PT_LANES(kStackTiny, true, true, true);
            else PT_LANES(kStackTiny, false, true, true);
        } else if (shortst) {
            if (L.trav_stats) PT_LANES(kStackShort, true, true, true);
            else PT_LANES(kStackShort, false, true, true);
        } else if (sw > 0) {
            if (L.trav_stats) {
                if (sw == 32) PT_LANES_W(32, true);
                else if (sw == 48) PT_LANES_W(48, true);
                else PT_LANES_W(64, true);
            } else {
                if (sw == 32) PT_LANES_W(32, false);
                else if (sw == 48) PT_LANES_W(48, false);
                else PT_LANES_W(64, false);
            }
        } else if (L.trav_stats) {
            if (st == 24) PT_LANES(24, true, false, false);
            else PT_LANES(64, true, false, false);
        } else {
            if (st == 24) PT_LANES(24, false, false, false);
            else PT_LANES(64, false, false, false);
        }
#undef PT_LANES_W
#undef PT_LANES
    } else if (L.trav_stats) {
        if (st == 24) hipLaunchKernelGGL((k_render<24, true>), dim3(blocks), dim3(256), 0, ctx->stream, P);
        else hipLaunchKernelGGL((k_render<64, true>), dim3(blocks), dim3(256), 0, ctx->stream, P);
    } else {
        if (st == 24) hipLaunchKernelGGL((k_render<24, false>), dim3(blocks), dim3(256), 0, ctx->stream, P);
        else hipLaunchKernelGGL((k_render<64, false>), dim3(blocks), dim3(256), 0, ctx->stream, P);
    }
    PT_HIP(hipGetLastError());
    return PTGPU_OK;
}

int launch_trace(ptgpu_ctx *ctx, const DevAccel &acc, unsigned int shapeCnt, const Ray *d_rays, const float *d_maxt,
                 int n, float *d_t, int *d_id, int *d_hit, bool any) {
    (void)shapeCnt;
    if (n <= 0) return PTGPU_OK;
    PT_CHECK(!acc.height_check, PTGPU_ERR_UNSUPPORTED,
             "batch queries cannot be recorded before the tree height was read back");
    const int st = stack_for(acc, ctx->stack_min);
    PT_CHECK(st > 0, PTGPU_ERR_UNSUPPORTED, "tree depth %d exceeds the 62-level traversal stack", acc.depth);
    if (st == 24)
        hipLaunchKernelGGL(k_trace<24>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, acc, d_rays, d_maxt, n, d_t,
                           d_id, d_hit, any ? 1 : 0, (unsigned *)nullptr);
    else
        hipLaunchKernelGGL(k_trace<64>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, acc, d_rays, d_maxt, n, d_t,
                           d_id, d_hit, any ? 1 : 0, (unsigned *)nullptr);
    PT_HIP(hipGetLastError());
    return PTGPU_OK;
}

}  // namespace ptgpu

using namespace ptgpu;

// Resolve the traversal structure for (bvh, kd, shapes); NONE mode builds index-order prims.
static int resolve_accel(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const Shape *d_shapes, unsigned n,
                         DevAccel *acc, bool exact = false);

extern "C" int ptgpu_kdtree_device_accel(ptgpu_kdtree *kd, ptgpu::DevAccel *acc);

static int resolve_accel(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const Shape *d_shapes, unsigned n,
                         DevAccel *acc, bool exact) {
    PT_CHECK(!(bvh && kd), PTGPU_ERR_INVALID, "pass at most one of bvh / kd");
    if (bvh) {
        PT_CHECK(bvh->state >= 3, PTGPU_ERR_INVALID, "BVH not built (call build_bvh_tree)");
        PT_CHECK((unsigned)bvh->n == n, PTGPU_ERR_INVALID, "shapeCnt %u differs from the BVH's %d", n, bvh->n);
        PT_TRY(clbvh_prepare_traversal(bvh, ctx->capturing, exact));
        PT_CHECK(bvh->ctx->device == ctx->device, PTGPU_ERR_INVALID, "BVH built on device %d, render on device %d",
                 bvh->ctx->device, ctx->device);
        if (bvh->ctx != ctx && bvh->ctx->stream != ctx->stream) {
            // a handle built on another context's stream: this stream waits for that stream's work on
            // the handle (the build and the derived traversal layout) before it traverses
            PT_HIP(hipEventRecord(bvh->ev[10], bvh->ctx->stream));
            PT_HIP(hipStreamWaitEvent(ctx->stream, bvh->ev[10], 0));
        }
        *acc = bvh->acc;
        return PTGPU_OK;
    }
    if (kd) {
        PT_TRY(ptgpu_kdtree_device_accel(kd, acc));
        PT_CHECK((unsigned)acc->nshapes == n, PTGPU_ERR_INVALID, "shapeCnt %u differs from the KD tree's %d", n,
                 acc->nshapes);
        return PTGPU_OK;
    }
    PT_TRY(build_index_prims(ctx, d_shapes, (int)n, ctx->tmp_prims));
    acc->nodes = nullptr;
    acc->prims = ctx->tmp_prims.as<Prim>();
    acc->nprims = (int)n;
    acc->nnodes = 0;
    acc->nshapes = (int)n;
    acc->root = ~0;
    acc->root_cnt = (int)n;
    acc->depth = 0;
    acc->mode = PTGPU_ACCEL_NONE;
    return PTGPU_OK;
}

// RadianceGPU (box == NULL) and RadianceBoxGPU (box = {x, y, bwidth, bheight}) over reference-layout trees.
static int radiance_gpu_impl(ptgpu_ctx *ctx, int accel, const ptgpu_radiance_args *a, const int *box) {
    PT_CHECK(ctx && a, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(a->colors && a->seeds && a->camera && a->shapes && a->pixels, PTGPU_ERR_INVALID, "null buffer");
    PT_CHECK(a->shapeCnt >= 1, PTGPU_ERR_INVALID, "shapeCnt is 0");
    PT_CHECK(a->width > 0 && a->height > 0, PTGPU_ERR_INVALID, "bad frame size %dx%d", a->width, a->height);
    PT_CHECK((unsigned long long)a->width * (unsigned long long)a->height < (1ull << 30), PTGPU_ERR_UNSUPPORTED,
             "frame %dx%d: 2^30 pixels or more (seed indices 2*pixel must stay within int)", a->width,
             a->height);
    PT_CHECK(a->currentSample >= 0, PTGPU_ERR_INVALID, "currentSample < 0");
    PT_HIP(hipSetDevice(ctx->device));
    const int n = (int)a->shapeCnt;
    DevAccel acc;
    if (accel == PTGPU_ACCEL_BVH) {
        PT_CHECK(a->btn && a->btl, PTGPU_ERR_INVALID, "BVH mode needs btn and btl");
        PT_TRY(convert_bvh_ref(ctx, a->shapes, n, a->btn, a->btl, ctx->tmp_nodes, ctx->tmp_prims, &acc));
        PT_TRY(build_wide(ctx, &acc, ctx->tmp_nodes4));
    } else if (accel == PTGPU_ACCEL_KD) {
        PT_CHECK(a->kng && a->kn, PTGPU_ERR_INVALID, "KD mode needs kng and kn");
        PT_TRY(convert_kd_ref(ctx, a->shapes, n, a->kng, a->kngCnt, a->kn, a->knCnt, ctx->tmp_nodes, ctx->tmp_prims,
                              &acc));
        PT_TRY(build_wide(ctx, &acc, ctx->tmp_nodes4));
    } else {
        PT_CHECK(accel == PTGPU_ACCEL_NONE, PTGPU_ERR_INVALID, "unknown accel mode %d", accel);
        PT_TRY(resolve_accel(ctx, nullptr, nullptr, a->shapes, a->shapeCnt, &acc));
    }
    int nl = 0;
    PT_TRY(build_light_list(ctx, a->shapes, n, (int)a->lightCnt, ctx->tmp_lights, ctx->tmp_scan, &nl));
    RenderLaunch L;
    L.colors = a->colors;
    L.seeds = a->seeds;
    L.camera = a->camera;
    L.shapes = a->shapes;
    L.shapeCnt = a->shapeCnt;
    L.lights = ctx->tmp_lights.as<int>();
    L.nlights = nl;
    L.width = a->width;
    L.height = a->height;
    L.currentSample = a->currentSample;
    L.spp = 1;
    L.max_depth = ctx->max_depth;
    L.pixels = a->pixels;
    L.ray_counter = nullptr;
    L.trav_stats = nullptr;
    if (box) {   // the one box, as a one-entry tile list
        PT_TRY(ctx->tmp_tiles.ensure(64));
        PT_HIP(hipMemcpyAsync(ctx->tmp_tiles.p, box, 8, hipMemcpyHostToDevice, ctx->stream));
        L.tiles = ctx->tmp_tiles.as<int2>();
        L.ntiles = 1;
        L.tile_w = box[2];
        L.tile_h = box[3];
    }
    return launch_render(ctx, acc, L);
}

extern "C" {

int ptgpu_radiance_gpu(ptgpu_ctx *ctx, int accel, const ptgpu_radiance_args *a) {
    return radiance_gpu_impl(ctx, accel, a, nullptr);
}

int ptgpu_radiance_box_gpu(ptgpu_ctx *ctx, int accel, const ptgpu_radiance_args *a, int x, int y, int bwidth,
                           int bheight) {
    PT_CHECK(a, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(x >= 0 && y >= 0 && bwidth >= 1 && bheight >= 1 && x < a->width && y < a->height &&
                 (long long)bwidth * bheight < (1ll << 30),
             PTGPU_ERR_INVALID, "bad box (%d,%d) %dx%d in a %dx%d frame", x, y, bwidth, bheight, a->width, a->height);
    const int box[4] = {x, y, bwidth, bheight};
    return radiance_gpu_impl(ctx, accel, a, box);
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// Allocation-free RadianceGPU (ptgpu_radiance_gpu_noalloc)
// ---------------------------------------------------------------------------------------------
namespace ptgpu {
constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;
constexpr int kScratchDevices = 64;
}  // namespace ptgpu

// The work space of ptgpu_radiance_gpu_noalloc: the traversal layout converted from the caller's tree
// (binary nodes, primitives, four-wide nodes), the parent map of a caller's BVH, the light list and
// the persistent waves' pixel counter.  Each GPU holds its own copy until the process ends; host code
// reaches it only through the address hipGetSymbolAddress returns for that GPU.
static __device__ __align__(256) unsigned char scratch[ptgpu::kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[ptgpu::kScratchDevices] = {};

// carves 256-byte-aligned pieces out of one device range
struct Carve {
    char *base;
    size_t used;
    template <class T> T *take(size_t bytes) {
        T *p = reinterpret_cast<T *>(base + used);
        used += (bytes + 255) & ~size_t(255);
        return p;
    }
};
size_t up256(size_t b) { return (b + 255) & ~size_t(255); }
}  // namespace

// Device address of this GPU's scratch array, looked up once per device with that device current
// (the context's set-up calls this; the entry point calls it again if the look-up failed then).
int ptgpu::noalloc_scratch_init(int device) {
    PT_CHECK(device >= 0 && device < kScratchDevices, PTGPU_ERR_UNSUPPORTED, "device %d: scratch table holds %d",
             device, kScratchDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return PTGPU_OK;
    void *ptr = nullptr;
    PT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    PT_CHECK(ptr != nullptr, PTGPU_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = ptr;
    return PTGPU_OK;
}

extern "C" {

int ptgpu_radiance_gpu_noalloc(ptgpu_ctx *ctx, int accel, const ptgpu_radiance_args *a) {
    PT_CHECK(ctx && a, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(a->colors && a->seeds && a->camera && a->shapes && a->pixels, PTGPU_ERR_INVALID, "null buffer");
    PT_CHECK(a->shapeCnt >= 1 && a->shapeCnt < (1u << 30), PTGPU_ERR_INVALID, "shapeCnt %u out of range", a->shapeCnt);
    PT_CHECK(a->width > 0 && a->height > 0, PTGPU_ERR_INVALID, "bad frame size %dx%d", a->width, a->height);
    PT_CHECK((unsigned long long)a->width * (unsigned long long)a->height < (1ull << 30), PTGPU_ERR_UNSUPPORTED,
             "frame %dx%d: 2^30 pixels or more (seed indices 2*pixel must stay within int)", a->width, a->height);
    PT_CHECK(a->currentSample >= 0, PTGPU_ERR_INVALID, "currentSample < 0");
    PT_CHECK(accel == PTGPU_ACCEL_NONE || accel == PTGPU_ACCEL_BVH || accel == PTGPU_ACCEL_KD, PTGPU_ERR_INVALID,
             "unknown accel mode %d", accel);
    PT_CHECK(ctx->engine != 2, PTGPU_ERR_UNSUPPORTED,
             "the allocation-free entry runs the megakernel engines (the wavefront engine allocates its queues)");
    const int n = (int)a->shapeCnt;
    // the scratch this call needs, checked before any device work
    size_t nodes_b = 0, prims_b = 0, wide_b = 0, par_b = 0;
    if (accel == PTGPU_ACCEL_BVH) {
        PT_CHECK(a->btn && a->btl, PTGPU_ERR_INVALID, "BVH mode needs btn and btl");
        nodes_b = (size_t)(n > 1 ? n - 1 : 1) * sizeof(TNode);
        prims_b = (size_t)n * sizeof(Prim);
        wide_b = (size_t)(n > 1 ? n - 1 : 0) * (sizeof(TNode4) + sizeof(TNode4H));
        par_b = ref_parents_bytes(n);
    } else if (accel == PTGPU_ACCEL_KD) {
        PT_CHECK(a->kng && a->kn, PTGPU_ERR_INVALID, "KD mode needs kng and kn");
        PT_CHECK(a->kngCnt >= 2 && a->knCnt >= 0, PTGPU_ERR_INVALID, "kngCnt %d / knCnt %d", a->kngCnt, a->knCnt);
        nodes_b = (size_t)a->kngCnt * sizeof(TNode);
        prims_b = (size_t)(a->knCnt > 0 ? a->knCnt : 1) * sizeof(Prim);
        wide_b = (size_t)a->kngCnt * (sizeof(TNode4) + sizeof(TNode4H));
    } else {
        prims_b = (size_t)n * sizeof(Prim);
    }
    const size_t lights_b = (size_t)(a->lightCnt > 0 ? a->lightCnt : 1) * 4;
    const size_t need = up256(64) + up256(nodes_b) + up256(prims_b) + up256(wide_b) + up256(par_b) + up256(lights_b) +
                        up256(light_scan_bytes(n));
    PT_CHECK(need <= kNoAllocScratchBytes, PTGPU_ERR_UNSUPPORTED,
             "this frame's tree needs %zu bytes of scratch; the allocation-free entry holds %zu "
             "(ptgpu_radiance_gpu has no such limit)",
             need, kNoAllocScratchBytes);
    PT_HIP(hipSetDevice(ctx->device));
    PT_TRY(noalloc_scratch_init(ctx->device));
    Carve cv{static_cast<char *>(g_scratch_ptr[ctx->device]), 0};
    unsigned *counter = cv.take<unsigned>(64);
    TNode *nodes = cv.take<TNode>(nodes_b);
    Prim *prims = cv.take<Prim>(prims_b);
    TNode4 *wide = cv.take<TNode4>(wide_b);
    int *parents = cv.take<int>(par_b);
    int *lights = cv.take<int>(lights_b);
    unsigned *scan = cv.take<unsigned>(light_scan_bytes(n));
    DevAccel acc;
    if (accel == PTGPU_ACCEL_BVH) {
        PT_TRY(convert_bvh_ref_into(ctx, a->shapes, n, a->btn, a->btl, nodes, prims, parents, &acc));
        PT_TRY(build_wide_into(ctx, &acc, wide));
    } else if (accel == PTGPU_ACCEL_KD) {
        PT_TRY(convert_kd_ref_into(ctx, a->shapes, n, a->kng, a->kngCnt, a->kn, a->knCnt, nodes, prims, &acc));
        PT_TRY(build_wide_into(ctx, &acc, wide));
    } else {
        PT_TRY(build_index_prims_into(ctx, a->shapes, n, prims));
        acc.nodes = nullptr;
        acc.prims = prims;
        acc.nprims = n;
        acc.nnodes = 0;
        acc.nshapes = n;
        acc.root = ~0;
        acc.root_cnt = n;
        acc.depth = 0;
        acc.mode = PTGPU_ACCEL_NONE;
    }
    int nl = 0;
    PT_TRY(build_light_list_into(ctx, a->shapes, n, (int)a->lightCnt, lights, scan, &nl));
    RenderLaunch L;
    L.colors = a->colors;
    L.seeds = a->seeds;
    L.camera = a->camera;
    L.shapes = a->shapes;
    L.shapeCnt = a->shapeCnt;
    L.lights = lights;
    L.nlights = nl;
    L.width = a->width;
    L.height = a->height;
    L.currentSample = a->currentSample;
    L.spp = 1;
    L.max_depth = ctx->max_depth;
    L.pixels = a->pixels;
    L.ray_counter = nullptr;
    L.trav_stats = nullptr;
    L.work_counter = counter;
    L.no_alloc = true;
    return launch_render(ctx, acc, L);
}

}  // extern "C"

// Checks a ptgpu_render argument set, resolves its traversal structure and light list (the handle's
// prepared layout, or index-order primitives without a tree) and fills the launch description.
int ptgpu::prepare_render(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const ptgpu_render_args *a,
                          DevAccel *acc, RenderLaunch *L) {
    PT_CHECK(ctx && a, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(a->colors && a->seeds && a->camera && a->shapes, PTGPU_ERR_INVALID, "null buffer");
    PT_CHECK(a->shapeCnt >= 1, PTGPU_ERR_INVALID, "shapeCnt is 0");
    PT_CHECK(a->width > 0 && a->height > 0, PTGPU_ERR_INVALID, "bad frame size %dx%d", a->width, a->height);
    PT_CHECK((unsigned long long)a->width * (unsigned long long)a->height < (1ull << 30), PTGPU_ERR_UNSUPPORTED,
             "frame %dx%d: 2^30 pixels or more (seed indices 2*pixel must stay within int)", a->width,
             a->height);
    PT_CHECK(a->spp >= 1 && a->currentSample >= 0, PTGPU_ERR_INVALID, "bad spp/currentSample");
    PT_CHECK(a->max_depth >= 0 && a->max_depth < 1024, PTGPU_ERR_INVALID, "bad max_depth");
    PT_HIP(hipSetDevice(ctx->device));
    PT_TRY(resolve_accel(ctx, bvh, kd, a->shapes, a->shapeCnt, acc));
    const int *lights = nullptr;
    int nl = 0;
    if (bvh) {
        if (bvh->lights_for != (int)a->lightCnt) {
            PT_TRY(build_light_list(ctx, a->shapes, (int)a->shapeCnt, (int)a->lightCnt, bvh->lights, bvh->scan,
                                    &bvh->nlights));
            bvh->lights_for = (int)a->lightCnt;
        }
        lights = bvh->lights.as<int>();
        nl = bvh->nlights;
    } else if (kd) {
        PT_TRY(kdtree_lights(kd, a->shapes, (int)a->shapeCnt, (int)a->lightCnt, &lights, &nl));
    } else {
        PT_TRY(build_light_list(ctx, a->shapes, (int)a->shapeCnt, (int)a->lightCnt, ctx->tmp_lights, ctx->tmp_scan,
                                &nl));
        lights = ctx->tmp_lights.as<int>();
    }
    L->colors = a->colors;
    L->seeds = a->seeds;
    L->camera = a->camera;
    L->shapes = a->shapes;
    L->shapeCnt = a->shapeCnt;
    L->lights = lights;
    L->nlights = nl;
    L->width = a->width;
    L->height = a->height;
    L->currentSample = a->currentSample;
    L->spp = a->spp;
    L->max_depth = a->max_depth;
    L->pixels = a->pixels;
    L->ray_counter = a->ray_counter;
    L->trav_stats = a->trav_stats;
    return PTGPU_OK;
}

int ptgpu::render_tiles(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const ptgpu_render_args *a,
                        const int2 *d_tiles, int ntiles, int tile_w, int tile_h, DevAccel *acc, RenderLaunch *L) {
    PT_CHECK(tile_w >= 1 && tile_h >= 1 && ntiles >= 0 && (long long)ntiles * tile_w * tile_h < (1ll << 30),
             PTGPU_ERR_INVALID, "bad tiles (%d of %dx%d)", ntiles, tile_w, tile_h);
    PT_TRY(prepare_render(ctx, bvh, kd, a, acc, L));
    if (ntiles == 0) return PTGPU_OK;
    PT_CHECK(d_tiles, PTGPU_ERR_INVALID, "null tile list");
    L->tiles = d_tiles;
    L->ntiles = ntiles;
    L->tile_w = tile_w;
    L->tile_h = tile_h;
    return launch_render(ctx, *acc, *L);
}

extern "C" {

int ptgpu_render(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const ptgpu_render_args *a) {
    DevAccel acc;
    RenderLaunch L;
    PT_TRY(prepare_render(ctx, bvh, kd, a, &acc, &L));
    return launch_render(ctx, acc, L);
}

int ptgpu_render_tiles(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const ptgpu_render_args *a,
                       const int *d_tile_origins, int ntiles, int tile_w, int tile_h) {
    DevAccel acc;
    RenderLaunch L;
    return render_tiles(ctx, bvh, kd, a, reinterpret_cast<const int2 *>(d_tile_origins), ntiles, tile_w, tile_h,
                        &acc, &L);
}

int ptgpu_render_batch(ptgpu_ctx *ctx, const ptgpu_render_job *jobs, int njobs, unsigned long long *d_ray_counter) {
    PT_CHECK(ctx && (jobs || njobs == 0) && njobs >= 0, PTGPU_ERR_INVALID, "null argument");
    if (njobs == 0) return PTGPU_OK;
    PT_HIP(hipSetDevice(ctx->device));
    ctx->batch.stream = ctx->stream;
    return launch_batch(ctx, ctx->batch, jobs, njobs, d_ray_counter);
}

}  // extern "C"

namespace ptgpu {
int launch_batch(ptgpu_ctx *ctx, BatchChannel &ch, const ptgpu_render_job *jobs, int njobs,
                 unsigned long long *d_ray_counter) {
    std::vector<RenderJob> table((size_t)njobs);
    std::vector<unsigned> begin((size_t)njobs + 1, 0u);
    std::vector<int> jst((size_t)njobs, 0);   // each job's traversal-stack class (32, 48 or 64 entries)
    int st = 0;
    bool half = ctx->wide == 2;   // binary16 nodes only when every job's tree has them
    bool checks = false;          // some job's tree height is checked on the device (recorded after a build)
    unsigned long long total = 0;
    for (int j = 0; j < njobs; j++) {
        const ptgpu_render_job &J = jobs[j];
        const ptgpu_render_args *a = &J.args;
        PT_CHECK(!!J.bvh != !!J.kd, PTGPU_ERR_INVALID, "job %d: pass exactly one of bvh / kd", j);
        for (int i = 0; i < j; i++)   // a handle holds one light list: jobs sharing it must agree on lightCnt
            PT_CHECK(!((jobs[i].bvh && jobs[i].bvh == J.bvh) || (jobs[i].kd && jobs[i].kd == J.kd)) ||
                         jobs[i].args.lightCnt == a->lightCnt,
                     PTGPU_ERR_INVALID, "jobs %d and %d share a tree but differ in lightCnt", i, j);
        PT_CHECK(a->colors && a->seeds && a->camera && a->shapes, PTGPU_ERR_INVALID, "job %d: null buffer", j);
        PT_CHECK(a->shapeCnt >= 2, PTGPU_ERR_UNSUPPORTED, "job %d: a batch needs at least two shapes per scene", j);
        PT_CHECK(a->width > 0 && a->height > 0, PTGPU_ERR_INVALID, "job %d: bad frame size %dx%d", j, a->width,
                 a->height);
        PT_CHECK((unsigned long long)a->width * (unsigned long long)a->height < (1ull << 30), PTGPU_ERR_UNSUPPORTED,
                 "job %d: frame %dx%d has 2^30 pixels or more", j, a->width, a->height);
        PT_CHECK(a->spp >= 1 && a->currentSample >= 0, PTGPU_ERR_INVALID, "job %d: bad spp/currentSample", j);
        PT_CHECK(a->max_depth >= 0 && a->max_depth < 1024, PTGPU_ERR_INVALID, "job %d: bad max_depth", j);
        RenderJob &R = table[(size_t)j];
        PT_TRY(resolve_accel(ctx, J.bvh, J.kd, a->shapes, a->shapeCnt, &R.acc));
        PT_CHECK(R.acc.root >= 0 && R.acc.nodes4, PTGPU_ERR_UNSUPPORTED, "job %d: the tree has no internal node", j);
        const int sw = stack_for_wide(R.acc, ctx->stack_min);
        PT_CHECK(sw > 0, PTGPU_ERR_UNSUPPORTED, "job %d: tree height %d exceeds the traversal stack", j, R.acc.depth);
        st = sw > st ? sw : st;
        jst[(size_t)j] = sw;
        half = half && R.acc.nodes4h;
        checks = checks || R.acc.height_check;
        if (J.bvh) {
            if (J.bvh->lights_for != (int)a->lightCnt) {
                PT_TRY(build_light_list(ctx, a->shapes, (int)a->shapeCnt, (int)a->lightCnt, J.bvh->lights,
                                        J.bvh->scan, &J.bvh->nlights));
                J.bvh->lights_for = (int)a->lightCnt;
            }
            R.lights = J.bvh->lights.as<int>();
            R.nlights = J.bvh->nlights;
        } else {
            PT_TRY(kdtree_lights(J.kd, a->shapes, (int)a->shapeCnt, (int)a->lightCnt, &R.lights, &R.nlights));
        }
        R.colors = a->colors;
        R.seeds = a->seeds;
        R.camera = a->camera;
        R.shapes = a->shapes;
        R.width = a->width;
        R.height = a->height;
        R.currentSample = a->currentSample;
        R.spp = a->spp;
        R.max_depth = a->max_depth;
        R.pixels = a->pixels;
        begin[(size_t)j] = (unsigned)total;
        total += (unsigned long long)a->width * (unsigned long long)a->height;
        PT_CHECK(total < (1ull << 31), PTGPU_ERR_UNSUPPORTED, "batch of more than 2^31 pixels");
    }
    begin[(size_t)njobs] = (unsigned)total;
    (void)st;
    (void)checks;
    // One launch per stack class, deepest first, in job order within a class.  The LDS stack of every
    // workgroup is sized for its launch's class, so a few deep trees (64 entries, 64 KB: two
    // workgroups per CU) do not hold the whole batch at that occupancy; the classes run at once, on
    // the channel's stream and two more forked from it.
    struct Cls {
        int st;
        std::vector<int> jobs;
        unsigned long long total;
        bool checks;
        size_t jo, bo;   // offsets of its jobs and its prefix sums in the table image
    };
    std::vector<Cls> cls;
    for (int c : {64, 48, 32}) {
        Cls K{c, {}, 0, false, 0, 0};
        for (int j = 0; j < njobs; j++)
            if (jst[(size_t)j] == c) {
                K.jobs.push_back(j);
                K.total += (unsigned long long)table[(size_t)j].width * (unsigned long long)table[(size_t)j].height;
                K.checks = K.checks || table[(size_t)j].acc.height_check;
            }
        if (!K.jobs.empty()) cls.push_back(std::move(K));
    }
    size_t img_bytes = 0;
    for (Cls &K : cls) {
        K.jo = img_bytes;
        img_bytes = (K.jo + K.jobs.size() * sizeof(RenderJob) + 63) & ~size_t(63);
        K.bo = img_bytes;
        img_bytes = (K.bo + (K.jobs.size() + 1) * sizeof(unsigned) + 63) & ~size_t(63);
    }
    std::vector<char> img(img_bytes, 0);
    for (Cls &K : cls) {
        unsigned run = 0;
        for (size_t q = 0; q < K.jobs.size(); q++) {
            const RenderJob &R = table[(size_t)K.jobs[q]];
            std::memcpy(img.data() + K.jo + q * sizeof(RenderJob), &R, sizeof(RenderJob));
            std::memcpy(img.data() + K.bo + q * sizeof(unsigned), &run, sizeof(unsigned));
            run += (unsigned)(R.width * R.height);
        }
        std::memcpy(img.data() + K.bo + K.jobs.size() * sizeof(unsigned), &run, sizeof(unsigned));
    }
    hipStream_t stream = ch.stream;
    if (!ch.copied) PT_HIP(hipEventCreateWithFlags(&ch.copied, hipEventDisableTiming));
    if (!ch.ready) PT_HIP(hipEventCreateWithFlags(&ch.ready, hipEventDisableTiming));
    if (stream != ctx->stream) {   // builds and layout conversion ran on the context's stream
        PT_HIP(hipEventRecord(ch.ready, ctx->stream));
        PT_HIP(hipStreamWaitEvent(stream, ch.ready, 0));
    }
    void *d_table = nullptr;
    if (ctx->capturing) {
        // recorded: the table gets a pinned copy and a device buffer of its own, kept with the graph
        PT_TRY(record_stage(ctx, stream, img.data(), img_bytes, img_bytes, &d_table));
    } else {
        // stage the table through pinned memory; the channel's previous copy must be done before reuse
        PT_HIP(hipEventSynchronize(ch.copied));
        if (ch.pinned_bytes < img_bytes) {
            if (ch.pinned) PT_HIP(hipHostFree(ch.pinned));
            ch.pinned = nullptr;
            ch.pinned_bytes = 0;
            PT_HIP(hipHostMalloc(&ch.pinned, img_bytes * 2, hipHostMallocDefault));
            ch.pinned_bytes = img_bytes * 2;
        }
        std::memcpy(ch.pinned, img.data(), img_bytes);
        PT_TRY(ch.table.ensure(img_bytes));
        PT_HIP(hipMemcpyAsync(ch.table.p, ch.pinned, img_bytes, hipMemcpyHostToDevice, stream));
        PT_HIP(hipEventRecord(ch.copied, stream));
        d_table = ch.table.p;
    }
    const int ncls = (int)cls.size();
    PT_TRY(ch.counter.ensure((size_t)64 * ncls));
    PT_TRY(fill_u32(stream, ch.counter.p, 0u, (size_t)16 * ncls));   // one work counter per class, 64 B apart
    if (ncls > 1) {
        if (!ch.fork) PT_HIP(hipEventCreateWithFlags(&ch.fork, hipEventDisableTiming));
        PT_HIP(hipEventRecord(ch.fork, stream));
    }
    for (int k = 0; k < ncls; k++) {
        const Cls &K = cls[(size_t)k];
        hipStream_t s = stream;
        if (k > 0) {
            if (!ch.sub[k - 1]) PT_HIP(hipStreamCreateWithFlags(&ch.sub[k - 1], hipStreamNonBlocking));
            if (!ch.join[k - 1]) PT_HIP(hipEventCreateWithFlags(&ch.join[k - 1], hipEventDisableTiming));
            s = ch.sub[k - 1];
            PT_HIP(hipStreamWaitEvent(s, ch.fork, 0));
        }
        RenderParams P;
        std::memset(static_cast<void *>(&P), 0, sizeof(P));
        P.acc = table[(size_t)K.jobs[0]].acc;   // unused by the batch variant
        P.ray_counter = d_ray_counter;
        P.trav_stats = nullptr;
        P.work_counter = ch.counter.as<unsigned>() + 16 * k;
        P.total = (int)K.total;
        P.regroup = ctx->batch_regroup;
        P.quad = ctx->quad;
        P.jobs = reinterpret_cast<const RenderJob *>(static_cast<char *>(d_table) + K.jo);
        P.job_begin = reinterpret_cast<const unsigned *>(static_cast<char *>(d_table) + K.bo);
        P.njobs = (int)K.jobs.size();
        P.one = 1.0f;
        P.depth_cap = wide_depth_cap(K.st);
        P.status = K.checks ? ctx->rec_status.as<unsigned>() : nullptr;
        long long blocks = (long long)ctx->cu_count * ctx->batch_blocks_per_cu;
        const long long need = ((long long)K.total + 255) / 256;
        if (blocks > need) blocks = need;
        if (blocks < 1) blocks = 1;
#define PT_BATCH(ST, HALF_) \
    hipLaunchKernelGGL((k_render_lanes<ST, false, true, true, HALF_>), dim3((unsigned)blocks), dim3(256), 0, s, P)
        if (half) {
            if (K.st == 32) PT_BATCH(32, true);
            else if (K.st == 48) PT_BATCH(48, true);
            else PT_BATCH(64, true);
        } else {
            if (K.st == 32) PT_BATCH(32, false);
            else if (K.st == 48) PT_BATCH(48, false);
            else PT_BATCH(64, false);
        }
#undef PT_BATCH
        PT_HIP(hipGetLastError());
        if (k > 0) {
            PT_HIP(hipEventRecord(ch.join[k - 1], s));
            PT_HIP(hipStreamWaitEvent(stream, ch.join[k - 1], 0));
        }
    }
    return PTGPU_OK;
}
}  // namespace ptgpu

extern "C" {

int ptgpu_trace_closest(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const Shape *d_shapes,
                        unsigned int shapeCnt, const Ray *d_rays, int n, float *d_t, int *d_id) {
    PT_CHECK(ctx && d_shapes && d_rays && d_t && d_id, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(shapeCnt >= 1 && n >= 0, PTGPU_ERR_INVALID, "bad sizes");
    PT_HIP(hipSetDevice(ctx->device));
    DevAccel acc;
    PT_TRY(resolve_accel(ctx, bvh, kd, d_shapes, shapeCnt, &acc, true));
    return launch_trace(ctx, acc, shapeCnt, d_rays, nullptr, n, d_t, d_id, nullptr, false);
}

int ptgpu_trace_any(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const Shape *d_shapes, unsigned int shapeCnt,
                    const Ray *d_rays, const float *d_maxt, int n, int *d_hit) {
    PT_CHECK(ctx && d_shapes && d_rays && d_maxt && d_hit, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(shapeCnt >= 1 && n >= 0, PTGPU_ERR_INVALID, "bad sizes");
    PT_HIP(hipSetDevice(ctx->device));
    DevAccel acc;
    PT_TRY(resolve_accel(ctx, bvh, kd, d_shapes, shapeCnt, &acc, true));
    return launch_trace(ctx, acc, shapeCnt, d_rays, d_maxt, n, nullptr, nullptr, d_hit, true);
}

}  // extern "C"
