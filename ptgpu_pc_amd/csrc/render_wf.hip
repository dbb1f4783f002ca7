// render_wf.hip -- wavefront path-tracing engine for gfx950.
//
// Same result, bit for bit, as the megakernel (render.hip) and the reference's RadianceGPU
// (rendering_kernel.cl:1072-1150): every path segment runs the same device functions in the same
// order on the same per-pixel RNG stream.  What changes is the schedule:
//
//   init      per pixel: seeds, first camera ray, path state; all paths queued.
//   repeat spp x (MAX_DEPTH+1) times:
//     trace     closest hit for the queued paths.  Persistent waves: a lane whose ray finished takes
//               the next ray from the queue (wave-wide ballot + one atomic), so lanes stay busy
//               however unequal the traversal lengths are.
//     shade     emission, specular/refractive continuation, or -- for diffuse hits -- the light
//               samples of SampleLights (randoms drawn in the reference's order) written as shadow
//               records plus the next bounce ray; finished samples update the running mean and
//               start the pixel's next sample (path regeneration).
//     shadow    any-hit for the shadow records (persistent waves as above).
//     resolve   Ld = sum over unoccluded records in light order, rad += throughput * Ld, then
//               continue or finish the sample.
//     advance   swap and reset the queue counters.
//   Queues are compacted with wave-aggregated atomics; the launch sequence has fixed grids and no
//   host synchronisation, so it is captured once into a HIP graph and replayed.
#include <cmath>
#include <cstring>

#include "lbvh.hpp"
#include "render_device.hpp"

namespace ptgpu {

struct WfBuffers {
    float4 *ro, *rd, *thr, *rad, *col;   // per path (pixel)
    uint2 *rng;
    int2 *meta;                           // x: depth | spec << 16, y: sample index k
    int2 *hit;                            // x: t bits, y: shape id or -1
    float4 *sh_o, *sh_d, *sh_c;           // per path x light: origin+maxt, dir+valid, contribution
    int *occl;
    unsigned *q[2], *qdiff, *qshadow;
    unsigned *cnt;                        // [0],[1] path queues, [2] diff, [3] shadow, [4],[5] fetch, [6],[7] rays
};

struct WfFrame {
    DevAccel acc;
    Vec *colors;
    unsigned *seeds;
    const Camera *camera;
    const Shape *shapes;
    const int *lights;
    int nlights;
    int width, height, currentSample, spp, max_depth;
    unsigned *pixels;
    int total;
};

__device__ __forceinline__ f3 ld3(const float4 &v) { return mk(v.x, v.y, v.z); }
__device__ __forceinline__ float4 st3(f3 v, float w = 0.f) { return make_float4(v.x, v.y, v.z, w); }

// Wave-aggregated append: every lane with `want` gets a distinct slot in queue.
__device__ __forceinline__ void wave_append(bool want, unsigned item, unsigned *queue, unsigned *count) {
    const unsigned long long m = __ballot(want);
    if (!m) return;
    const int leader = __ffsll((long long)m) - 1;
    unsigned base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(count, (unsigned)__popcll(m));
    base = __shfl(base, leader, kWave);
    if (want) queue[base + (unsigned)__popcll(m & lanemask_lt())] = item;
}

// ---------------------------------------------------------------------------------------------
// persistent traversal with per-lane refill (closest hit, or any hit for shadow records)
// ---------------------------------------------------------------------------------------------
template <int STACK, bool ANY>
__global__ __launch_bounds__(256) void k_wf_trace(DevAccel A, const unsigned *queue, const unsigned *count_ptr,
                                                  unsigned *fetch, const float4 *ray_o, const float4 *ray_d,
                                                  int2 *hit_out, int *occl_out, unsigned long long *ray_counter) {
    __shared__ int s_stack[STACK * 256];
    int *stk = s_stack + threadIdx.x;
    const unsigned count = *count_ptr;
    const unsigned long long lt = lanemask_lt();
    bool active = false, drained = count == 0;
    unsigned item = 0, taken = 0;
    f3 o = mk(0.f, 0.f, 0.f), d = mk(0.f, 0.f, 1.f), inv = mk(0.f, 0.f, 0.f);
    float tb = kInf;
    int bid = -1, node = 0, sp = 0;
    unsigned guard = 0;

    for (;;) {
        if (!drained) {
            const unsigned long long need = __ballot(!active);
            if (__popcll(need) >= 8 || need == __ballot(true)) {
                const unsigned cnt = (unsigned)__popcll(need);
                const int leader = __ffsll((long long)need) - 1;
                unsigned base = 0;
                if ((int)__lane_id() == leader) base = atomicAdd(fetch, cnt);
                base = __shfl(base, leader, kWave);
                if (base + cnt >= count) drained = true;
                if (!active) {
                    const unsigned my = base + (unsigned)__popcll(need & lt);
                    if (my < count) {
                        item = queue[my];
                        taken++;
                        const float4 a = ray_o[item], b = ray_d[item];
                        o = ld3(a);
                        d = ld3(b);
                        inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
                        tb = ANY ? a.w : kInf;   // any hit: maxt rides in ray_o.w
                        bid = -1;
                        sp = 0;
                        guard = 0;
                        node = A.root;
                        active = true;
                        if (A.root < 0) {        // the whole scene is one leaf
                            const int first = ~A.root;
                            bool hitany = false;
                            for (int q = 0; q < A.root_cnt; q++) {
                                int id;
                                const float t = prim_isect(A.prims, first + q, o, d, &id);
                                if (ANY) {
                                    if (t != 0.f && t < tb && id != A.nshapes - 1) hitany = true;
                                } else if (t != 0.f && (t < tb || (t == tb && id > bid))) {
                                    tb = t;
                                    bid = id;
                                }
                            }
                            if (ANY) occl_out[item] = hitany ? 1 : 0;
                            else hit_out[item] = make_int2(__float_as_int(bid >= 0 ? tb : 0.f), bid);
                            active = false;
                        }
                    }
                }
            }
        }
        if (__ballot(active) == 0) {
            if (drained) break;
            continue;
        }
        if (!active) continue;
        // one node visit
        bool done = false;
        if ((unsigned)node >= (unsigned)A.nnodes || ++guard > kMaxVisits) {
            done = true;   // malformed caller tree: give up on this ray
        } else {
            const TNode *np = A.nodes + node;
            const float4 a = np->a, b = np->b, c = np->c;
            const int4 r = np->r;
            float t0 = slab(a.x, a.y, a.z, a.w, b.x, b.y, o, inv, tb);
            float t1 = slab(b.z, b.w, c.x, c.y, c.z, c.w, o, inv, tb);
            bool found = false;
            for (int side = 0; side < 2; side++) {
                const int ref = side ? r.y : r.x;
                const int cntl = side ? r.w : r.z;
                float &ts = side ? t1 : t0;
                if (ref < 0 && ts != __builtin_inff()) {
                    const int first = ~ref;
                    if (first >= 0 && cntl <= A.nprims - first) {
                        for (int q = 0; q < cntl; q++) {
                            int id;
                            const float t = prim_isect(A.prims, first + q, o, d, &id);
                            if (ANY) {
                                if (t != 0.f && t < tb && id != A.nshapes - 1) found = true;
                            } else if (t != 0.f && (t < tb || (t == tb && id > bid))) {
                                tb = t;
                                bid = id;
                            }
                        }
                    }
                    ts = __builtin_inff();
                }
            }
            if (ANY && found) {
                occl_out[item] = 1;
                active = false;
            } else {
                const bool h0 = t0 != __builtin_inff(), h1 = t1 != __builtin_inff();
                if (h0 && h1) {
                    const bool first0 = ANY || t0 <= t1;
                    const int nearc = first0 ? r.x : r.y, farc = first0 ? r.y : r.x;
                    if (sp < STACK) stk[(sp++) * 256] = farc;
                    node = nearc;
                } else if (h0) {
                    node = r.x;
                } else if (h1) {
                    node = r.y;
                } else if (sp > 0) {
                    node = stk[(--sp) * 256];
                } else {
                    done = true;
                }
            }
        }
        if (active && done) {
            if (ANY) occl_out[item] = 0;
            else hit_out[item] = make_int2(__float_as_int(bid >= 0 ? tb : 0.f), bid);
            active = false;
        }
    }
    if (ray_counter) {
        unsigned long long v = taken;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
        if (__lane_id() == 0 && v) atomicAdd(ray_counter, v);
    }
}

// ---------------------------------------------------------------------------------------------
// init, shade, resolve, advance
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void start_path(const WfFrame &F, int gid, unsigned &s0, unsigned &s1, f3 &ro, f3 &rd) {
    const int x = gid % F.width, y = gid / F.width;
    camera_ray(F.camera, s0, s1, F.width, F.height, x, y, ro, rd);
}

__global__ __launch_bounds__(256) void k_wf_init(WfFrame F, WfBuffers B) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0) {
        B.cnt[0] = (unsigned)F.total;
        for (int k = 1; k < 6; k++) B.cnt[k] = 0;
    }
    if (gid >= F.total) return;
    const int x = gid % F.width, y = gid / F.width;
    const int ci = (F.height - y - 1) * F.width + x;
    unsigned s0 = F.seeds[2 * gid], s1 = F.seeds[2 * gid + 1];
    f3 col = mk(0.f, 0.f, 0.f);
    if (F.currentSample > 0) col = ldvec(F.colors + ci);
    f3 ro, rd;
    start_path(F, gid, s0, s1, ro, rd);
    B.ro[gid] = st3(ro);
    B.rd[gid] = st3(rd);
    B.thr[gid] = make_float4(1.f, 1.f, 1.f, 0.f);
    B.rad[gid] = make_float4(0.f, 0.f, 0.f, 0.f);
    B.col[gid] = st3(col);
    B.rng[gid] = make_uint2(s0, s1);
    B.meta[gid] = make_int2(1 << 16, 0);   // depth 0, specularBounce 1, sample 0
    B.q[0][gid] = (unsigned)gid;
}

// Finish the current sample of path p: running mean, then either the pixel's next sample (returns
// true: queue it) or the final write of colors / pixels / seeds.
__device__ __forceinline__ bool finish_sample(const WfFrame &F, const WfBuffers &B, int p, f3 rad, f3 &col,
                                              unsigned &s0, unsigned &s1, int &k, f3 &ro, f3 &rd, f3 &thr,
                                              int &depth, int &spec) {
    col = running_mean(col, rad, F.currentSample + k);
    k++;
    if (k < F.spp) {
        start_path(F, p, s0, s1, ro, rd);
        thr = mk(1.f, 1.f, 1.f);
        depth = 0;
        spec = 1;
        return true;
    }
    const int x = p % F.width, y = p / F.width;
    const int ci = (F.height - y - 1) * F.width + x;
    float *cp = reinterpret_cast<float *>(F.colors + ci);
    cp[0] = col.x; cp[1] = col.y; cp[2] = col.z;
    if (F.pixels) F.pixels[y * F.width + x] = pack_rgba(col);
    F.seeds[2 * p] = s0;
    F.seeds[2 * p + 1] = s1;
    return false;
}

__global__ __launch_bounds__(256) void k_wf_shade(WfFrame F, WfBuffers B, int cur) {
    const unsigned count = B.cnt[cur];
    const unsigned stride = gridDim.x * blockDim.x;
    const unsigned rounds = (count + stride - 1) / stride;
    for (unsigned r = 0; r < rounds; r++) {   // uniform trip count: ballots see every lane
        const unsigned e = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        const bool valid = e < count;
        bool to_next = false, to_diff = false;
        unsigned shadow_mask = 0;
        int p = 0;
        if (valid) {
            p = (int)B.q[cur][e];
            const int2 h = B.hit[p];
            const int id = h.y;
            f3 ro = ld3(B.ro[p]), rd = ld3(B.rd[p]), thr = ld3(B.thr[p]), rad = ld3(B.rad[p]);
            int2 meta = B.meta[p];
            int depth = meta.x & 0xFFFF, spec = meta.x >> 16, k = meta.y;
            uint2 rng = B.rng[p];
            unsigned s0 = rng.x, s1 = rng.y;
            bool finish = false;
            if (id < 0) {
                finish = true;
            } else {
                const float t = __int_as_float(h.x);
                const f3 hp = add3(ro, smul(t, rd));
                const Surface S = surface_at(F.shapes + id, hp, rd);
                if (!iszero3(S.e)) {
                    if (spec) rad = add3(rad, mul3(thr, smul(fabsf(S.dp), S.e)));
                    finish = true;
                } else if (S.refl == DIFF) {
                    spec = 0;
                    thr = mul3(thr, S.c);
                    for (int li = 0; li < F.nlights; li++) {
                        f3 sd, contrib;
                        float len;
                        const size_t slot = (size_t)p * F.nlights + li;
                        if (sample_light(F.shapes + F.lights[li], s0, s1, hp, S.nl, &sd, &len, &contrib)) {
                            B.sh_o[slot] = st3(hp, len - kEpsilon);
                            B.sh_d[slot] = st3(sd, 1.f);
                            B.sh_c[slot] = st3(contrib);
                            shadow_mask |= 1u << li;
                        } else {
                            B.sh_d[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
                        }
                    }
                    rd = diffuse_dir(S.nl, s0, s1);
                    ro = hp;
                    depth++;
                    to_diff = true;
                } else {
                    spec = 1;
                    specular_or_refract(S, rd, thr, s0, s1);
                    ro = hp;
                    depth++;
                    if (depth > F.max_depth) finish = true;
                    else to_next = true;
                }
            }
            f3 col = ld3(B.col[p]);
            if (finish) {
                to_next = finish_sample(F, B, p, rad, col, s0, s1, k, ro, rd, thr, depth, spec);
                rad = mk(0.f, 0.f, 0.f);
                B.col[p] = st3(col);
            }
            B.ro[p] = st3(ro);
            B.rd[p] = st3(rd);
            B.thr[p] = st3(thr);
            B.rad[p] = st3(rad);
            B.rng[p] = make_uint2(s0, s1);
            B.meta[p] = make_int2(depth | (spec << 16), k);
        }
        wave_append(to_next, (unsigned)p, B.q[cur ^ 1], &B.cnt[cur ^ 1]);
        wave_append(to_diff, (unsigned)p, B.qdiff, &B.cnt[2]);
        for (int li = 0; li < F.nlights; li++)
            wave_append((shadow_mask >> li) & 1u, (unsigned)((size_t)p * F.nlights + li), B.qshadow, &B.cnt[3]);
    }
}

__global__ __launch_bounds__(256) void k_wf_resolve(WfFrame F, WfBuffers B, int cur) {
    const unsigned count = B.cnt[2];
    const unsigned stride = gridDim.x * blockDim.x;
    const unsigned rounds = (count + stride - 1) / stride;
    for (unsigned r = 0; r < rounds; r++) {
        const unsigned e = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        bool to_next = false;
        int p = 0;
        if (e < count) {
            p = (int)B.qdiff[e];
            f3 thr = ld3(B.thr[p]), rad = ld3(B.rad[p]);
            f3 Ld = mk(0.f, 0.f, 0.f);
            for (int li = 0; li < F.nlights; li++) {
                const size_t slot = (size_t)p * F.nlights + li;
                if (B.sh_d[slot].w != 0.f && !B.occl[slot]) Ld = add3(Ld, ld3(B.sh_c[slot]));
            }
            Ld = mul3(thr, Ld);
            rad = add3(rad, Ld);
            int2 meta = B.meta[p];
            int depth = meta.x & 0xFFFF, spec = meta.x >> 16, k = meta.y;
            if (depth > F.max_depth) {
                uint2 rng = B.rng[p];
                unsigned s0 = rng.x, s1 = rng.y;
                f3 ro, rd;
                f3 col = ld3(B.col[p]);
                to_next = finish_sample(F, B, p, rad, col, s0, s1, k, ro, rd, thr, depth, spec);
                rad = mk(0.f, 0.f, 0.f);
                B.col[p] = st3(col);
                B.rng[p] = make_uint2(s0, s1);
                B.ro[p] = st3(ro);
                B.rd[p] = st3(rd);
                B.thr[p] = st3(thr);
                B.meta[p] = make_int2(depth | (spec << 16), k);
            } else {
                to_next = true;
            }
            B.rad[p] = st3(rad);
        }
        wave_append(to_next, (unsigned)p, B.q[cur ^ 1], &B.cnt[cur ^ 1]);
    }
}

__global__ void k_wf_advance(unsigned *cnt, int cur, unsigned long long *ray_counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        cnt[cur] = 0;
        cnt[2] = 0;
        cnt[3] = 0;
        cnt[4] = 0;
        cnt[5] = 0;
    }
}

// ---------------------------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------------------------
struct WfWorkspace {
    DevBuf path, shadow, queues, counters;
    size_t paths = 0, slots = 0;
};

static WfWorkspace &workspace(ptgpu_ctx *ctx) { return *reinterpret_cast<WfWorkspace *>(ctx->wf_workspace); }

int wf_supported(int nlights) { return nlights <= 8; }

int launch_render_wavefront(ptgpu_ctx *ctx, const DevAccel &acc, const RenderLaunch &L) {
    const int total = L.width * L.height;
    if (total <= 0) return PTGPU_OK;
    int st = 24;
    if (acc.mode != PTGPU_ACCEL_NONE) {
        if (acc.depth + 2 <= 24) st = 24;
        else if (acc.depth + 2 <= 64) st = 64;
        else PT_CHECK(false, PTGPU_ERR_UNSUPPORTED, "tree depth %d exceeds the traversal stack", acc.depth);
    }
    PT_CHECK(acc.mode != PTGPU_ACCEL_NONE, PTGPU_ERR_UNSUPPORTED, "wavefront engine needs a tree");
    if (!ctx->wf_workspace) ctx->wf_workspace = new WfWorkspace();
    WfWorkspace &W = workspace(ctx);
    const int nl = L.nlights;
    const size_t P = (size_t)total;
    const size_t S = P * (size_t)(nl > 0 ? nl : 1);
    PT_TRY(W.path.ensure(P * (5 * sizeof(float4) + sizeof(uint2) + 2 * sizeof(int2))));
    PT_TRY(W.shadow.ensure(S * (3 * sizeof(float4) + sizeof(int))));
    PT_TRY(W.queues.ensure((3 * P + S) * sizeof(unsigned)));
    PT_TRY(W.counters.ensure(64));
    WfBuffers B;
    char *pp = static_cast<char *>(W.path.p);
    B.ro = reinterpret_cast<float4 *>(pp); pp += P * sizeof(float4);
    B.rd = reinterpret_cast<float4 *>(pp); pp += P * sizeof(float4);
    B.thr = reinterpret_cast<float4 *>(pp); pp += P * sizeof(float4);
    B.rad = reinterpret_cast<float4 *>(pp); pp += P * sizeof(float4);
    B.col = reinterpret_cast<float4 *>(pp); pp += P * sizeof(float4);
    B.rng = reinterpret_cast<uint2 *>(pp); pp += P * sizeof(uint2);
    B.meta = reinterpret_cast<int2 *>(pp); pp += P * sizeof(int2);
    B.hit = reinterpret_cast<int2 *>(pp);
    char *sp = static_cast<char *>(W.shadow.p);
    B.sh_o = reinterpret_cast<float4 *>(sp); sp += S * sizeof(float4);
    B.sh_d = reinterpret_cast<float4 *>(sp); sp += S * sizeof(float4);
    B.sh_c = reinterpret_cast<float4 *>(sp); sp += S * sizeof(float4);
    B.occl = reinterpret_cast<int *>(sp);
    unsigned *qp = W.queues.as<unsigned>();
    B.q[0] = qp; B.q[1] = qp + P; B.qdiff = qp + 2 * P; B.qshadow = qp + 3 * P;
    B.cnt = W.counters.as<unsigned>();

    WfFrame F;
    F.acc = acc;
    F.colors = L.colors;
    F.seeds = L.seeds;
    F.camera = L.camera;
    F.shapes = L.shapes;
    F.lights = L.lights;
    F.nlights = nl;
    F.width = L.width;
    F.height = L.height;
    F.currentSample = L.currentSample;
    F.spp = L.spp;
    F.max_depth = L.max_depth;
    F.pixels = L.pixels;
    F.total = total;

    hipStream_t s = ctx->stream;
    const unsigned init_blocks = (unsigned)((P + 255) / 256);
    const unsigned grid_blocks = (unsigned)(ctx->cu_count * 8);
    const unsigned trace_blocks = (unsigned)(ctx->cu_count * (st == 24 ? 6 : 2));
    unsigned long long *rc0 = L.ray_counter, *rc1 = L.ray_counter ? L.ray_counter + 1 : nullptr;
    hipLaunchKernelGGL(k_wf_init, dim3(init_blocks), dim3(256), 0, s, F, B);
    PT_HIP(hipGetLastError());
    const int iters = L.spp * (L.max_depth + 1);
    for (int it = 0; it < iters; it++) {
        const int cur = it & 1;
        if (st == 24) {
            hipLaunchKernelGGL((k_wf_trace<24, false>), dim3(trace_blocks), dim3(256), 0, s, acc, B.q[cur],
                               B.cnt + cur, B.cnt + 4, (const float4 *)B.ro, (const float4 *)B.rd, B.hit,
                               (int *)nullptr, rc0);
        } else {
            hipLaunchKernelGGL((k_wf_trace<64, false>), dim3(trace_blocks), dim3(256), 0, s, acc, B.q[cur],
                               B.cnt + cur, B.cnt + 4, (const float4 *)B.ro, (const float4 *)B.rd, B.hit,
                               (int *)nullptr, rc0);
        }
        hipLaunchKernelGGL(k_wf_shade, dim3(grid_blocks), dim3(256), 0, s, F, B, cur);
        if (nl > 0) {
            if (st == 24) {
                hipLaunchKernelGGL((k_wf_trace<24, true>), dim3(trace_blocks), dim3(256), 0, s, acc, B.qshadow,
                                   B.cnt + 3, B.cnt + 5, (const float4 *)B.sh_o, (const float4 *)B.sh_d,
                                   (int2 *)nullptr, B.occl, rc1);
            } else {
                hipLaunchKernelGGL((k_wf_trace<64, true>), dim3(trace_blocks), dim3(256), 0, s, acc, B.qshadow,
                                   B.cnt + 3, B.cnt + 5, (const float4 *)B.sh_o, (const float4 *)B.sh_d,
                                   (int2 *)nullptr, B.occl, rc1);
            }
        }
        hipLaunchKernelGGL(k_wf_resolve, dim3(grid_blocks), dim3(256), 0, s, F, B, cur);
        hipLaunchKernelGGL(k_wf_advance, dim3(1), dim3(64), 0, s, B.cnt, cur, (unsigned long long *)nullptr);
    }
    PT_HIP(hipGetLastError());
    return PTGPU_OK;
}

void wf_release(ptgpu_ctx *ctx) {
    if (!ctx->wf_workspace) return;
    WfWorkspace *W = reinterpret_cast<WfWorkspace *>(ctx->wf_workspace);
    W->path.release();
    W->shadow.release();
    W->queues.release();
    W->counters.release();
    delete W;
    ctx->wf_workspace = nullptr;
}

}  // namespace ptgpu
