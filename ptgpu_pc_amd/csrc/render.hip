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

#include "device_util.hpp"
#include "lbvh.hpp"

namespace ptgpu {

constexpr float kSlack = 1.0f + 1.0e-5f;   // relative slack on the slab test's exit distance
constexpr float kInf = 1e20f;              // the reference's initial t (rendering_kernel.cl:228)
constexpr unsigned kMaxVisits = 1u << 24;  // node visits per query; only a cyclic tree reaches it

struct f3 { float x, y, z; };
__device__ __forceinline__ f3 mk(float x, float y, float z) { f3 v; v.x = x; v.y = y; v.z = z; return v; }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 mul3(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 smul(float k, f3 b) { return mk(k * b.x, k * b.y, k * b.z); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ f3 norm3(f3 v) { float l = 1.f / sqrtf(dot3(v, v)); return smul(l, v); }
__device__ __forceinline__ bool iszero3(f3 v) { return v.x == 0.f && v.y == 0.f && v.z == 0.f; }
__device__ __forceinline__ f3 ldvec(const Vec *p) {
    const float *f = reinterpret_cast<const float *>(p);
    return mk(f[0], f[1], f[2]);
}

// ---- GetRandom, rendering_kernel.cl:4-17 ----
__device__ __forceinline__ float get_random(unsigned &s0, unsigned &s1) {
    s0 = 36969u * (s0 & 65535u) + (s0 >> 16);
    s1 = 18000u * (s1 & 65535u) + (s1 >> 16);
    unsigned ires = (s0 << 16) + s1;
    float f = __uint_as_float((ires & 0x007fffffu) | 0x40000000u);
    return (f - 2.f) / 2.f;
}

// ---- primitive tests (rendering_kernel.cl:26-125) on the Prim layout ----
__device__ __forceinline__ float tri_isect(const float4 &p0, const float4 &p1, const float4 &p2, f3 o, f3 d) {
    const f3 v0 = mk(p0.x, p0.y, p0.z), e1 = mk(p1.x, p1.y, p1.z), e2 = mk(p2.x, p2.y, p2.z);
    const f3 pvec = cross3(d, e2);
    const float det = dot3(e1, pvec);
    if (det == 0.f) return 0.f;
    const float inv = 1.0f / det;
    const f3 tvec = sub3(o, v0);
    const float u = dot3(tvec, pvec) * inv;
    if (u < 0.0f || u > 1.0f) return 0.f;
    const f3 qvec = cross3(tvec, e1);
    const float v = dot3(d, qvec) * inv;
    if (v < 0.0f || u + v > 1.0f) return 0.f;
    const float t = dot3(e2, qvec) * inv;
    return t > kEpsilon ? t : 0.f;
}
__device__ __forceinline__ float sph_isect(const float4 &p0, float rad, f3 o, f3 d) {
    const f3 op = sub3(mk(p0.x, p0.y, p0.z), o);
    const float b = dot3(op, d);
    float det = b * b - dot3(op, op) + rad * rad;
    if (det < 0.f) return 0.f;
    det = sqrtf(det);
    float t = b - det;
    if (t > kEpsilon) return t;
    t = b + det;
    return t > kEpsilon ? t : 0.f;
}
__device__ __forceinline__ float prim_isect(const Prim *__restrict__ prims, int k, f3 o, f3 d, int *id) {
    const float4 p0 = prims[k].p0, p1 = prims[k].p1;
    *id = __float_as_int(p0.w);
    if (__float_as_int(p1.w) == SPHERE) return sph_isect(p0, p1.x, o, d);
    const float4 p2 = prims[k].p2;
    return tri_isect(p0, p1, p2, o, d);
}

// slab test of one box; returns entry distance (>= 0) or +inf on a miss
__device__ __forceinline__ float slab(float lx, float ly, float lz, float hx, float hy, float hz, f3 o, f3 inv,
                                      float tlimit) {
    const float tx0 = (lx - o.x) * inv.x, tx1 = (hx - o.x) * inv.x;
    const float ty0 = (ly - o.y) * inv.y, ty1 = (hy - o.y) * inv.y;
    const float tz0 = (lz - o.z) * inv.z, tz1 = (hz - o.z) * inv.z;
    const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.f));
    const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tlimit));
    return tn <= tf * kSlack ? tn : __builtin_inff();
}

// Closest hit over a prepared structure.  Returns the hit shape index or -1; t in *tout.
template <int STACK>
__device__ int trace_closest(const DevAccel &A, f3 o, f3 d, float *tout, int *stk, unsigned *nodes_visited,
                             unsigned *prims_tested) {
    float tb = kInf;
    int bid = -1;
    const f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    auto test_leaf = [&](int first, int cnt) {
        if (first < 0 || cnt > A.nprims - first) return;   // malformed caller tree
        for (int q = 0; q < cnt; q++) {
            int id;
            const float t = prim_isect(A.prims, first + q, o, d, &id);
            if (prims_tested) (*prims_tested)++;
            if (t != 0.f && (t < tb || (t == tb && id > bid))) { tb = t; bid = id; }
        }
    };
    if (A.root < 0) {
        test_leaf(~A.root, A.root_cnt);
        *tout = tb;
        return bid;
    }
    int node = A.root, sp = 0;
    for (unsigned guard = 0; guard < kMaxVisits; guard++) {
        if ((unsigned)node >= (unsigned)A.nnodes) break;   // malformed caller tree
        const TNode *np = A.nodes + node;
        const float4 a = np->a, b = np->b, c = np->c;
        const int4 r = np->r;
        if (nodes_visited) (*nodes_visited)++;
        float t0 = slab(a.x, a.y, a.z, a.w, b.x, b.y, o, inv, tb);
        float t1 = slab(b.z, b.w, c.x, c.y, c.z, c.w, o, inv, tb);
        if (r.x < 0 && t0 != __builtin_inff()) { test_leaf(~r.x, r.z); t0 = __builtin_inff(); }
        if (r.y < 0 && t1 != __builtin_inff()) { test_leaf(~r.y, r.w); t1 = __builtin_inff(); }
        const bool h0 = t0 != __builtin_inff(), h1 = t1 != __builtin_inff();
        if (h0 && h1) {
            const bool first0 = t0 <= t1;
            const int nearc = first0 ? r.x : r.y, farc = first0 ? r.y : r.x;
            if (sp < STACK) stk[(sp++) * 256] = farc;
            node = nearc;
        } else if (h0) {
            node = r.x;
        } else if (h1) {
            node = r.y;
        } else {
            if (sp == 0) break;
            node = stk[(--sp) * 256];
        }
    }
    *tout = tb;
    return bid;
}

// Any hit with 0 != t < maxt over shapes other than `skip_id` (IntersectP semantics).
template <int STACK>
__device__ bool trace_any(const DevAccel &A, f3 o, f3 d, float maxt, int skip_id, int *stk,
                          unsigned *nodes_visited, unsigned *prims_tested) {
    const f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    auto test_leaf = [&](int first, int cnt) -> bool {
        if (first < 0 || cnt > A.nprims - first) return false;   // malformed caller tree
        for (int q = 0; q < cnt; q++) {
            int id;
            const float t = prim_isect(A.prims, first + q, o, d, &id);
            if (prims_tested) (*prims_tested)++;
            if (t != 0.f && t < maxt && id != skip_id) return true;
        }
        return false;
    };
    if (A.root < 0) return test_leaf(~A.root, A.root_cnt);
    int node = A.root, sp = 0;
    for (unsigned guard = 0; guard < kMaxVisits; guard++) {
        if ((unsigned)node >= (unsigned)A.nnodes) break;   // malformed caller tree
        const TNode *np = A.nodes + node;
        const float4 a = np->a, b = np->b, c = np->c;
        const int4 r = np->r;
        if (nodes_visited) (*nodes_visited)++;
        float t0 = slab(a.x, a.y, a.z, a.w, b.x, b.y, o, inv, maxt);
        float t1 = slab(b.z, b.w, c.x, c.y, c.z, c.w, o, inv, maxt);
        if (r.x < 0 && t0 != __builtin_inff()) { if (test_leaf(~r.x, r.z)) return true; t0 = __builtin_inff(); }
        if (r.y < 0 && t1 != __builtin_inff()) { if (test_leaf(~r.y, r.w)) return true; t1 = __builtin_inff(); }
        const bool h0 = t0 != __builtin_inff(), h1 = t1 != __builtin_inff();
        if (h0 && h1) {
            if (sp < STACK) stk[(sp++) * 256] = r.y;
            node = r.x;
        } else if (h0) {
            node = r.x;
        } else if (h1) {
            node = r.y;
        } else {
            if (sp == 0) break;
            node = stk[(--sp) * 256];
        }
    }
    return false;
}

// Brute-force mode (ACCELSTR 0): prims in shape-index order.
__device__ int brute_closest(const Prim *prims, int n, f3 o, f3 d, float *tout, unsigned *prims_tested) {
    float tb = kInf;
    int bid = -1;
    for (int i = n - 1; i >= 0; i--) {
        int id;
        const float t = prim_isect(prims, i, o, d, &id);
        if (t != 0.f && t < tb) { tb = t; bid = id; }
    }
    if (prims_tested) *prims_tested += (unsigned)n;
    *tout = tb;
    return bid;
}
__device__ bool brute_any(const Prim *prims, int n, f3 o, f3 d, float maxt, unsigned *prims_tested) {
    for (int i = n - 2; i >= 0; i--) {
        int id;
        if (prims_tested) (*prims_tested)++;
        const float t = prim_isect(prims, i, o, d, &id);
        if (t != 0.f && t < maxt) return true;
    }
    return false;
}

template <int STACK>
__device__ __forceinline__ int closest(const DevAccel &A, f3 o, f3 d, float *t, int *stk, unsigned *nv, unsigned *pt) {
    if (A.mode == PTGPU_ACCEL_NONE) return brute_closest(A.prims, A.nprims, o, d, t, pt);
    return trace_closest<STACK>(A, o, d, t, stk, nv, pt);
}
template <int STACK>
__device__ __forceinline__ bool anyhit(const DevAccel &A, f3 o, f3 d, float maxt, int *stk, unsigned *nv, unsigned *pt) {
    if (A.mode == PTGPU_ACCEL_NONE) return brute_any(A.prims, A.nprims, o, d, maxt, pt);
    return trace_any<STACK>(A, o, d, maxt, A.nshapes - 1, stk, nv, pt);
}

// ---- camera ray, rendering_kernel.cl:1042-1070 ----
__device__ __forceinline__ void camera_ray(const Camera *cam, unsigned &s0, unsigned &s1, int width, int height,
                                           int x, int y, f3 &ro, f3 &rd) {
    const float invW = 1.f / (float)width, invH = 1.f / (float)height;
    const float r1 = get_random(s0, s1) - .5f;
    const float r2 = get_random(s0, s1) - .5f;
    const float kcx = ((float)x + r1) * invW - .5f;
    const float kcy = ((float)y + r2) * invH - .5f;
    const f3 cx = ldvec(&cam->x), cy = ldvec(&cam->y), cd = ldvec(&cam->dir), co = ldvec(&cam->orig);
    f3 rdir = mk(cx.x * kcx + cy.x * kcy + cd.x, cx.y * kcx + cy.y * kcy + cd.y, cx.z * kcx + cy.z * kcy + cd.z);
    ro = add3(smul(0.1f, rdir), co);
    rd = norm3(rdir);
}

__device__ __forceinline__ int to_int(float x) {   // vec.h:60 with the OpenCL clamp builtin
    return (int)(powf(fminf(fmaxf(x, 0.f), 1.f), 1.f / 2.2f) * 255.f + .5f);
}

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
};

// The path-tracing megakernel.  Lane state: one pixel (gid), its remaining samples, the RNG state
// and running mean; path state: ray, radiance, throughput, specularBounce, depth.
template <int STACK, bool STATS>
__global__ __launch_bounds__(256) void k_render(RenderParams P) {
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
                    if (my < (unsigned)P.total) {
                        active = true;
                        gid = (int)my;
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
                const Shape *sh = P.shapes + id;
                const float *w = reinterpret_cast<const float *>(sh);
                const int type = reinterpret_cast<const int *>(sh)[0];
                f3 nrm;
                if (type == SPHERE) {
                    nrm = sub3(hp, mk(w[2], w[3], w[4]));
                } else {
                    const f3 p1 = mk(w[1], w[2], w[3]);
                    nrm = cross3(sub3(mk(w[4], w[5], w[6]), p1), sub3(mk(w[7], w[8], w[9]), p1));
                }
                nrm = norm3(nrm);
                const float dp = dot3(nrm, rd);
                const float sgn = dp > 0.f ? 1.f : (dp < 0.f ? -1.f : (dp == dp ? dp : 0.f));  // OpenCL sign
                const f3 nl = smul(-1.f * sgn, nrm);
                const f3 e = mk(w[10], w[11], w[12]);
                const f3 c = mk(w[13], w[14], w[15]);
                const int refl = reinterpret_cast<const int *>(sh)[16];
                if (!iszero3(e)) {
                    if (spec) rad = add3(rad, mul3(thr, smul(fabsf(dp), e)));
                    ended = true;
                } else if (refl == DIFF) {
                    spec = 0;
                    thr = mul3(thr, c);
                    // SampleLights
                    f3 Ld = mk(0.f, 0.f, 0.f);
                    for (int li = 0; li < P.nlights; li++) {
                        const Shape *L = P.shapes + P.lights[li];
                        const float *lw = reinterpret_cast<const float *>(L);
                        const int ltype = reinterpret_cast<const int *>(L)[0];
                        f3 lp, ln;
                        if (ltype == TRIANGLE) {
                            const float u1 = get_random(s0, s1);
                            const float u2 = get_random(s0, s1);
                            const float su1 = sqrtf(u1);
                            const float bu = 1.0f - su1, bv = u2 * su1;
                            const f3 p1 = mk(lw[1], lw[2], lw[3]);
                            f3 e1 = sub3(mk(lw[4], lw[5], lw[6]), p1), e2 = sub3(mk(lw[7], lw[8], lw[9]), p1);
                            e1 = smul(bu, e1);
                            e2 = smul(bv, e2);
                            lp = add3(add3(p1, e1), e2);
                            ln = norm3(cross3(e1, e2));
                        } else {
                            const float u1 = get_random(s0, s1);
                            const float u2 = get_random(s0, s1);
                            const float zz = 1.f - 2.f * u1;
                            const float q = 1.f - zz * zz;
                            const float rr = sqrtf(0.f < q ? q : 0.f);
                            const float phi = 2.f * kPi * u2;
                            const f3 usp = mk(rr * cosf(phi), rr * sinf(phi), zz);
                            lp = add3(smul(lw[1], usp), mk(lw[2], lw[3], lw[4]));
                            ln = usp;
                        }
                        f3 sd = sub3(lp, hp);
                        const float len = sqrtf(dot3(sd, sd));
                        sd = smul(1.f / len, sd);
                        float wo = dot3(sd, ln);
                        if (wo > 0.f) continue;
                        wo = -wo;
                        wo = fminf(fmaxf(wo, 0.f), 1.f);
                        const float wi = dot3(sd, nl);
                        if (wi > 0.f) {
                            n_any++;
                            if (!anyhit<STACK>(A, hp, sd, len - kEpsilon, stk, pnv, ppt)) {
                                const float s = lw[25] * wi * wo / (len * len);
                                Ld = add3(Ld, smul(s, mk(lw[10], lw[11], lw[12])));
                            }
                        }
                    }
                    Ld = mul3(thr, Ld);
                    rad = add3(rad, Ld);
                    // cosine-weighted bounce
                    const float r1 = 2.f * kPi * get_random(s0, s1);
                    const float r2 = get_random(s0, s1);
                    const float r2s = sqrtf(r2);
                    f3 ww = nl;
                    const f3 aa = fabsf(ww.x) > .1f ? mk(0.f, 1.f, 0.f) : mk(1.f, 0.f, 0.f);
                    f3 u = norm3(cross3(aa, ww));
                    f3 v = cross3(ww, u);
                    u = smul(cosf(r1) * r2s, u);
                    v = smul(sinf(r1) * r2s, v);
                    f3 nd = add3(u, v);
                    ww = smul(sqrtf(1 - r2), ww);
                    nd = add3(nd, ww);
                    ro = hp;
                    rd = nd;
                } else if (refl == SPEC) {
                    spec = 1;
                    f3 nd = smul(2.f * dot3(nrm, rd), nrm);
                    nd = sub3(rd, nd);
                    thr = mul3(thr, c);
                    ro = hp;
                    rd = nd;
                } else {
                    spec = 1;
                    f3 nd = smul(2.f * dot3(nrm, rd), nrm);
                    nd = sub3(rd, nd);
                    const bool into = dot3(nrm, nl) > 0;
                    const float nc = 1.f, nt = 1.5f;
                    const float nnt = into ? nc / nt : nt / nc;
                    const float ddn = dot3(rd, nl);
                    const float cos2t = 1.f - nnt * nnt * (1.f - ddn * ddn);
                    if (cos2t < 0.f) {
                        thr = mul3(thr, c);
                        ro = hp;
                        rd = nd;
                    } else {
                        const float kk = (float)(into ? 1 : -1) * (ddn * nnt + sqrtf(cos2t));
                        const f3 nkk = smul(kk, nrm);
                        f3 td = smul(nnt, rd);
                        td = sub3(td, nkk);
                        td = norm3(td);
                        const float a = nt - nc, b = nt + nc;
                        const float R0 = a * a / (b * b);
                        const float cc = 1 - (into ? -ddn : dot3(td, nrm));
                        const float Re = R0 + (1 - R0) * cc * cc * cc * cc * cc;
                        const float Tr = 1.f - Re;
                        const float Pp = .25f + .5f * Re;
                        const float RP = Re / Pp;
                        const float TP = Tr / (1.f - Pp);
                        if (get_random(s0, s1) < Pp) {
                            thr = smul(RP, thr);
                            thr = mul3(thr, c);
                            ro = hp;
                            rd = nd;
                        } else {
                            thr = smul(TP, thr);
                            thr = mul3(thr, c);
                            ro = hp;
                            rd = td;
                        }
                    }
                }
                depth++;
            }
        }
        if (ended) {
            // RadianceGPU accumulation, rendering_kernel.cl:1127-1136
            const int cs = P.currentSample + k;
            if (cs == 0) {
                col = rad;
            } else {
                const float k1 = (float)cs;
                const float k2 = 1.f / ((float)cs + 1.f);
                col = mk((col.x * k1 + rad.x) * k2, (col.y * k1 + rad.y) * k2, (col.z * k1 + rad.z) * k2);
            }
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
                    P.pixels[y * P.width + x] = (unsigned)to_int(col.x) | ((unsigned)to_int(col.y) << 8) |
                                                ((unsigned)to_int(col.z) << 16) | 0xff000000u;
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
__device__ __forceinline__ Prim prim_of(const Shape *s, int id) {
    const float *w = reinterpret_cast<const float *>(s);
    const int type = reinterpret_cast<const int *>(s)[0];
    Prim p;
    if (type == SPHERE) {
        p.p0 = make_float4(w[2], w[3], w[4], __int_as_float(id));
        p.p1 = make_float4(w[1], 0.f, 0.f, __int_as_float(type));
        p.p2 = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
        const float x1 = w[1], y1 = w[2], z1 = w[3];
        p.p0 = make_float4(x1, y1, z1, __int_as_float(id));
        p.p1 = make_float4(w[4] - x1, w[5] - y1, w[6] - z1, __int_as_float(type));
        p.p2 = make_float4(w[7] - x1, w[8] - y1, w[9] - z1, 0.f);
    }
    return p;
}

__global__ void k_index_prims(const Shape *shapes, int n, Prim *prims) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) prims[i] = prim_of(shapes + i, i);
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

// reference BVH layout -> TNode/Prim (child is a leaf iff nLeft == min / nRight == max)
__global__ void k_convert_bvh(const Shape *shapes, int n, const BVHNodeGPU *btn, const BVHNodeGPU *btl,
                              TNode *nodes, Prim *prims) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) prims[i] = prim_of(shapes + btl[i].nShape, btl[i].nShape);
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

// reference KD heap layout (1-based, children 2i/2i+1, leaves own kn[min..max)) -> TNode/Prim
__global__ void k_convert_kd(const Shape *shapes, const KDNodeGPU *kng, int kngCnt, const int *kn, int knCnt,
                             TNode *nodes, Prim *prims) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h < knCnt) prims[h] = prim_of(shapes + kn[h], kn[h]);
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
int build_index_prims(ptgpu_ctx *ctx, const Shape *d_shapes, int n, DevBuf &out_prims) {
    PT_TRY(out_prims.ensure((size_t)n * sizeof(Prim)));
    hipLaunchKernelGGL(k_index_prims, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, d_shapes, n,
                       out_prims.as<Prim>());
    PT_HIP(hipGetLastError());
    return PTGPU_OK;
}

int build_light_list(ptgpu_ctx *ctx, const Shape *d_shapes, int n, int light_cnt, DevBuf &out, DevBuf &scan,
                     int *found) {
    *found = 0;
    if (light_cnt <= 0 || n <= 0) {
        PT_TRY(out.ensure(4));
        return PTGPU_OK;
    }
    const int nb = (n + kLightChunk - 1) / kLightChunk;
    PT_TRY(scan.ensure(((size_t)nb + 4) * 4));
    PT_TRY(out.ensure((size_t)light_cnt * 4));
    unsigned *counts = scan.as<unsigned>();
    unsigned *total = counts + nb;
    hipLaunchKernelGGL(k_light_count, dim3(nb), dim3(256), 0, ctx->stream, d_shapes, n, counts);
    PT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_light_scan, dim3(1), dim3(64), 0, ctx->stream, counts, nb, total);
    PT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_light_write, dim3(nb), dim3(256), 0, ctx->stream, d_shapes, n, counts, light_cnt,
                       out.as<int>());
    PT_HIP(hipGetLastError());
    unsigned h_total = 0;
    PT_HIP(hipMemcpyAsync(&h_total, total, 4, hipMemcpyDeviceToHost, ctx->stream));
    PT_HIP(hipStreamSynchronize(ctx->stream));
    *found = (int)h_total < light_cnt ? (int)h_total : light_cnt;
    return PTGPU_OK;
}

int convert_bvh_ref(ptgpu_ctx *ctx, const Shape *d_shapes, int n, const BVHNodeGPU *d_btn, const BVHNodeGPU *d_btl,
                    DevBuf &nodes, DevBuf &prims, DevAccel *acc) {
    PT_TRY(nodes.ensure((size_t)(n > 1 ? n - 1 : 1) * sizeof(TNode)));
    PT_TRY(prims.ensure((size_t)n * sizeof(Prim)));
    hipLaunchKernelGGL(k_convert_bvh, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, d_shapes, n, d_btn, d_btl,
                       nodes.as<TNode>(), prims.as<Prim>());
    PT_HIP(hipGetLastError());
    acc->nodes = nodes.as<TNode>();
    acc->prims = prims.as<Prim>();
    acc->nprims = n;
    acc->nnodes = n - 1;
    acc->nshapes = n;
    acc->mode = PTGPU_ACCEL_BVH;
    if (n >= 2) {
        acc->root = 0;
        acc->root_cnt = 0;
        int d = 0;
        while ((1ll << d) < n) d++;
        acc->depth = d;
    } else {
        acc->root = ~0;
        acc->root_cnt = 1;
        acc->depth = 0;
    }
    return PTGPU_OK;
}

int convert_kd_ref(ptgpu_ctx *ctx, const Shape *d_shapes, int n, const KDNodeGPU *d_kng, int kngCnt, const int *d_kn,
                   int knCnt, DevBuf &nodes, DevBuf &prims, DevAccel *acc) {
    PT_CHECK(kngCnt >= 2, PTGPU_ERR_INVALID, "kngCnt %d < 2", kngCnt);
    PT_TRY(nodes.ensure((size_t)kngCnt * sizeof(TNode)));
    PT_TRY(prims.ensure((size_t)(knCnt > 0 ? knCnt : 1) * sizeof(Prim)));
    const int m = kngCnt > knCnt ? kngCnt : knCnt;
    hipLaunchKernelGGL(k_convert_kd, dim3((m + 255) / 256), dim3(256), 0, ctx->stream, d_shapes, d_kng, kngCnt, d_kn,
                       knCnt, nodes.as<TNode>(), prims.as<Prim>());
    PT_HIP(hipGetLastError());
    KDNodeGPU root;
    PT_HIP(hipMemcpyAsync(&root, d_kng + 1, sizeof(KDNodeGPU), hipMemcpyDeviceToHost, ctx->stream));
    PT_HIP(hipStreamSynchronize(ctx->stream));
    acc->nodes = nodes.as<TNode>();
    acc->prims = prims.as<Prim>();
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

static int stack_for(const DevAccel &acc) {
    if (acc.mode == PTGPU_ACCEL_NONE) return 24;
    if (acc.depth + 2 <= 24) return 24;
    if (acc.depth + 2 <= 64) return 64;
    return -1;
}

int launch_render(ptgpu_ctx *ctx, const DevAccel &acc, const RenderLaunch &L) {
    const int total = L.width * L.height;
    if (total <= 0) return PTGPU_OK;
    const int st = stack_for(acc);
    PT_CHECK(st > 0, PTGPU_ERR_UNSUPPORTED, "tree depth %d exceeds the 62-level traversal stack", acc.depth);
    PT_TRY(ctx->tmp_counter.ensure(64));
    PT_TRY(fill_u32(ctx->stream, ctx->tmp_counter.p, 0u, 1));
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
    P.work_counter = ctx->tmp_counter.as<unsigned>();
    P.total = total;
    // persistent waves: enough workgroups to fill every CU several times, never more than needed
    int blocks = ctx->cu_count * 8;
    const int need = (total + 255) / 256;
    if (blocks > need) blocks = need;
    if (blocks < 1) blocks = 1;
    if (L.trav_stats) {
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
    const int st = stack_for(acc);
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
                         DevAccel *acc);

extern "C" int ptgpu_kdtree_device_accel(ptgpu_kdtree *kd, ptgpu::DevAccel *acc);
namespace ptgpu { int clbvh_prepare_traversal(ptgpu_clbvh *h); }

static int resolve_accel(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const Shape *d_shapes, unsigned n,
                         DevAccel *acc) {
    PT_CHECK(!(bvh && kd), PTGPU_ERR_INVALID, "pass at most one of bvh / kd");
    if (bvh) {
        PT_CHECK(bvh->state >= 3, PTGPU_ERR_INVALID, "BVH not built (call build_bvh_tree)");
        PT_CHECK((unsigned)bvh->n == n, PTGPU_ERR_INVALID, "shapeCnt %u differs from the BVH's %d", n, bvh->n);
        PT_TRY(clbvh_prepare_traversal(bvh));
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

extern "C" {

int ptgpu_radiance_gpu(ptgpu_ctx *ctx, int accel, const ptgpu_radiance_args *a) {
    PT_CHECK(ctx && a, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(a->colors && a->seeds && a->camera && a->shapes && a->pixels, PTGPU_ERR_INVALID, "null buffer");
    PT_CHECK(a->shapeCnt >= 1, PTGPU_ERR_INVALID, "shapeCnt is 0");
    PT_CHECK(a->width > 0 && a->height > 0, PTGPU_ERR_INVALID, "bad frame size %dx%d", a->width, a->height);
    PT_CHECK(a->currentSample >= 0, PTGPU_ERR_INVALID, "currentSample < 0");
    PT_HIP(hipSetDevice(ctx->device));
    const int n = (int)a->shapeCnt;
    DevAccel acc;
    if (accel == PTGPU_ACCEL_BVH) {
        PT_CHECK(a->btn && a->btl, PTGPU_ERR_INVALID, "BVH mode needs btn and btl");
        PT_TRY(convert_bvh_ref(ctx, a->shapes, n, a->btn, a->btl, ctx->tmp_nodes, ctx->tmp_prims, &acc));
    } else if (accel == PTGPU_ACCEL_KD) {
        PT_CHECK(a->kng && a->kn, PTGPU_ERR_INVALID, "KD mode needs kng and kn");
        PT_TRY(convert_kd_ref(ctx, a->shapes, n, a->kng, a->kngCnt, a->kn, a->knCnt, ctx->tmp_nodes, ctx->tmp_prims,
                              &acc));
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
    return launch_render(ctx, acc, L);
}

int ptgpu_render(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const ptgpu_render_args *a) {
    PT_CHECK(ctx && a, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(a->colors && a->seeds && a->camera && a->shapes, PTGPU_ERR_INVALID, "null buffer");
    PT_CHECK(a->shapeCnt >= 1, PTGPU_ERR_INVALID, "shapeCnt is 0");
    PT_CHECK(a->width > 0 && a->height > 0, PTGPU_ERR_INVALID, "bad frame size %dx%d", a->width, a->height);
    PT_CHECK(a->spp >= 1 && a->currentSample >= 0, PTGPU_ERR_INVALID, "bad spp/currentSample");
    PT_CHECK(a->max_depth >= 0 && a->max_depth < 1024, PTGPU_ERR_INVALID, "bad max_depth");
    PT_HIP(hipSetDevice(ctx->device));
    DevAccel acc;
    PT_TRY(resolve_accel(ctx, bvh, kd, a->shapes, a->shapeCnt, &acc));
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
    } else {
        PT_TRY(build_light_list(ctx, a->shapes, (int)a->shapeCnt, (int)a->lightCnt, ctx->tmp_lights, ctx->tmp_scan,
                                &nl));
        lights = ctx->tmp_lights.as<int>();
    }
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
    L.spp = a->spp;
    L.max_depth = a->max_depth;
    L.pixels = a->pixels;
    L.ray_counter = a->ray_counter;
    L.trav_stats = a->trav_stats;
    return launch_render(ctx, acc, L);
}

int ptgpu_trace_closest(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const Shape *d_shapes,
                        unsigned int shapeCnt, const Ray *d_rays, int n, float *d_t, int *d_id) {
    PT_CHECK(ctx && d_shapes && d_rays && d_t && d_id, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(shapeCnt >= 1 && n >= 0, PTGPU_ERR_INVALID, "bad sizes");
    PT_HIP(hipSetDevice(ctx->device));
    DevAccel acc;
    PT_TRY(resolve_accel(ctx, bvh, kd, d_shapes, shapeCnt, &acc));
    return launch_trace(ctx, acc, shapeCnt, d_rays, nullptr, n, d_t, d_id, nullptr, false);
}

int ptgpu_trace_any(ptgpu_ctx *ctx, ptgpu_clbvh *bvh, ptgpu_kdtree *kd, const Shape *d_shapes, unsigned int shapeCnt,
                    const Ray *d_rays, const float *d_maxt, int n, int *d_hit) {
    PT_CHECK(ctx && d_shapes && d_rays && d_maxt && d_hit, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(shapeCnt >= 1 && n >= 0, PTGPU_ERR_INVALID, "bad sizes");
    PT_HIP(hipSetDevice(ctx->device));
    DevAccel acc;
    PT_TRY(resolve_accel(ctx, bvh, kd, d_shapes, shapeCnt, &acc));
    return launch_trace(ctx, acc, shapeCnt, d_rays, d_maxt, n, nullptr, nullptr, d_hit, true);
}

}  // extern "C"
