// render_device.hpp -- device math shared by the path-tracing engines (megakernel and wavefront).
//
// Every function restates a piece of rendering_kernel.cl in FP32 with the expression order of the
// reference (the library is built with -ffp-contract=off and IEEE division/sqrt), so both engines
// produce the same bits for the same path.
#pragma once

#include "device_util.hpp"

// The per-path math below is also compiled for the host: the hybrid CPU/GPU renderer runs it on host
// cores for its CPU tiles, so those tiles get the kernel's bits.
#define PT_HD __host__ __device__ __forceinline__

namespace ptgpu {

constexpr float kSlack = 1.0f + 1.0e-5f;   // relative slack on the slab test's exit distance
constexpr float kInf = 1e20f;              // the reference's initial t (rendering_kernel.cl:228)
constexpr unsigned kMaxVisits = 1u << 24;  // node visits per query; only a cyclic tree reaches it

struct f3 { float x, y, z; };
PT_HD f3 mk(float x, float y, float z) { f3 v; v.x = x; v.y = y; v.z = z; return v; }
PT_HD f3 sub3(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
PT_HD f3 add3(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
PT_HD f3 mul3(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
PT_HD f3 smul(float k, f3 b) { return mk(k * b.x, k * b.y, k * b.z); }
PT_HD float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PT_HD f3 cross3(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
PT_HD f3 norm3(f3 v) { float l = 1.f / sqrtf(dot3(v, v)); return smul(l, v); }
PT_HD bool iszero3(f3 v) { return v.x == 0.f && v.y == 0.f && v.z == 0.f; }
PT_HD f3 ldvec(const Vec *p) {
    const float *f = reinterpret_cast<const float *>(p);
    return mk(f[0], f[1], f[2]);
}

// ---- sin and cos of the sampling angles (rendering_kernel.cl:131-132, 652-654) ----
// The reference calls OpenCL's sin/cos, whose results the OpenCL specification bounds (4 ulp) but
// does not fix.  This library evaluates them one documented way, in binary32: Cody-Waite reduction
// of x by k = rint(x * 2/pi) multiples of pi/2 split in three parts (P1 and P2 with few enough
// significant bits that k * P1 and k * P2 are exact), then the Taylor series of sin (to r^9) and cos
// (to r^10) on [-pi/4, pi/4] in Horner form; the truncation error there is below 0.05 ulp and the
// whole evaluation stays within about one ulp.  The same IEEE binary32 operations in the same order
// give the same bits on any conforming implementation (the CPU restatement evaluates them so too).
PT_HD void sincos_f32(float x, float *s, float *c) {
    const float k = rintf(x * 0.6366197466850281f);                        // 2/pi
    const float r = ((x - k * 1.5703125f) - k * 4.837512969970703e-4f) - k * 7.549790126404332e-8f;
    const float r2 = r * r;
    const float sp =
        r + r * r2 *
                (-0.1666666716337204f +
                 r2 * (0.008333333767950535f + r2 * (-0.00019841270113829523f + r2 * 2.7557318844628753e-06f)));
    const float cp =
        1.0f + r2 * (-0.5f + r2 * (0.0416666679084301f +
                                   r2 * (-0.0013888889225199819f +
                                         r2 * (2.4801587642286904e-05f + r2 * -2.755731998149713e-07f))));
    const int q = (int)k & 3;                                              // quadrant
    const float sv = (q & 1) ? cp : sp, cv = (q & 1) ? sp : cp;
    *s = (q & 2) ? -sv : sv;
    *c = ((q + 1) & 2) ? -cv : cv;
}

// The traversal record of shape `s` (index id): v0 and the FP32 edges e1 = v1 - v0, e2 = v2 - v0 the
// reference forms in its triangle test, or the sphere's centre and radius; `flags` (kLastInLeaf)
// joins the type word.
PT_HD Prim prim_of(const Shape *s, int id, int flags) {
    const float *w = reinterpret_cast<const float *>(s);
    const int type = reinterpret_cast<const int *>(s)[0] | flags;
    const float fid = __builtin_bit_cast(float, id), ftype = __builtin_bit_cast(float, type);
    Prim p;
    if ((type & 0xff) == SPHERE) {
        p.p0 = make_float4(w[2], w[3], w[4], fid);
        p.p1 = make_float4(w[1], 0.f, 0.f, ftype);
        p.p2 = make_float4(0.f, 0.f, 0.f, ftype);
    } else {
        const float x1 = w[1], y1 = w[2], z1 = w[3];
        p.p0 = make_float4(x1, y1, z1, fid);
        p.p1 = make_float4(w[4] - x1, w[5] - y1, w[6] - z1, ftype);
        p.p2 = make_float4(w[7] - x1, w[8] - y1, w[9] - z1, ftype);
    }
    return p;
}

// ---- GetRandom, rendering_kernel.cl:4-17 ----
PT_HD float get_random(unsigned &s0, unsigned &s1) {
    s0 = 36969u * (s0 & 65535u) + (s0 >> 16);
    s1 = 18000u * (s1 & 65535u) + (s1 >> 16);
    unsigned ires = (s0 << 16) + s1;
    float f = __builtin_bit_cast(float, (ires & 0x007fffffu) | 0x40000000u);
    return (f - 2.f) / 2.f;
}

// ---- primitive tests (rendering_kernel.cl:26-125) on the Prim layout ----
PT_HD float tri_isect(const float4 &p0, const float4 &p1, const float4 &p2, f3 o, f3 d) {
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
PT_HD float sph_isect(const float4 &p0, float rad, f3 o, f3 d) {
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
PT_HD float prim_isect(const Prim *__restrict__ prims, int k, f3 o, f3 d, int *id) {
    const float4 p0 = prims[k].p0, p1 = prims[k].p1, p2 = prims[k].p2;
    *id = __builtin_bit_cast(int, p0.w);
    if ((__builtin_bit_cast(int, p2.w) & 0xff) == SPHERE) return sph_isect(p0, p1.x, o, d);
    return tri_isect(p0, p1, p2, o, d);
}

// prim_isect that also reports whether primitive k ends its leaf run (kLastInLeaf)
__device__ __forceinline__ float prim_isect_vals(const float4 &p0, const float4 &p1, const float4 &p2, f3 o, f3 d,
                                                 int *id, bool *last) {
    *id = __float_as_int(p0.w);
    *last = (__float_as_int(p1.w) & kLastInLeaf) != 0;
    if ((__float_as_int(p2.w) & 0xff) == SPHERE) return sph_isect(p0, p1.x, o, d);
    return tri_isect(p0, p1, p2, o, d);
}
__device__ __forceinline__ float prim_isect_run(const Prim *__restrict__ prims, int k, f3 o, f3 d, int *id,
                                                bool *last) {
    const float4 p0 = prims[k].p0, p1 = prims[k].p1, p2 = prims[k].p2;
    return prim_isect_vals(p0, p1, p2, o, d, id, last);
}

// the two binary16 halves of a TNode4H word, widened exactly to fp32
__device__ __forceinline__ float half_lo(unsigned w) {
    return (float)__builtin_bit_cast(_Float16, (unsigned short)(w & 0xffffu));
}
__device__ __forceinline__ float half_hi(unsigned w) { return (float)__builtin_bit_cast(_Float16, (unsigned short)(w >> 16)); }

// Lane exchange inside a quad (four consecutive lanes) with DPP quad_perm; every lane of the quad
// must be active.  qrot<K>: the value of lane (i + K) mod 4; qxor<K>: the value of lane i ^ K.
template <int K>
__device__ __forceinline__ int qrot(int v) {
    static_assert(K >= 1 && K <= 3, "qrot");
    constexpr int ctrl = K == 1 ? 0x39 : K == 2 ? 0x4E : 0x93;   // quad_perm [1,2,3,0] [2,3,0,1] [3,0,1,2]
    return __builtin_amdgcn_mov_dpp(v, ctrl, 0xF, 0xF, false);
}
template <int K>
__device__ __forceinline__ float qrotf(float v) { return __int_as_float(qrot<K>(__float_as_int(v))); }
template <int K>
__device__ __forceinline__ int qxor(int v) {
    static_assert(K == 1 || K == 2, "qxor");
    constexpr int ctrl = K == 1 ? 0xB1 : 0x4E;   // quad_perm [1,0,3,2] and [2,3,0,1]
    return __builtin_amdgcn_mov_dpp(v, ctrl, 0xF, 0xF, false);
}
template <int K>
__device__ __forceinline__ float qxorf(float v) { return __int_as_float(qxor<K>(__float_as_int(v))); }

// slab test of one box; returns entry distance (>= 0) or +inf on a miss
PT_HD float slab(float lx, float ly, float lz, float hx, float hy, float hz, f3 o, f3 inv,
                                      float tlimit) {
    const float tx0 = (lx - o.x) * inv.x, tx1 = (hx - o.x) * inv.x;
    const float ty0 = (ly - o.y) * inv.y, ty1 = (hy - o.y) * inv.y;
    const float tz0 = (lz - o.z) * inv.z, tz1 = (hz - o.z) * inv.z;
    const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.f));
    const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tlimit));
    return tn <= tf * kSlack ? tn : __builtin_inff();
}

// The same test on a binary16 box: each word holds a bound pair, lo in its low half and hi in its
// high half.  Each difference `bound - o` is one mixed-precision fma, fma(bound, one, -o) with the
// binary16 bound read from its half of the word: the bound converts to binary32 exactly and the
// product by one is exact, so the result is the rounded difference, bit for bit what a conversion
// and a subtraction give.  `one` is 1.0f held where the compiler cannot see it (a literal would be
// folded into that conversion and subtraction, two instructions instead of one).
//
// A ray's side of the test, set once per query (RaySlab): per axis, the rotation that brings the
// near bound of a bound word into its low half (0 when the reciprocal direction is >= 0, where lo is
// nearer; 16 otherwise), and the reciprocal for the far bound scaled by the slack.  Choosing near and
// far by the direction's sign gives the values a min and a max of the two distances would (rounding
// is monotone), with no min/max pair per axis.  The slack applied to the far reciprocal and to the
// limit, rather than to the exit distance as slab() does, differs by rounding far below the slack
// itself, so the test stays conservative.  Where a distance is NaN (a ray in a bound's plane with an
// infinite reciprocal) the test ignores that axis's bound, which only accepts more boxes.
struct RaySlab {
    unsigned rx, ry, rz;
    f3 inv_far;
};
__device__ __forceinline__ RaySlab ray_slab(f3 inv) {
    RaySlab r;
    r.rx = inv.x >= 0.f ? 0u : 16u;
    r.ry = inv.y >= 0.f ? 0u : 16u;
    r.rz = inv.z >= 0.f ? 0u : 16u;
    r.inv_far = mk(inv.x * kSlack, inv.y * kSlack, inv.z * kSlack);
    return r;
}

// tlimit_s: the query's limit times kSlack
__device__ __forceinline__ float slab_hn(unsigned wx, unsigned wy, unsigned wz, f3 o, f3 inv, const RaySlab &rs,
                                         float tlimit_s, float one) {
    const unsigned nx = __builtin_amdgcn_alignbit(wx, wx, rs.rx);
    const unsigned ny = __builtin_amdgcn_alignbit(wy, wy, rs.ry);
    const unsigned nz = __builtin_amdgcn_alignbit(wz, wz, rs.rz);
    const float tnx = __builtin_fmaf(half_lo(nx), one, -o.x) * inv.x;
    const float tfx = __builtin_fmaf(half_hi(nx), one, -o.x) * rs.inv_far.x;
    const float tny = __builtin_fmaf(half_lo(ny), one, -o.y) * inv.y;
    const float tfy = __builtin_fmaf(half_hi(ny), one, -o.y) * rs.inv_far.y;
    const float tnz = __builtin_fmaf(half_lo(nz), one, -o.z) * inv.z;
    const float tfz = __builtin_fmaf(half_hi(nz), one, -o.z) * rs.inv_far.z;
    const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.f));
    const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tlimit_s));
    return tn <= tf ? tn : __builtin_inff();
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
        if (nodes_visited) (*nodes_visited) += 2;   // box tests
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
        if (nodes_visited) (*nodes_visited) += 2;   // box tests
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
PT_HD void camera_ray(const Camera *cam, unsigned &s0, unsigned &s1, int width, int height,
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

PT_HD int to_int(float x) {   // vec.h:60 with the OpenCL clamp builtin
    return (int)(powf(fminf(fmaxf(x, 0.f), 1.f), 1.f / 2.2f) * 255.f + .5f);
}


// ---- surface at a hit (rendering_kernel.cl:550-598) ----
struct Surface {
    f3 nrm, nl, e, c;
    float dp;
    int refl;
};
PT_HD Surface surface_at(const Shape *sh, f3 hp, f3 rd) {
    // words 0-16 loaded together before the type is looked at: one memory round trip, where loads
    // placed in the type's branches would wait for the type word first
    float w[17];
    const float *src = reinterpret_cast<const float *>(sh);
#pragma unroll
    for (int i = 0; i < 17; i++) w[i] = src[i];
    asm volatile("" ::: "memory");   // compiler-only: keeps the loads here, ahead of the branches
    const int type = __builtin_bit_cast(int, w[0]);
    Surface S;
    f3 nrm;
    if (type == SPHERE) {
        nrm = sub3(hp, mk(w[2], w[3], w[4]));
    } else {
        const f3 p1 = mk(w[1], w[2], w[3]);
        nrm = cross3(sub3(mk(w[4], w[5], w[6]), p1), sub3(mk(w[7], w[8], w[9]), p1));
    }
    nrm = norm3(nrm);
    S.nrm = nrm;
    S.dp = dot3(nrm, rd);
    const float sgn = S.dp > 0.f ? 1.f : (S.dp < 0.f ? -1.f : (S.dp == S.dp ? S.dp : 0.f));  // OpenCL sign
    S.nl = smul(-1.f * sgn, nrm);
    S.e = mk(w[10], w[11], w[12]);
    S.c = mk(w[13], w[14], w[15]);
    S.refl = __builtin_bit_cast(int, w[16]);
    return S;
}

// ---- one light of SampleLights (rendering_kernel.cl:421-490): draws two randoms (u1 first),
// returns true when a shadow ray must decide the contribution s * e; sd/len describe that ray.
// lw: words 0-12 of the light's Shape (type, geometry, emission); area: its word 25
PT_HD bool sample_light_words(const float *lw, float area, unsigned &s0, unsigned &s1, f3 hp,
                                                   f3 nl, f3 *sd_out, float *len_out, f3 *contrib) {
    const int ltype = __builtin_bit_cast(int, lw[0]);
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
        float sphi, cphi;
        sincos_f32(phi, &sphi, &cphi);
        const f3 usp = mk(rr * cphi, rr * sphi, zz);
        lp = add3(smul(lw[1], usp), mk(lw[2], lw[3], lw[4]));
        ln = usp;
    }
    f3 sd = sub3(lp, hp);
    const float len = sqrtf(dot3(sd, sd));
    sd = smul(1.f / len, sd);
    float wo = dot3(sd, ln);
    if (wo > 0.f) return false;
    wo = -wo;
    wo = fminf(fmaxf(wo, 0.f), 1.f);
    const float wi = dot3(sd, nl);
    if (!(wi > 0.f)) return false;
    const float s = area * wi * wo / (len * len);
    *sd_out = sd;
    *len_out = len;
    *contrib = smul(s, mk(lw[10], lw[11], lw[12]));
    return true;
}
PT_HD bool sample_light(const Shape *L, unsigned &s0, unsigned &s1, f3 hp, f3 nl, f3 *sd_out,
                                             float *len_out, f3 *contrib) {
    // the words the sample reads, loaded together ahead of the branch on the light's type (one
    // memory round trip instead of the type word first, then the geometry)
    const float *src = reinterpret_cast<const float *>(L);
    float lw[13];
#pragma unroll
    for (int i = 0; i < 13; i++) lw[i] = src[i];
    const float area = src[25];
    asm volatile("" ::: "memory");   // compiler-only: keeps the loads here, ahead of the branches
    return sample_light_words(lw, area, s0, s1, hp, nl, sd_out, len_out, contrib);
}

// ---- cosine-weighted diffuse bounce (rendering_kernel.cl:629-661) ----
PT_HD f3 diffuse_dir(f3 nl, unsigned &s0, unsigned &s1) {
    const float r1 = 2.f * kPi * get_random(s0, s1);
    const float r2 = get_random(s0, s1);
    const float r2s = sqrtf(r2);
    f3 ww = nl;
    const f3 aa = fabsf(ww.x) > .1f ? mk(0.f, 1.f, 0.f) : mk(1.f, 0.f, 0.f);
    f3 u = norm3(cross3(aa, ww));
    f3 v = cross3(ww, u);
    float sr1, cr1;
    sincos_f32(r1, &sr1, &cr1);
    u = smul(cr1 * r2s, u);
    v = smul(sr1 * r2s, v);
    f3 nd = add3(u, v);
    ww = smul(sqrtf(1 - r2), ww);
    return add3(nd, ww);
}

// ---- specular and refractive continuation (rendering_kernel.cl:666-752); updates thr/rd ----
PT_HD void specular_or_refract(const Surface &S, f3 &rd, f3 &thr, unsigned &s0, unsigned &s1) {
    f3 nd = smul(2.f * dot3(S.nrm, rd), S.nrm);
    nd = sub3(rd, nd);
    if (S.refl == SPEC) {
        thr = mul3(thr, S.c);
        rd = nd;
        return;
    }
    const bool into = dot3(S.nrm, S.nl) > 0;
    const float nc = 1.f, nt = 1.5f;
    const float nnt = into ? nc / nt : nt / nc;
    const float ddn = dot3(rd, S.nl);
    const float cos2t = 1.f - nnt * nnt * (1.f - ddn * ddn);
    if (cos2t < 0.f) {
        thr = mul3(thr, S.c);
        rd = nd;
        return;
    }
    const float kk = (float)(into ? 1 : -1) * (ddn * nnt + sqrtf(cos2t));
    const f3 nkk = smul(kk, S.nrm);
    f3 td = smul(nnt, rd);
    td = sub3(td, nkk);
    td = norm3(td);
    const float a = nt - nc, b = nt + nc;
    const float R0 = a * a / (b * b);
    const float cc = 1 - (into ? -ddn : dot3(td, S.nrm));
    const float Re = R0 + (1 - R0) * cc * cc * cc * cc * cc;
    const float Tr = 1.f - Re;
    const float Pp = .25f + .5f * Re;
    const float RP = Re / Pp;
    const float TP = Tr / (1.f - Pp);
    if (get_random(s0, s1) < Pp) {
        thr = smul(RP, thr);
        thr = mul3(thr, S.c);
        rd = nd;
    } else {
        thr = smul(TP, thr);
        thr = mul3(thr, S.c);
        rd = td;
    }
}

// ---- running mean of RadianceGPU (rendering_kernel.cl:1127-1136) ----
PT_HD f3 running_mean(f3 col, f3 rad, int cs) {
    if (cs == 0) return rad;
    const float k1 = (float)cs;
    const float k2 = 1.f / ((float)cs + 1.f);
    return mk((col.x * k1 + rad.x) * k2, (col.y * k1 + rad.y) * k2, (col.z * k1 + rad.z) * k2);
}

PT_HD unsigned pack_rgba(f3 col) {
    return (unsigned)to_int(col.x) | ((unsigned)to_int(col.y) << 8) | ((unsigned)to_int(col.z) << 16) | 0xff000000u;
}

}  // namespace ptgpu
