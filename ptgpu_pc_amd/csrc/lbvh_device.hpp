// lbvh_device.hpp -- device bodies of the LBVH build stages.  Each body takes the block index it
// plays and the block count of its scene, so the same code runs as a single-scene kernel (block =
// blockIdx.x) and inside the batched kernels, which map a launch's blocks onto many scenes.
#pragma once

#include <cfloat>

#include "device_util.hpp"
#include "lbvh.hpp"

namespace ptgpu {

__device__ __forceinline__ float shape_f(const Shape *s, int word) {
    return reinterpret_cast<const float *>(s)[word];
}
__device__ __forceinline__ int shape_i(const Shape *s, int word) {
    return reinterpret_cast<const int *>(s)[word];
}

// CLBVH::getBound CLBVH.cpp:49-75 from the Shape words (type @0, p1..p3 / rad,p @1..9).
__device__ __forceinline__ Bound bound_of_words(const float *w, int *type_out) {
    int type = __float_as_int(w[0]);
    *type_out = type;
    Bound b;
    if (type == SPHERE) {
        float rad = w[1], px = w[2], py = w[3], pz = w[4];
        b.min_x = px - rad; b.max_x = px + rad;
        b.min_y = py - rad; b.max_y = py + rad;
        b.min_z = pz - rad; b.max_z = pz + rad;
    } else {
        float x1 = w[1], y1 = w[2], z1 = w[3], x2 = w[4], y2 = w[5], z2 = w[6], x3 = w[7], y3 = w[8], z3 = w[9];
        b.min_x = sel_min(sel_min(x1, x2), x3); b.max_x = sel_max(sel_max(x1, x2), x3);
        b.min_y = sel_min(sel_min(y1, y2), y3); b.max_y = sel_max(sel_max(y1, y2), y3);
        b.min_z = sel_min(sel_min(z1, z2), z3); b.max_z = sel_max(sel_max(z1, z2), z3);
    }
    return b;
}
__device__ __forceinline__ Bound shape_bound(const Shape *s, int *type_out) {
    return bound_of_words(reinterpret_cast<const float *>(s), type_out);
}

// -------------------------------------------------------------------------------------------
// Stage 1: AABBs, scene centroid bounds, Morton codes (CLBVH.cpp:90-164)
// -------------------------------------------------------------------------------------------
constexpr int kBoundsThreads = 256;

__device__ __forceinline__ void reduce6(float *v, float *sh /*6*kBoundsThreads*/) {
    const int tid = threadIdx.x;
    for (int k = 0; k < 6; k++) sh[k * kBoundsThreads + tid] = v[k];
    __syncthreads();
    for (int s = kBoundsThreads / 2; s > 0; s >>= 1) {
        if (tid < s) {
            for (int k = 0; k < 6; k++) {
                float a = sh[k * kBoundsThreads + tid], b = sh[k * kBoundsThreads + tid + s];
                sh[k * kBoundsThreads + tid] = (k & 1) ? sel_max(a, b) : sel_min(a, b);
            }
        }
        __syncthreads();
    }
    for (int k = 0; k < 6; k++) v[k] = sh[k * kBoundsThreads];
}

// Per shape: bound -> Shape.b, index -> Shape.index, centroid; block partial of the centroid
// min/max (max starts at FLT_MIN as CLBVH.cpp:92-96 does) -> partials[bid].
// (block `bid` of `nblocks` working on one scene; the kernels map blocks to scenes)
// A shape's box record, 64 aligned bytes in input order: words 0-9 of the Shape (type, then the
// triangle's vertices or the sphere's radius and centre) and its AABB (min_x, max_x, min_y, max_y,
// min_z, max_z).  The Morton pass reads it instead of the 104-byte Shape and the leaf refit gathers
// it (one aligned 64-byte record per leaf instead of a whole Shape).
struct alignas(16) ShapeBox {
    float4 q[4];
};
static_assert(sizeof(ShapeBox) == 64, "ShapeBox");
__device__ __forceinline__ const float *box_words(const ShapeBox &r) { return reinterpret_cast<const float *>(&r); }

// Contiguous runs of fixed-size records moved between global memory and a wave's LDS staging area
// with 16-byte accesses: lane l's record occupies dwords [l * WORDS, (l + 1) * WORDS) of the stage,
// and the run of `count` (<= 64) records starts at a 16-byte aligned address.  A wave then reads or
// writes its run in WORDS / 4 instructions instead of WORDS dword accesses per lane at a WORDS * 4-byte
// stride (each of which touches as many cache lines as the run does).
template <int WORDS>
__device__ __forceinline__ void wave_store_records(const unsigned *stage, unsigned *dst, int count, int lane) {
    const int total = count * WORDS;
    const int full = total >> 2;
    for (int c = lane; c < full; c += kWave)
        *reinterpret_cast<uint4 *>(dst + 4 * c) = *reinterpret_cast<const uint4 *>(stage + 4 * c);
    const int rem = total - 4 * full;
    if (lane < rem) dst[4 * full + lane] = stage[4 * full + lane];
}
template <int WORDS>
__device__ __forceinline__ void wave_load_records(const unsigned *src, unsigned *stage, int count, int lane) {
    const int total = count * WORDS;
    const int full = total >> 2;
    for (int c = lane; c < full; c += kWave)
        *reinterpret_cast<uint4 *>(stage + 4 * c) = *reinterpret_cast<const uint4 *>(src + 4 * c);
    const int rem = total - 4 * full;
    if (lane < rem) stage[4 * full + lane] = src[4 * full + lane];
}

__device__ __forceinline__ void bounds_body(const Shape *shapes, int n, float *partials, ShapeBox *boxes, int bid,
                                            int nblocks) {
    __shared__ float sh[6 * kBoundsThreads];
    float acc[6] = {FLT_MAX, FLT_MIN, FLT_MAX, FLT_MIN, FLT_MAX, FLT_MIN};
    for (int i = bid * blockDim.x + threadIdx.x; i < n; i += nblocks * blockDim.x) {
        const Shape *s = shapes + i;
        int type;
        Bound b = shape_bound(s, &type);
        const float *wf = reinterpret_cast<const float *>(s);
        ShapeBox r;
        r.q[0] = make_float4(wf[0], wf[1], wf[2], wf[3]);
        r.q[1] = make_float4(wf[4], wf[5], wf[6], wf[7]);
        r.q[2] = make_float4(wf[8], wf[9], b.min_x, b.max_x);
        r.q[3] = make_float4(b.min_y, b.max_y, b.min_z, b.max_z);
        boxes[i] = r;
        float px, py, pz;
        if (type == SPHERE) {
            px = wf[2]; py = wf[3]; pz = wf[4];
        } else {
            px = (b.min_x + b.max_x) / 2.0f;
            py = (b.min_y + b.max_y) / 2.0f;
            pz = (b.min_z + b.max_z) / 2.0f;
        }
        if (px < acc[0]) acc[0] = px;
        if (px > acc[1]) acc[1] = px;
        if (py < acc[2]) acc[2] = py;
        if (py > acc[3]) acc[3] = py;
        if (pz < acc[4]) acc[4] = pz;
        if (pz > acc[5]) acc[5] = pz;
    }
    reduce6(acc, sh);
    if (threadIdx.x < 6) partials[bid * 6 + threadIdx.x] = acc[threadIdx.x];
}

// One block folds a scene's block partials into scene6 = {min_x, max_x, min_y, max_y, min_z, max_z}.
// A separate launch rather than a last-block hand-off: on gfx950 a device-scope release per block
// writes back the L2, which holds every Shape.b just written.
__device__ __forceinline__ void bounds_final_body(const float *partials, int nblocks, float *scene6) {
    __shared__ float sh[6 * kBoundsThreads];
    float fin[6] = {FLT_MAX, FLT_MIN, FLT_MAX, FLT_MIN, FLT_MAX, FLT_MIN};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int k = 0; k < 6; k++) {
            const float v = partials[b * 6 + k];
            fin[k] = (k & 1) ? sel_max(fin[k], v) : sel_min(fin[k], v);
        }
    reduce6(fin, sh);
    if (threadIdx.x < 6) scene6[threadIdx.x] = fin[threadIdx.x];
}

__device__ __forceinline__ unsigned f2u_code(float v) { return v >= 0.f ? (unsigned)v : 0u; }

__device__ __forceinline__ unsigned expand10(unsigned x) {
    x &= 0x3FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// CLBVH.cpp:125-164: quantise each centroid to 10 bits per axis (1024 wraps to 0 through the
// low-10-bit interleave) and interleave x above y above z.  The Morton pass also emits the sort key
// ~code (30 bits), so that an ascending stable sort yields the reference's stable descending order,
// and stores the reference's per-shape outputs: index @68, morton_code @72, b @76.
__device__ __forceinline__ unsigned morton_code_of(float px, float py, float pz, const float *scene6) {
    const float mnx = scene6[0], mxx = scene6[1], mny = scene6[2], mxy = scene6[3], mnz = scene6[4],
                mxz = scene6[5];
    unsigned a = f2u_code(((px - mnx) / (mxx - mnx)) * 1024.f);
    unsigned b = f2u_code(((py - mny) / (mxy - mny)) * 1024.f);
    unsigned c = f2u_code(((pz - mnz) / (mxz - mnz)) * 1024.f);
    return (expand10(a) << 2) | (expand10(b) << 1) | expand10(c);
}

// A wave handles 64 consecutive shapes.  When the Shape array is 16-byte aligned, the wave loads its
// 64 whole records (26 words each) into LDS with 16-byte loads, fills in index, morton_code and b,
// and stores the records back whole: full cache lines, where writing only each record's 32-byte tail
// at a 104-byte stride left every line partly written.  The centroid then comes from the record's
// own words; otherwise (a caller's unaligned array) each lane reads its box record and stores its tail.
__device__ __forceinline__ void morton_body(Shape *shapes, int n, const float *scene6, unsigned *keys,
                                            const ShapeBox *boxes, int bid) {
    constexpr int kW = 26;   // words per Shape
    __shared__ __align__(16) unsigned stage[kBoundsThreads / kWave][kWave * kW];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const long long first = (long long)bid * blockDim.x + (long long)wv * kWave;
    if (first >= n) return;   // the whole wave is past the end
    const int count = (long long)n - first < kWave ? (int)((long long)n - first) : kWave;
    const int i = (int)first + lane;
    if ((reinterpret_cast<size_t>(shapes) & 15) == 0) {
        unsigned *st = stage[wv];
        wave_load_records<kW>(reinterpret_cast<const unsigned *>(shapes + first), st, count, lane);
        __builtin_amdgcn_wave_barrier();
        if (lane < count) {
            unsigned *rw = st + lane * kW;
            float w[10];
#pragma unroll
            for (int f = 0; f < 10; f++) w[f] = __uint_as_float(rw[f]);
            int type;
            const Bound b = bound_of_words(w, &type);
            float px, py, pz;
            if (type == SPHERE) {
                px = w[2]; py = w[3]; pz = w[4];
            } else {
                px = (b.min_x + b.max_x) / 2.0f;
                py = (b.min_y + b.max_y) / 2.0f;
                pz = (b.min_z + b.max_z) / 2.0f;
            }
            const unsigned code = morton_code_of(px, py, pz, scene6);
            rw[17] = (unsigned)i;   // index @68
            rw[18] = code;          // morton_code @72
            rw[19] = __float_as_uint(b.min_x); rw[20] = __float_as_uint(b.max_x);   // b @76
            rw[21] = __float_as_uint(b.min_y); rw[22] = __float_as_uint(b.max_y);
            rw[23] = __float_as_uint(b.min_z); rw[24] = __float_as_uint(b.max_z);
            keys[i] = (~code) & 0x3FFFFFFFu;
        }
        __builtin_amdgcn_wave_barrier();
        wave_store_records<kW>(st, reinterpret_cast<unsigned *>(shapes + first), count, lane);
        return;
    }
    if (lane >= count) return;
    const ShapeBox r = boxes[i];
    const float *w = box_words(r);   // w[10..15]: the AABB
    const int type = __float_as_int(w[0]);
    float px, py, pz;
    if (type == SPHERE) {
        px = w[2]; py = w[3]; pz = w[4];
    } else {
        px = (w[10] + w[11]) / 2.0f;
        py = (w[12] + w[13]) / 2.0f;
        pz = (w[14] + w[15]) / 2.0f;
    }
    const unsigned code = morton_code_of(px, py, pz, scene6);
    unsigned *ws = reinterpret_cast<unsigned *>(shapes + i);
    ws[17] = (unsigned)i;   // index @68
    ws[18] = code;          // morton_code @72
    for (int f = 0; f < 6; f++) ws[19 + f] = __float_as_uint(w[10 + f]);   // b @76
    keys[i] = (~code) & 0x3FFFFFFFu;
}

// -------------------------------------------------------------------------------------------
// Stage 2: onesweep LSD radix sort (Adinets & Merrill 2022) for wave64.
//   k_sort_hist:   one read of the keys builds the digit histograms of every pass.
//   k_sort_scan:   exclusive scan of each pass histogram (global digit bases).
//   k_sort_pass:   per 4096-key tile: wave-level match ranking in input order (stable), tile
//                  histogram, decoupled look-back over earlier tiles per digit, LDS shuffle into
//                  digit order, coalesced scatter.  Tile ids come from an atomic counter so every
//                  tile a block waits on has already started.
// -------------------------------------------------------------------------------------------
constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / kWave;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr unsigned kFlagAgg = 1u << 30, kFlagPrefix = 2u << 30, kValMask = (1u << 30) - 1;
constexpr unsigned kSpinLimit = 1u << 26;
constexpr int kLookback = 8;   // predecessor status words per look-back round

template <int BITS> struct SortCfg {
    static constexpr int kRadix = 1 << BITS;
    static constexpr unsigned kMask = (unsigned)kRadix - 1;
    static constexpr int kPasses = (30 + BITS - 1) / BITS;
};

template <int BITS>
__device__ __forceinline__ void sort_hist_body(const unsigned *keys, int n, unsigned *hist, int bid, int nblocks) {
    using C = SortCfg<BITS>;
    __shared__ unsigned sh[C::kPasses * C::kRadix];
    for (int j = threadIdx.x; j < C::kPasses * C::kRadix; j += kSortThreads) sh[j] = 0;
    __syncthreads();
    for (int i = bid * kSortThreads + threadIdx.x; i < n; i += nblocks * kSortThreads) {
        unsigned k = keys[i];
#pragma unroll
        for (int p = 0; p < C::kPasses; p++) atomicAdd(&sh[p * C::kRadix + ((k >> (p * BITS)) & C::kMask)], 1u);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < C::kPasses * C::kRadix; j += kSortThreads)
        if (sh[j]) atomicAdd(&hist[j], sh[j]);
}

// exclusive scan of `v` across the 256 threads of a block; returns the exclusive prefix
__device__ __forceinline__ unsigned block_exscan256(unsigned v, unsigned *s_w, unsigned *total) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    unsigned x = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        unsigned y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    if (lane == kWave - 1) s_w[w] = x;
    __syncthreads();
    unsigned pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kSortWaves; k++) {
        unsigned t = s_w[k];
        if (k < w) pre += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

template <int BITS>
__device__ __forceinline__ void sort_scan_body(unsigned *hist) {
    using C = SortCfg<BITS>;
    __shared__ unsigned s_w[kSortWaves];
    for (int p = 0; p < C::kPasses; p++) {
        for (int base = 0; base < C::kRadix; base += kSortThreads) {
            int d = base + threadIdx.x;
            unsigned v = d < C::kRadix ? hist[p * C::kRadix + d] : 0u;
            unsigned tot;
            unsigned ex = block_exscan256(v, s_w, &tot);
            if (d < C::kRadix) hist[p * C::kRadix + d] = ex;   // kRadix <= 256: one round
        }
    }
}

// The element forms of a pass's input and output.  kSplit: a key array and a value array (an input
// without a value array takes each element's index as its value).  kPacked: one 8-byte word per
// element, the key in its low half and the value in its high half, read from / written to the
// keys_in / keys_out pointer (vals_in / vals_out unused): one load and one store per element, and each
// tile's digit runs are twice as long in bytes as in either split array, so the scatter leaves fewer
// partly written cache lines.
enum SortForm { kSplit = 0, kPacked = 1 };

template <int BITS, int IN = kSplit, int OUT = kSplit>
__device__ __forceinline__ void sort_pass_body(const unsigned *__restrict__ keys_in, const unsigned *__restrict__ vals_in,
                                               unsigned *__restrict__ keys_out, unsigned *__restrict__ vals_out, int n,
                                               int shift, const unsigned *__restrict__ digit_base, unsigned *status,
                                               unsigned tile, unsigned *err) {
    using C = SortCfg<BITS>;
    __shared__ unsigned long long s_kv[kSortTile];   // the tile in digit order, key | value << 32
    __shared__ unsigned s_wcnt[kSortWaves][C::kRadix];
    __shared__ unsigned s_lofs[C::kRadix];
    __shared__ unsigned s_gofs[C::kRadix];
    __shared__ unsigned s_w[kSortWaves];

    const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
    for (int j = tid; j < kSortWaves * C::kRadix; j += kSortThreads) (&s_wcnt[0][0])[j] = 0;
    __syncthreads();
    const long long base = (long long)tile * kSortTile;
    const long long wbase = base + (long long)w * kWave * kSortItems;

    unsigned k[kSortItems], v[kSortItems], rank[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        long long idx = wbase + j * kWave + lane;
        bool valid = idx < n;
        if (IN == kPacked) {
            const unsigned long long x = valid ? reinterpret_cast<const unsigned long long *>(keys_in)[idx] : ~0ull;
            k[j] = (unsigned)x;
            v[j] = (unsigned)(x >> 32);
        } else {
            k[j] = valid ? keys_in[idx] : 0xFFFFFFFFu;
            v[j] = valid ? (vals_in ? vals_in[idx] : (unsigned)idx) : 0u;
        }
    }
    const unsigned long long lt = lanemask_lt();
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        long long idx = wbase + j * kWave + lane;
        bool valid = idx < n;
        unsigned digit = (k[j] >> shift) & C::kMask;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < BITS; b++) {
            bool bit = (digit >> b) & 1u;
            unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        unsigned pre = 0;
        if (valid) pre = s_wcnt[w][digit];
        __builtin_amdgcn_wave_barrier();
        if (valid) {
            rank[j] = pre + (unsigned)__popcll(peers & lt);
            if ((peers & lt) == 0) s_wcnt[w][digit] = pre + (unsigned)__popcll(peers);
        }
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // per digit: wave-exclusive offsets, tile count, tile-local offsets, global base (look-back)
    for (int d0 = 0; d0 < C::kRadix; d0 += kSortThreads) {
        const int d = d0 + tid;
        unsigned cnt = 0;
        if (d < C::kRadix) {
            unsigned run = 0;
#pragma unroll
            for (int q = 0; q < kSortWaves; q++) {
                unsigned c = s_wcnt[q][d];
                s_wcnt[q][d] = run;
                run += c;
            }
            cnt = run;
        }
        unsigned tot;
        unsigned lofs = block_exscan256(cnt, s_w, &tot);
        if (d < C::kRadix) {
            s_lofs[d] = lofs;
            unsigned *my = status + (size_t)tile * C::kRadix + d;
            unsigned prefix = 0;
            if (tile == 0) {
                st_relaxed(my, kFlagPrefix | cnt);
            } else {
                st_relaxed(my, kFlagAgg | cnt);
                // Windowed look-back: the status words of kLookback predecessors are loaded together
                // and summed nearest first up to the first inclusive prefix, so a walk over tiles
                // that have published only their counts costs one round trip per window instead of
                // one per tile.  A word not yet published stops the window; the walk resumes there.
                long long t = (long long)tile - 1;
                unsigned spins = 0;
                while (t >= 0) {
                    unsigned w[kLookback];
#pragma unroll
                    for (int k = 0; k < kLookback; k++)
                        w[k] = t - k >= 0 ? ld_relaxed(status + (size_t)(t - k) * C::kRadix + d) : kFlagPrefix;
                    bool stop = false, found = false;
                    int adv = 0;
#pragma unroll
                    for (int k = 0; k < kLookback; k++) {
                        if (stop) continue;
                        const unsigned flag = w[k] & ~kValMask;
                        if (flag == 0) {
                            stop = true;
                        } else {
                            prefix += w[k] & kValMask;
                            adv++;
                            if (flag == kFlagPrefix) stop = found = true;
                        }
                    }
                    if (found) break;
                    t -= adv;
                    if (adv < kLookback && ++spins > kSpinLimit) { atomicOr(err, 1u); break; }
                }
                st_relaxed(my, kFlagPrefix | ((prefix + cnt) & kValMask));
            }
            s_gofs[d] = digit_base[d] + prefix;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        long long idx = wbase + j * kWave + lane;
        if (idx < n) {
            unsigned digit = (k[j] >> shift) & C::kMask;
            unsigned lpos = s_lofs[digit] + s_wcnt[w][digit] + rank[j];
            s_kv[lpos] = (unsigned long long)k[j] | ((unsigned long long)v[j] << 32);
        }
    }
    __syncthreads();
    long long rem = (long long)n - base;
    const int count = rem < kSortTile ? (int)rem : kSortTile;
    for (int i = tid; i < count; i += kSortThreads) {
        const unsigned long long x = s_kv[i];
        unsigned key = (unsigned)x;
        unsigned digit = (key >> shift) & C::kMask;
        unsigned dest = s_gofs[digit] + (unsigned)i - s_lofs[digit];
        if (OUT == kPacked) {
            reinterpret_cast<unsigned long long *>(keys_out)[dest] = x;
        } else {
            keys_out[dest] = key;
            vals_out[dest] = (unsigned)(x >> 32);
        }
    }
}

// -------------------------------------------------------------------------------------------
// Stage 3: Karras radix tree over leaf indices (BVH.cl:127-208).  The prefix length of two leaves
// is clz(i ^ j); the tree is therefore the binary radix tree of 0..N-1.
// -------------------------------------------------------------------------------------------
__device__ __forceinline__ int delta_idx(int i, int j, int n) {
    return (j >= 0 && j < n) ? __clz(i ^ j) : -1;
}

// Each thread also writes the reference's initial contents (CLBVH.cpp:83-88: internal nodes zeroed,
// leaves filled with 0xFF bytes) into every field of internal node i and leaf i but nParent, which
// the parent's thread writes (the root's is 0); thread n-2 fills leaf n-1 as well.  One pass of
// whole records instead of two fills followed by scattered field stores.
__device__ __forceinline__ void fill_leaf_record(BVHNodeGPU *l) {
    int *w = reinterpret_cast<int *>(l);
    for (int f = 0; f < 15; f++)
        if (f != 6) w[f] = -1;   // nParent (field 6) comes from the parent
}

__device__ __forceinline__ void karras_body(int n, BVHNodeGPU *btn, BVHNodeGPU *btl, int bid) {
    const int i = bid * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    fill_leaf_record(btl + i);
    if (i == n - 2) fill_leaf_record(btl + n - 1);
    const int d = (delta_idx(i, i + 1, n) - delta_idx(i, i - 1, n)) > 0 ? 1 : -1;
    const int dmin = delta_idx(i, i - d, n);
    int lmax = 2;
    while (delta_idx(i, i + lmax * d, n) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta_idx(i, i + (l + t) * d, n) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta_idx(i, j, n);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta_idx(i, i + (s + t) * d, n) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? d : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    int leafflag = 0;
    if (lo == gamma) { leafflag = 2; btl[gamma].nParent = i; }
    else btn[gamma].nParent = i;
    if (hi == gamma + 1) { leafflag = 2; btl[gamma + 1].nParent = i; }
    else btn[gamma + 1].nParent = i;
    int *w = reinterpret_cast<int *>(btn + i);
    w[0] = lo;           // min
    w[1] = hi;           // max
    w[2] = 0;            // nShape
    w[3] = leafflag;     // leaf
    w[4] = gamma;        // nLeft
    w[5] = gamma + 1;    // nRight
    if (i == 0) w[6] = 0;   // the root's nParent
    for (int f = 7; f < 15; f++) w[f] = 0;   // bound, cost, area
}

// -------------------------------------------------------------------------------------------
// Stage 4: leaves + bottom-up refit (BVH.cl:214-258).  Leaf k takes the AABB and index of shape
// sorted[k]; a parent's bound is the merge of its (left, right) child bounds, where a child is a
// leaf iff nLeft == min / nRight == max.  Also writes the traversal layout (TNode per internal
// node, Prim per leaf).
// -------------------------------------------------------------------------------------------
// w: words 0-9 of a Shape (a ShapeBox's first ten)
__device__ __forceinline__ Prim make_prim(const float *w, int id) {
    const int type = __float_as_int(w[0]) | kLastInLeaf;   // one primitive per leaf
    Prim p;
    if ((type & 0xff) == SPHERE) {
        p.p0 = make_float4(w[2], w[3], w[4], __int_as_float(id));
        p.p1 = make_float4(w[1], 0.f, 0.f, __int_as_float(type));
        p.p2 = make_float4(0.f, 0.f, 0.f, __int_as_float(type));
    } else {
        float x1 = w[1], y1 = w[2], z1 = w[3];
        p.p0 = make_float4(x1, y1, z1, __int_as_float(id));
        p.p1 = make_float4(w[4] - x1, w[5] - y1, w[6] - z1, __int_as_float(type));
        p.p2 = make_float4(w[7] - x1, w[8] - y1, w[9] - z1, __int_as_float(type));
    }
    return p;
}

__device__ __forceinline__ void store_bound(BVHNodeGPU *node, const Bound &b) {
    float *f = reinterpret_cast<float *>(&node->bound);
    f[0] = b.min_x; f[1] = b.max_x; f[2] = b.min_y; f[3] = b.max_y; f[4] = b.min_z; f[5] = b.max_z;
}

__device__ __forceinline__ TNode make_tnode(const Bound &L, const Bound &R, int rl, int rr, int cl, int cr) {
    TNode t;
    t.a = make_float4(L.min_x, L.min_y, L.min_z, L.max_x);
    t.b = make_float4(L.max_y, L.max_z, R.min_x, R.min_y);
    t.c = make_float4(R.min_z, R.max_x, R.max_y, R.max_z);
    t.r = make_int4(rl, rr, cl, cr);
    return t;
}

// -------------------------------------------------------------------------------------------
// Stage 4 (index-keyed trees): fence-free refit by level passes.
// The Karras tree over leaf indices is the binary radix tree of 0..N-1, so every internal node
// covers an aligned block [j*2^s, min((j+1)*2^s, N) - 1] of leaves, with s = ceil(log2(len)),
// and its Karras index is a closed form of (lo, hi, N): the node is the LEFT child of its parent
// (index hi) iff bit s of lo is 0 and leaves remain after the block, otherwise it is the right
// child or the root (index lo).  One workgroup reduces 1024 consecutive items (leaves, or subtree
// roots of the previous pass) bottom-up in LDS, merging (left, right) exactly as BVH.cl:112-122,
// writes every internal node's bound and TNode on the way, and hands its block root to the next
// pass.  No atomics, no cross-workgroup visibility, deterministic; identical results to the
// reference's visit-flag refit (BVH.cl:214-258), whose arrival order does not change a bound.
// -------------------------------------------------------------------------------------------
constexpr int kRefitLog = 10;
constexpr int kRefitItems = 1 << kRefitLog;

__device__ __forceinline__ int karras_index(long long lo, long long hi, int n) {
    const int len = (int)(hi - lo + 1);
    const int s = 32 - __clz(len - 1);   // ceil(log2(len)), len >= 2
    return (lo + (1ll << s) <= n - 1 && !((lo >> s) & 1)) ? (int)hi : (int)lo;
}
__device__ __forceinline__ int child_ref(long long lo, long long hi, int n) {
    return lo == hi ? ~(int)lo : karras_index(lo, hi, n);
}
// Parent of the node (or leaf, s = 0) that covers the aligned block of level s starting at lo: the
// first enclosing block whose right half is not empty (a block with an empty right half is the
// same node one level up).  The root's nParent is 0 (the reference zeroes internal nodes), a lone
// leaf's stays 0xFF-filled.
__device__ __forceinline__ int index_tree_parent(long long lo, int s, int n) {
    if (n < 2) return -1;
    for (int t = s + 1; t < 40; t++) {
        const long long plo = lo & ~((1ll << t) - 1);
        const long long pmid = plo + (1ll << (t - 1));
        if (pmid <= (long long)n - 1) {
            const long long phi_full = plo + (1ll << t) - 1;
            return karras_index(plo, phi_full < n - 1 ? phi_full : (long long)n - 1, n);
        }
    }
    return 0;
}
// Whole reference records (BVHNodeGPU.h:4-16, 15 words) in one pass: a leaf keeps the 0xFF fill
// (CLBVH.cpp:83-88) in every field but nShape, nParent and bound; an internal node gets its Karras
// fields (min, max, leaf = 2 when a child is a leaf, nLeft = split, nRight = split + 1, nParent)
// with nShape, cost and area zero.
__device__ __forceinline__ void write_leaf_record(BVHNodeGPU *l, int sid, int parent, const Bound &B) {
    int *w = reinterpret_cast<int *>(l);
    w[0] = -1; w[1] = -1; w[2] = sid; w[3] = -1; w[4] = -1; w[5] = -1; w[6] = parent;
    store_bound(l, B);
    w[13] = -1; w[14] = -1;
}
__device__ __forceinline__ void write_node_record(BVHNodeGPU *d, int lo, int hi, int mid, int parent, const Bound &B) {
    int *w = reinterpret_cast<int *>(d);
    w[0] = lo; w[1] = hi; w[2] = 0; w[3] = (lo == mid - 1 || hi == mid) ? 2 : 0;
    w[4] = mid - 1; w[5] = mid; w[6] = parent;
    store_bound(d, B);
    w[13] = 0; w[14] = 0;
}

template <bool LEAF>
__device__ __forceinline__ void refit_levels_body(const ShapeBox *boxes, const unsigned *sorted, int n, int s0, int levels,
                                                  const float *in_bounds, float *out_bounds, BVHNodeGPU *btn,
                                                  BVHNodeGPU *btl, TNode *tnodes, Prim *prims, bool tnodes_on, int bid) {
    __shared__ float sb[6][kRefitItems];
    // per wave: 64 leaf records (15 words) or 64 Prims (12 words), staged for contiguous stores
    __shared__ __align__(16) unsigned stage[LEAF ? 4 : 1][LEAF ? 64 * 15 : 4];
    const long long item_base = (long long)bid * kRefitItems;
    const long long m = ((long long)n + (1ll << s0) - 1) >> s0;   // items at level s0
    // leaf level: all of this thread's gathers are issued before any store (a store between them
    // would make each gather wait for the stores before it)
    constexpr int kPer = kRefitItems / 256;
    int sids[kPer];
    ShapeBox recs[kPer];
    if (LEAF) {
#pragma unroll
        for (int q = 0; q < kPer; q++) {
            const long long it = item_base + threadIdx.x + q * 256;
            sids[q] = it < m ? (int)sorted[it] : -1;
        }
#pragma unroll
        for (int q = 0; q < kPer; q++)
            if (sids[q] >= 0) recs[q] = boxes[sids[q]];
    }
#pragma unroll
    for (int q = 0; q < kPer; q++) {
        const int j = threadIdx.x + q * 256;
        const long long it = item_base + j;
        if (LEAF) {
            // the wave's 64 items are consecutive leaves: their records go out as one contiguous run
            const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
            const long long first = it - lane;
            const int count = first >= m ? 0 : (m - first < kWave ? (int)(m - first) : kWave);
            unsigned *st = stage[wv];
            Bound B;
            const float *w = box_words(recs[q]);
            if (it < m) {
                B.min_x = w[10]; B.max_x = w[11]; B.min_y = w[12]; B.max_y = w[13]; B.min_z = w[14]; B.max_z = w[15];
                BVHNodeGPU rec;
                write_leaf_record(&rec, sids[q], index_tree_parent(it, 0, n), B);
                const unsigned *rw = reinterpret_cast<const unsigned *>(&rec);
#pragma unroll
                for (int f = 0; f < 15; f++) st[lane * 15 + f] = rw[f];
                sb[0][j] = B.min_x; sb[1][j] = B.max_x; sb[2][j] = B.min_y;
                sb[3][j] = B.max_y; sb[4][j] = B.min_z; sb[5][j] = B.max_z;
            }
            __builtin_amdgcn_wave_barrier();
            wave_store_records<15>(st, reinterpret_cast<unsigned *>(btl + first), count, lane);
            __builtin_amdgcn_wave_barrier();
            if (it < m) {
                const Prim p = make_prim(w, sids[q]);
                const unsigned *pw = reinterpret_cast<const unsigned *>(&p);
#pragma unroll
                for (int f = 0; f < 12; f++) st[lane * 12 + f] = pw[f];
            }
            __builtin_amdgcn_wave_barrier();
            wave_store_records<12>(st, reinterpret_cast<unsigned *>(prims + first), count, lane);
            __builtin_amdgcn_wave_barrier();
            continue;
        }
        if (it >= m) continue;
        Bound B;
        {
            const float *f = in_bounds + it * 6;
            B.min_x = f[0]; B.max_x = f[1]; B.min_y = f[2]; B.max_y = f[3]; B.min_z = f[4]; B.max_z = f[5];
        }
        sb[0][j] = B.min_x; sb[1][j] = B.max_x; sb[2][j] = B.min_y;
        sb[3][j] = B.max_y; sb[4][j] = B.min_z; sb[5][j] = B.max_z;
    }
    __syncthreads();
    for (int lev = 1; lev <= levels; lev++) {
        const int s = s0 + lev;
        const int nodes_here = kRefitItems >> lev;
        for (int j = threadIdx.x; j < nodes_here; j += 256) {
            const long long lo = (item_base + ((long long)j << lev)) << s0;
            const long long mid = lo + (1ll << (s - 1));
            if (mid > (long long)n - 1) continue;   // right half empty (or block past the end): collapsed
            const long long hi_full = lo + (1ll << s) - 1;
            const long long hi = hi_full < n - 1 ? hi_full : (long long)n - 1;
            const int ls = j << lev, rs = ls + (1 << (lev - 1));
            Bound L, R;
            L.min_x = sb[0][ls]; L.max_x = sb[1][ls]; L.min_y = sb[2][ls];
            L.max_y = sb[3][ls]; L.min_z = sb[4][ls]; L.max_z = sb[5][ls];
            R.min_x = sb[0][rs]; R.max_x = sb[1][rs]; R.min_y = sb[2][rs];
            R.max_y = sb[3][rs]; R.min_z = sb[4][rs]; R.max_z = sb[5][rs];
            const Bound M = merge_bounds(L, R);
            sb[0][ls] = M.min_x; sb[1][ls] = M.max_x; sb[2][ls] = M.min_y;
            sb[3][ls] = M.max_y; sb[4][ls] = M.min_z; sb[5][ls] = M.max_z;
            const int idx = karras_index(lo, hi, n);
            write_node_record(btn + idx, (int)lo, (int)hi, (int)mid, index_tree_parent(lo, s, n), M);
            if (tnodes_on) {
                const int rl = child_ref(lo, mid - 1, n), rr = child_ref(mid, hi, n);
                tnodes[idx] = make_tnode(L, R, rl, rr, rl < 0 ? 1 : 0, rr < 0 ? 1 : 0);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6 && out_bounds) out_bounds[(long long)bid * 6 + threadIdx.x] = sb[threadIdx.x][0];
}

// -------------------------------------------------------------------------------------------
// Traversal tree: Karras radix tree over the sorted Morton keys (ties broken by leaf index), the
// standard LBVH.  The reference exports the index-keyed tree (kept bit-exact in btn/btl); the
// traversal layout may use any tree over the same leaves, and the Morton-keyed one halves the SAH
// cost of the index-keyed one on the C2 soup (159 vs 318).  Closest-hit and any-hit results do
// not depend on the tree.
// -------------------------------------------------------------------------------------------
__device__ __forceinline__ int delta_key(const unsigned *keys, int i, int j, int n) {
    if (j < 0 || j >= n) return -1;
    const unsigned x = keys[i] ^ keys[j];
    return x ? __clz(x) : 32 + __clz(i ^ j);
}

__device__ __forceinline__ void karras_morton_body(const unsigned *keys, int n, int4 *mnode, int *mpar_int,
                                                   int *mpar_leaf, int bid) {
    const int i = bid * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta_key(keys, i, i + 1, n) - delta_key(keys, i, i - 1, n)) > 0 ? 1 : -1;
    const int dmin = delta_key(keys, i, i - d, n);
    long long lmax = 2;
    while (delta_key(keys, i, (int)(i + lmax * d), n) > dmin) lmax <<= 1;
    long long l = 0;
    for (long long t = lmax >> 1; t >= 1; t >>= 1)
        if (delta_key(keys, i, (int)(i + (l + t) * d), n) > dmin) l += t;
    const int j = (int)(i + l * d);
    const int dnode = delta_key(keys, i, j, n);
    long long s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta_key(keys, i, (int)(i + (s + t) * d), n) > dnode) s += t;
    } while (t > 1);
    const int gamma = (int)(i + s * d + (d < 0 ? d : 0));
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const int lref = (lo == gamma) ? ~gamma : gamma;
    const int rref = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
    mnode[i] = make_int4(lref, rref, lo, hi);
    if (lref < 0) mpar_leaf[gamma] = i; else mpar_int[gamma] = i;
    if (rref < 0) mpar_leaf[gamma + 1] = i; else mpar_int[gamma + 1] = i;
    if (i == 0) mpar_int[0] = -1;
}

// Pass 1: one workgroup per 512 consecutive leaves.  Every node whose leaf range lies inside the
// chunk has its Karras index inside the chunk too, so visit counters and bounds live in LDS and the
// bottom-up climb synchronises through workgroup-scope LDS atomics only.  A climb that reaches a
// node spanning chunks publishes its child's bound and height and queues the (node, child) arrival
// in the list of the 64-chunk super-chunk that holds the node's first leaf.  512-leaf chunks keep
// the pass at 40 KB of LDS per workgroup, so it runs beside a render's persistent workgroups (two
// per CU hold about 100 KB of the 160 KB) instead of waiting for them to retire.
//
// List capacities follow from the tree height h <= 62 (30-bit codes, ties split by a 32-bit index):
// at most h nodes cross any boundary and each node receives two arrivals, so a super-chunk's list
// holds at most 2 * 62 * 64 entries and the list of nodes spanning super-chunks 2 * 62 per boundary.
// An overflow (a malformed tree) sets the build's error word instead of writing out of bounds.
constexpr int kMChunk = 512;
constexpr unsigned kMSuper = 64u * kMChunk;    // leaves per super-chunk (64 chunks)
constexpr unsigned kXCap = 8192;               // arrivals per super-chunk list
// Climb steps of one thread: a tree over 30-bit codes with ties split by a 32-bit index is at most 62
// levels high, so a climb longer than this follows a malformed parent chain (a cycle): it stops and
// sets bit 4 of the build's error word instead of spinning.
constexpr int kMaxClimb = 64;
__host__ __device__ constexpr unsigned mtop_cap(unsigned nsuper) { return 128u * (nsuper + 1u); }

__device__ __forceinline__ void mrefit_chunk_body(int n, const int4 *mnode, const int *mpar_int, const int *mpar_leaf,
                                                  const BVHNodeGPU *btl, TNode *tnodes, float *gbi, int *gh,
                                                  int2 *xlist, unsigned *xcnt, unsigned *root_height, unsigned *err,
                                                  int bid) {
    __shared__ float lb[6][kMChunk];
    __shared__ float nb[6][kMChunk];
    __shared__ int nh[kMChunk];
    __shared__ unsigned cnt[kMChunk];
    // The chunk's internal nodes, their parents and the leaves' parents, loaded once with coalesced
    // reads: the climbs below then touch global memory only for their tnodes stores and, at their
    // end, for a parent spanning chunks.  (On gfx950 a wait for a global load also waits for every
    // earlier store of the thread, so a load inside the climb would serialise it on those stores.)
    __shared__ int4 smn[kMChunk];
    __shared__ int spar[kMChunk];
    __shared__ int slp[kMChunk];
    const int base = bid * kMChunk;
    const int cend = (n - base) < kMChunk ? n : base + kMChunk;
    for (int q = threadIdx.x; q < kMChunk; q += (int)blockDim.x) {
        cnt[q] = 0;
        const int k = base + q;
        if (k < cend) {
            const Bound b = load_bound(&btl[k].bound);
            lb[0][q] = b.min_x; lb[1][q] = b.max_x; lb[2][q] = b.min_y;
            lb[3][q] = b.max_y; lb[4][q] = b.min_z; lb[5][q] = b.max_z;
            slp[q] = n > 1 ? mpar_leaf[k] : -1;
        }
        if (k < n - 1) {
            smn[q] = mnode[k];
            spar[q] = mpar_int[k];
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < kMChunk; q += (int)blockDim.x) {
        const int k = base + q;
        if (k >= cend) break;
        int child = ~k, ch = 0;
        int p = slp[q];
        bool spans = false;
        int steps = 0;
        while (p >= 0) {
            if (++steps > kMaxClimb) {
                atomicOr(err, 4u);
                p = -1;
                child = -1;
                break;
            }
            // a node whose leaf range lies inside the chunk has its index there too (Karras numbering:
            // the index is an end of the range), so an index outside the chunk means a spanning node
            if (p < base || p >= base + kMChunk) {
                spans = true;
                break;
            }
            const int4 P = smn[p - base];
            if (P.z < base || P.w >= base + kMChunk) {
                spans = true;
                break;
            }
            // The sibling's bound and height are in LDS, so the counter orders LDS only: an LDS-scope
            // release / acquire, which (unlike a workgroup-scope one over all memory) does not wait
            // for this thread's earlier tnodes stores to reach memory.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            const unsigned old = __hip_atomic_fetch_add(&cnt[p - base], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
            if (old == 0) break;
            Bound L, R;
            int hl, hr;
            {
                const int a = P.x < 0 ? ~P.x - base : P.x - base;
                float (*src)[kMChunk] = P.x < 0 ? lb : nb;
                L.min_x = src[0][a]; L.max_x = src[1][a]; L.min_y = src[2][a];
                L.max_y = src[3][a]; L.min_z = src[4][a]; L.max_z = src[5][a];
                hl = P.x < 0 ? 0 : nh[a];
            }
            {
                const int a = P.y < 0 ? ~P.y - base : P.y - base;
                float (*src)[kMChunk] = P.y < 0 ? lb : nb;
                R.min_x = src[0][a]; R.max_x = src[1][a]; R.min_y = src[2][a];
                R.max_y = src[3][a]; R.min_z = src[4][a]; R.max_z = src[5][a];
                hr = P.y < 0 ? 0 : nh[a];
            }
            const Bound M = merge_bounds(L, R);
            const int c = p - base;
            nb[0][c] = M.min_x; nb[1][c] = M.max_x; nb[2][c] = M.min_y;
            nb[3][c] = M.max_y; nb[4][c] = M.min_z; nb[5][c] = M.max_z;
            ch = (hl > hr ? hl : hr) + 1;
            nh[c] = ch;
            tnodes[p] = make_tnode(L, R, P.x, P.y, P.x < 0 ? 1 : 0, P.y < 0 ? 1 : 0);
            child = p;
            p = spar[c];
        }
        if (spans) {
            // publish the child's bound and height, queue the arrival at its spanning parent
            const int pz = (p >= base && p < base + kMChunk) ? smn[p - base].z : mnode[p].z;
            if (child >= 0) {
                const int c = child - base;
                for (int f = 0; f < 6; f++) gbi[(size_t)child * 6 + f] = nb[f][c];
                gh[child] = ch;
            }
            const unsigned sc = (unsigned)pz / kMSuper;
            const unsigned slot = atomicAdd(xcnt + sc, 1u);
            if (slot < kXCap) xlist[(size_t)sc * kXCap + slot] = make_int2(p, child);
            else atomicOr(err, 2u);
        } else if (p < 0 && child >= 0) {
            *root_height = (unsigned)ch;   // the root closed inside this chunk
        }
    }
}

// Passes 2 and 3: child bounds written by an earlier pass are visible (earlier launch); those written
// by another thread of the same workgroup are ordered by the workgroup-scope acquire/release counter.
// One node's refit from its two children (the second arrival's work in the passes below): bound and
// height into gbi / gh, the traversal node into tnodes; returns the height.
__device__ __forceinline__ int mrefit_node(int p, const int4 &P, const BVHNodeGPU *btl, TNode *tnodes, float *gbi,
                                           int *gh) {
    Bound L, R;
    int hl = 0, hr = 0;
    if (P.x < 0) L = load_bound(&btl[~P.x].bound);
    else {
        const float *f = gbi + (size_t)P.x * 6;
        L.min_x = f[0]; L.max_x = f[1]; L.min_y = f[2]; L.max_y = f[3]; L.min_z = f[4]; L.max_z = f[5];
        hl = gh[P.x];
    }
    if (P.y < 0) R = load_bound(&btl[~P.y].bound);
    else {
        const float *f = gbi + (size_t)P.y * 6;
        R.min_x = f[0]; R.max_x = f[1]; R.min_y = f[2]; R.max_y = f[3]; R.min_z = f[4]; R.max_z = f[5];
        hr = gh[P.y];
    }
    const Bound M = merge_bounds(L, R);
    float *o = gbi + (size_t)p * 6;
    o[0] = M.min_x; o[1] = M.max_x; o[2] = M.min_y; o[3] = M.max_y; o[4] = M.min_z; o[5] = M.max_z;
    const int hp = (hl > hr ? hl : hr) + 1;
    gh[p] = hp;
    tnodes[p] = make_tnode(L, R, P.x, P.y, P.x < 0 ? 1 : 0, P.y < 0 ? 1 : 0);
    return hp;
}

// Pass 2: one workgroup per super-chunk climbs from its list's arrivals through the nodes whose leaf
// range lies inside the super-chunk (their visit counters see no other workgroup); a climb that
// reaches a node spanning super-chunks queues the arrival for pass 3.
__device__ __forceinline__ void mrefit_mid_body(const int4 *mnode, const int *mpar_int, const BVHNodeGPU *btl,
                                                TNode *tnodes, float *gbi, int *gh, const int2 *xlist,
                                                const unsigned *xcnt, unsigned *gcnt, int2 *xtop, unsigned *xtop_count,
                                                unsigned xtop_cap, unsigned *root_height, unsigned *err, unsigned sc) {
    const unsigned c = xcnt[sc];
    const unsigned count = c < kXCap ? c : kXCap;
    auto defer = [&](int node, int child) {
        const unsigned slot = atomicAdd(xtop_count, 1u);
        if (slot < xtop_cap) xtop[slot] = make_int2(node, child);
        else atomicOr(err, 2u);
    };
    for (unsigned e = threadIdx.x; e < count; e += blockDim.x) {
        const int2 E = xlist[(size_t)sc * kXCap + e];
        int p = E.x;
        int4 P = mnode[p];
        if ((unsigned)P.w / kMSuper != sc) {   // spans super-chunks: pass 3
            defer(p, E.y);
            continue;
        }
        for (int steps = 0;; steps++) {
            if (steps > kMaxClimb) {
                atomicOr(err, 4u);
                break;
            }
            const unsigned old = __hip_atomic_fetch_add(gcnt + p, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old == 0) break;
            const int hp = mrefit_node(p, P, btl, tnodes, gbi, gh);
            const int pp = mpar_int[p];
            if (pp < 0) {
                *root_height = (unsigned)hp;
                break;
            }
            const int4 PP = mnode[pp];
            if ((unsigned)PP.z / kMSuper != (unsigned)PP.w / kMSuper) {
                defer(pp, p);
                break;
            }
            p = pp;
            P = PP;
        }
    }
}

// Pass 3: the nodes spanning super-chunks, one workgroup, arrivals from pass 2's queue.
__device__ __forceinline__ void mrefit_top_body(const int4 *mnode, const int *mpar_int, const BVHNodeGPU *btl,
                                                TNode *tnodes, float *gbi, int *gh, const int2 *xlist,
                                                const unsigned *xcount, unsigned xcap, unsigned *gcnt,
                                                unsigned *root_height, unsigned *err) {
    const unsigned c = *xcount;
    const unsigned count = c < xcap ? c : xcap;
    for (unsigned e = threadIdx.x; e < count; e += blockDim.x) {
        int p = xlist[e].x;
        for (int steps = 0;; steps++) {
            if (steps > kMaxClimb) {
                atomicOr(err, 4u);
                break;
            }
            const unsigned old = __hip_atomic_fetch_add(gcnt + p, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old == 0) break;
            const int hp = mrefit_node(p, mnode[p], btl, tnodes, gbi, gh);
            const int pp = mpar_int[p];
            if (pp < 0) {
                *root_height = (unsigned)hp;
                break;
            }
            p = pp;
        }
    }
}

}  // namespace ptgpu
