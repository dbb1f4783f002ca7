// serve.cpp -- continuous batching of render requests on one GPU (no reference counterpart: the
// served form of the path).
//
// A request is one frame of one scene: `spp` samples from sample 0, the reference's RadianceGPU
// loop.  The server keeps three sets on the device:
//   * resident scenes, keyed by the caller's scene key: shapes + a built LBVH, shared by every request
//     on that scene, kept after their last request until the memory is needed (LRU);
//   * admitted requests: their frame buffers (colors, seeds, pixels, camera) stay resident while the
//     frame accumulates over ticks and until the caller takes the result;
//   * the queue, in arrival order.
// One tick retires the batches that have finished, admits queued requests while their footprint fits
// the HBM budget (head of line first), builds the scenes they bring in (all of them in one batched
// build), and launches one batched render in which every admitted request that is not already in
// flight advances by min(spp_per_tick, remaining) samples.  Two batches may be in flight on two
// streams, so a new batch fills the GPU while the previous one drains; a request is in at most one of
// them.  Splitting a frame's samples over launches gives the same bits as one launch (the running
// mean and the RNG state carry over in colors/seeds), so every request's result equals rendering it
// alone.  Frame buffers of taken results go to a pool for later requests (freeing device memory would
// wait for the batches in flight).
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "device_util.hpp"

namespace {

constexpr size_t kSceneBytesPerPrim = 600;   // estimate before a build: Shape 104 B + LBVH buffers ~454 B

// the host-side mirror (ctypes) depends on this layout
static_assert(sizeof(ptgpu_request) == 128, "ptgpu_request");
static_assert(offsetof(ptgpu_request, camera) == 24 && offsetof(ptgpu_request, width) == 100 &&
                  offsetof(ptgpu_request, h_seeds) == 120,
              "ptgpu_request offsets");

struct Scene {
    unsigned long long key = 0;
    int n = 0;
    unsigned light_cnt = 0;
    ptgpu::DevBuf shapes;
    ptgpu_clbvh *bvh = nullptr;
    int refs = 0;
    size_t bytes = 0;
    unsigned long long last_use = 0;
};

struct Request {
    int id = 0;
    ptgpu_request r{};
};

struct Slot {
    int id = 0;
    ptgpu_request r{};
    Scene *scene = nullptr;
    ptgpu::DevBuf frame;      // colors | seeds | pixels | camera
    size_t bytes = 0;
    int done = 0;             // samples accumulated
    int step = 0;             // samples in the launch in flight
    bool in_flight = false;
    bool finished = false;
};

constexpr int kChannels = 2;

size_t frame_bytes(int w, int h) {
    const size_t px = (size_t)w * (size_t)h;
    return ((px * 12 + 255) & ~size_t(255)) + ((px * 8 + 255) & ~size_t(255)) + ((px * 4 + 255) & ~size_t(255)) +
           256;
}

}  // namespace

struct ptgpu_server {
    ptgpu_ctx *ctx = nullptr;
    size_t budget = 0, used = 0;
    int spp_per_tick = 1;
    int next_id = 1;
    unsigned long long tick_no = 0, builds = 0;
    std::deque<Request> queue;
    std::vector<Slot *> active;                  // admitted, frame accumulating or waiting to be taken
    std::map<int, Slot *> by_id;
    std::map<unsigned long long, Scene *> scenes;
    std::vector<int> finished;                   // finished since the last ptgpu_server_finished call
    ptgpu::DevBuf rays;                          // u64[2] query counters
    std::vector<ptgpu_render_job> jobs;
    std::vector<Slot *> job_slots;
    ptgpu::BatchChannel ch[kChannels];
    hipEvent_t done[kChannels] = {};
    std::vector<Slot *> flight[kChannels];
    unsigned long long launched_at[kChannels] = {};   // launch sequence number, 0 = idle
    unsigned long long launches = 0;
    std::vector<ptgpu::DevBuf> pool;              // frame buffers of taken results
    std::vector<Scene *> unbuilt;                 // admitted this tick, built together after admission
};

namespace ptgpu {

__global__ void k_default_seeds(unsigned *seeds, unsigned n) {
    const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const unsigned s = g > 2u ? g : 2u;
    seeds[2 * g] = s;
    seeds[2 * g + 1] = s;
}

}  // namespace ptgpu

using namespace ptgpu;

static void release_scene(Scene *sc) {
    if (sc->bvh) ptgpu_clbvh_destroy(sc->bvh);
    sc->bvh = nullptr;
    sc->shapes.release();
}

// frees resident scenes nobody uses (least recently used first) until `need` more bytes fit
static void evict_for(ptgpu_server *s, size_t need) {
    while (s->used + need > s->budget && !s->pool.empty()) {   // idle frame buffers first
        s->used -= s->pool.back().bytes;
        s->pool.back().release();
        s->pool.pop_back();
    }
    while (s->used + need > s->budget) {
        Scene *victim = nullptr;
        for (auto &kv : s->scenes)
            if (kv.second->refs == 0 && (!victim || kv.second->last_use < victim->last_use)) victim = kv.second;
        if (!victim) return;
        s->used -= victim->bytes;
        release_scene(victim);
        s->scenes.erase(victim->key);
        delete victim;
    }
}

// a pooled frame buffer of at least `bytes` (the smallest such), moved into *out
static bool take_pooled(ptgpu_server *s, size_t bytes, ptgpu::DevBuf *out) {
    int best = -1;
    for (size_t i = 0; i < s->pool.size(); i++)
        if (s->pool[i].bytes >= bytes && (best < 0 || s->pool[i].bytes < s->pool[(size_t)best].bytes)) best = (int)i;
    if (best < 0) return false;
    *out = s->pool[(size_t)best];
    s->pool.erase(s->pool.begin() + best);
    return true;
}

// undoes the creation of a scene that admission brought in but could not use
static void drop_new_scene(ptgpu_server *s, Scene *sc) {
    s->unbuilt.erase(std::remove(s->unbuilt.begin(), s->unbuilt.end(), sc), s->unbuilt.end());
    s->scenes.erase(sc->key);
    s->used -= sc->bytes;
    release_scene(sc);
    delete sc;
}

static int admit(ptgpu_server *s, const Request &q, bool *fitted) {
    *fitted = false;
    const ptgpu_request &r = q.r;
    auto it = s->scenes.find(r.scene_key);
    Scene *sc = it == s->scenes.end() ? nullptr : it->second;
    PT_CHECK(!sc || (sc->light_cnt == r.lightCnt && sc->n == (int)r.shapeCnt), PTGPU_ERR_INVALID,
             "request %d: scene key %llu is resident with other shapes or lightCnt", q.id, r.scene_key);
    const size_t fb = frame_bytes(r.width, r.height);
    size_t sb = sc ? 0 : (size_t)r.shapeCnt * kSceneBytesPerPrim;
    // A pooled frame buffer needs no new memory.  It leaves the pool before the eviction below, which
    // frees pooled buffers first and must not free this one.
    ptgpu::DevBuf pooled;
    const bool have_pooled = take_pooled(s, fb, &pooled);
    const size_t fresh = have_pooled ? 0 : fb;
    // The request's own scene may be idle (refs 0) and the least recently used: it must survive.
    if (sc) sc->refs++;
    evict_for(s, fresh + sb);
    if (sc) sc->refs--;
    if (sc && sc->refs == 0 && s->used + fresh > s->budget) {
        // Only the request's own idle scene stands in the way (its built size can exceed the estimate
        // that let the request in): drop it and bring the scene in anew, as for a first request.
        s->used -= sc->bytes;
        release_scene(sc);
        s->scenes.erase(sc->key);
        delete sc;
        sc = nullptr;
        sb = (size_t)r.shapeCnt * kSceneBytesPerPrim;
        evict_for(s, fresh + sb);
    }
    if (s->used + fresh + sb > s->budget) {   // wait for memory
        if (have_pooled) s->pool.push_back(pooled);
        return PTGPU_OK;
    }
    hipStream_t st = s->ctx->stream;
    bool created = false;
    if (!sc) {
        sc = new Scene();
        sc->key = r.scene_key;
        sc->n = (int)r.shapeCnt;
        sc->light_cnt = r.lightCnt;
        int e = sc->shapes.ensure((size_t)r.shapeCnt * sizeof(Shape));
        if (e == PTGPU_OK) {
            hipError_t he = hipMemcpyAsync(sc->shapes.p, r.h_shapes, (size_t)r.shapeCnt * sizeof(Shape),
                                           hipMemcpyHostToDevice, st);
            if (he != hipSuccess) e = hip_fail(he, "hipMemcpyAsync", __FILE__, __LINE__);
        }
        if (e != PTGPU_OK) {
            release_scene(sc);
            delete sc;
            if (have_pooled) s->pool.push_back(pooled);
            return e;
        }
        sc->bytes = sb;   // the estimate until the build reports the real footprint
        s->used += sc->bytes;
        s->scenes[sc->key] = sc;
        s->unbuilt.push_back(sc);
        created = true;
    }
    Slot *sl = new Slot();
    sl->id = q.id;
    sl->r = r;
    sl->scene = sc;
    if (have_pooled) {
        sl->frame = pooled;
    } else {
        int e = sl->frame.ensure(fb);
        if (e != PTGPU_OK) {
            delete sl;
            if (created) drop_new_scene(s, sc);   // refs 0: nothing else holds it yet
            return e;
        }
        s->used += sl->frame.bytes;
    }
    sl->bytes = sl->frame.bytes;
    const size_t px = (size_t)r.width * (size_t)r.height;
    char *base = static_cast<char *>(sl->frame.p);
    const size_t oc = 0, os = (px * 12 + 255) & ~size_t(255), op = os + ((px * 8 + 255) & ~size_t(255));
    const size_t ocam = op + ((px * 4 + 255) & ~size_t(255));
    (void)oc;
    hipError_t he;
    if (r.h_seeds) {
        he = hipMemcpyAsync(base + os, r.h_seeds, px * 8, hipMemcpyHostToDevice, st);
    } else {
        hipLaunchKernelGGL(k_default_seeds, dim3((unsigned)((px + 255) / 256)), dim3(256), 0, st,
                           reinterpret_cast<unsigned *>(base + os), (unsigned)px);
        he = hipGetLastError();
    }
    if (he == hipSuccess) he = hipMemcpyAsync(base + ocam, &sl->r.camera, sizeof(Camera), hipMemcpyHostToDevice, st);
    if (he != hipSuccess) {   // the frame buffer stays counted in `used`, in the pool
        s->pool.push_back(sl->frame);
        delete sl;
        if (created) drop_new_scene(s, sc);
        return hip_fail(he, "admit: upload", __FILE__, __LINE__);
    }
    sc->refs++;
    sc->last_use = s->tick_no;
    s->active.push_back(sl);
    s->by_id[sl->id] = sl;
    *fitted = true;
    return PTGPU_OK;
}

// Builds every scene admitted since the last build in one batched build (one launch per stage for
// all of them, one wait); their requests are launched from the same tick on.
static int build_unbuilt(ptgpu_server *s) {
    if (s->unbuilt.empty()) return PTGPU_OK;
    const size_t k = s->unbuilt.size();
    std::vector<Shape *> shapes(k);
    std::vector<int> counts(k);
    std::vector<ptgpu_clbvh *> out(k, nullptr);
    for (size_t i = 0; i < k; i++) {
        shapes[i] = s->unbuilt[i]->shapes.as<Shape>();
        counts[i] = s->unbuilt[i]->n;
    }
    PT_TRY(ptgpu_clbvh_create_batch(s->ctx, shapes.data(), counts.data(), (int)k, out.data()));
    for (size_t i = 0; i < k; i++) {
        Scene *sc = s->unbuilt[i];
        sc->bvh = out[i];
        size_t hb = 0;
        PT_TRY(ptgpu_clbvh_device_bytes(sc->bvh, &hb));
        const size_t real = sc->shapes.bytes + hb;
        s->used = s->used - sc->bytes + real;
        sc->bytes = real;
        s->builds++;
    }
    s->unbuilt.clear();
    return PTGPU_OK;
}

extern "C" {

int ptgpu_server_create(ptgpu_ctx *ctx, size_t hbm_budget, int spp_per_tick, ptgpu_server **out) {
    PT_CHECK(ctx && out, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(spp_per_tick >= 1, PTGPU_ERR_INVALID, "spp_per_tick %d < 1", spp_per_tick);
    PT_CHECK(hbm_budget >= (1u << 20), PTGPU_ERR_INVALID, "HBM budget below 1 MiB");
    ptgpu_server *s = new ptgpu_server();
    s->ctx = ctx;
    s->budget = hbm_budget;
    s->spp_per_tick = spp_per_tick;
    int e = s->rays.ensure(64);
    if (e == PTGPU_OK) e = fill_u32(ctx->stream, s->rays.p, 0u, 4);
    if (e != PTGPU_OK) {
        delete s;
        return e;
    }
    *out = s;
    return PTGPU_OK;
}

int ptgpu_server_submit(ptgpu_server *s, const ptgpu_request *r, int *id) {
    PT_CHECK(s && r && id, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(r->h_shapes && r->shapeCnt >= 2, PTGPU_ERR_INVALID, "a request needs at least two shapes");
    PT_CHECK(r->width > 0 && r->height > 0 && r->spp >= 1, PTGPU_ERR_INVALID, "bad frame %dx%d x %d spp", r->width,
             r->height, r->spp);
    PT_CHECK(r->max_depth >= 0 && r->max_depth < 1024, PTGPU_ERR_INVALID, "bad max_depth");
    const size_t need = frame_bytes(r->width, r->height) + (size_t)r->shapeCnt * kSceneBytesPerPrim;
    PT_CHECK(need <= s->budget, PTGPU_ERR_UNSUPPORTED, "request needs %zu bytes, above the %zu-byte budget", need,
             s->budget);
    Request q;
    q.id = s->next_id++;
    q.r = *r;
    s->queue.push_back(q);
    *id = q.id;
    return PTGPU_OK;
}

static void retire(ptgpu_server *s, int c) {
    for (Slot *sl : s->flight[c]) {
        sl->done += sl->step;
        sl->step = 0;
        sl->in_flight = false;
        if (sl->done >= sl->r.spp) {
            sl->finished = true;
            s->finished.push_back(sl->id);
            if (--sl->scene->refs < 0) sl->scene->refs = 0;
        }
    }
    s->flight[c].clear();
    s->launched_at[c] = 0;
}

// waits for the oldest batch in flight and retires it; false if none is in flight
static int retire_oldest(ptgpu_server *s, bool *any) {
    int c = -1;
    for (int k = 0; k < kChannels; k++)
        if (s->launched_at[k] && (c < 0 || s->launched_at[k] < s->launched_at[c])) c = k;
    *any = c >= 0;
    if (c < 0) return PTGPU_OK;
    PT_HIP(hipEventSynchronize(s->done[c]));
    retire(s, c);
    return PTGPU_OK;
}

int ptgpu_server_tick(ptgpu_server *s, int *n_active, int *n_queued) {
    PT_CHECK(s, PTGPU_ERR_INVALID, "null server");
    PT_HIP(hipSetDevice(s->ctx->device));
    s->tick_no++;
    // 1. retire the batches that are done
    for (int c = 0; c < kChannels; c++) {
        if (!s->launched_at[c]) continue;
        const hipError_t q = hipEventQuery(s->done[c]);
        if (q == hipSuccess) retire(s, c);
        else if (q != hipErrorNotReady) return hip_fail(q, "hipEventQuery", __FILE__, __LINE__);
    }
    // 2. admission, head of line first; one wait at the end releases the admitted requests' host data
    bool admitted = false;
    while (!s->queue.empty()) {
        bool fitted = false;
        PT_TRY(admit(s, s->queue.front(), &fitted));
        if (!fitted) break;
        admitted = true;
        s->queue.pop_front();
    }
    PT_TRY(build_unbuilt(s));   // waits for the context's stream, which also covers the uploads above
    if (admitted) PT_HIP(hipStreamSynchronize(s->ctx->stream));
    // 3. one batch of every admitted request not in flight, on a free channel
    int c = -1;
    for (int k = 0; k < kChannels; k++)
        if (!s->launched_at[k]) {
            c = k;
            break;
        }
    s->jobs.clear();
    s->job_slots.clear();
    if (c >= 0) {
        for (Slot *sl : s->active) {
            if (sl->finished || sl->in_flight || !sl->scene->bvh) continue;
            const ptgpu_request &r = sl->r;
            const size_t px = (size_t)r.width * (size_t)r.height;
            char *base = static_cast<char *>(sl->frame.p);
            const size_t os = (px * 12 + 255) & ~size_t(255), op = os + ((px * 8 + 255) & ~size_t(255));
            const size_t ocam = op + ((px * 4 + 255) & ~size_t(255));
            sl->step = r.spp - sl->done < s->spp_per_tick ? r.spp - sl->done : s->spp_per_tick;
            ptgpu_render_job j{};
            j.bvh = sl->scene->bvh;
            j.kd = nullptr;
            j.args.colors = reinterpret_cast<Vec *>(base);
            j.args.seeds = reinterpret_cast<unsigned *>(base + os);
            j.args.pixels = reinterpret_cast<unsigned *>(base + op);
            j.args.camera = reinterpret_cast<const Camera *>(base + ocam);
            j.args.shapes = sl->scene->shapes.as<Shape>();
            j.args.shapeCnt = (unsigned)sl->scene->n;
            j.args.lightCnt = r.lightCnt;
            j.args.width = r.width;
            j.args.height = r.height;
            j.args.currentSample = sl->done;
            j.args.spp = sl->step;
            j.args.max_depth = r.max_depth;
            s->jobs.push_back(j);
            s->job_slots.push_back(sl);
            sl->scene->last_use = s->tick_no;
        }
    }
    if (!s->jobs.empty()) {
        // lanes take pixels in job order: the longest per-pixel chains (most samples this launch, then the
        // largest frames) go first, so the launch does not end waiting on a long chain started late
        std::vector<size_t> order(s->jobs.size());
        for (size_t i = 0; i < order.size(); i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
            const ptgpu_render_args &x = s->jobs[a].args, &y = s->jobs[b].args;
            if (x.spp != y.spp) return x.spp > y.spp;
            return (long long)x.width * x.height > (long long)y.width * y.height;
        });
        std::vector<ptgpu_render_job> sorted;
        sorted.reserve(order.size());
        for (size_t i : order) {
            sorted.push_back(s->jobs[i]);
            s->flight[c].push_back(s->job_slots[i]);
            s->job_slots[i]->in_flight = true;
        }
        if (!s->ch[c].stream) PT_HIP(hipStreamCreateWithFlags(&s->ch[c].stream, hipStreamNonBlocking));
        if (!s->done[c]) PT_HIP(hipEventCreateWithFlags(&s->done[c], hipEventDisableTiming));
        PT_TRY(launch_batch(s->ctx, s->ch[c], sorted.data(), (int)sorted.size(), s->rays.as<unsigned long long>()));
        PT_HIP(hipEventRecord(s->done[c], s->ch[c].stream));
        s->launched_at[c] = ++s->launches;
    } else {
        // 4. nothing new to launch: wait for the oldest batch so the caller's loop makes progress
        bool any = false;
        PT_TRY(retire_oldest(s, &any));
    }
    int act = 0;
    for (Slot *sl : s->active) act += sl->finished ? 0 : 1;
    if (n_active) *n_active = act;
    if (n_queued) *n_queued = (int)s->queue.size();
    return PTGPU_OK;
}

int ptgpu_server_finished(ptgpu_server *s, int *ids, int cap, int *n) {
    PT_CHECK(s && n && (ids || cap == 0), PTGPU_ERR_INVALID, "null argument");
    const int k = (int)s->finished.size() < cap ? (int)s->finished.size() : cap;
    for (int i = 0; i < k; i++) ids[i] = s->finished[(size_t)i];
    s->finished.erase(s->finished.begin(), s->finished.begin() + k);
    *n = k;
    return PTGPU_OK;
}

int ptgpu_server_result(ptgpu_server *s, int id, float *h_colors, unsigned int *h_pixels, unsigned int *h_seeds) {
    PT_CHECK(s, PTGPU_ERR_INVALID, "null server");
    auto it = s->by_id.find(id);
    PT_CHECK(it != s->by_id.end(), PTGPU_ERR_INVALID, "request %d is not admitted (or already taken)", id);
    Slot *sl = it->second;
    PT_CHECK(sl->finished, PTGPU_ERR_INVALID, "request %d has not finished", id);
    const size_t px = (size_t)sl->r.width * (size_t)sl->r.height;
    const char *base = static_cast<const char *>(sl->frame.p);
    const size_t os = (px * 12 + 255) & ~size_t(255), op = os + ((px * 8 + 255) & ~size_t(255));
    hipStream_t st = s->ctx->stream;
    if (h_colors) PT_HIP(hipMemcpyAsync(h_colors, base, px * 12, hipMemcpyDeviceToHost, st));
    if (h_seeds) PT_HIP(hipMemcpyAsync(h_seeds, base + os, px * 8, hipMemcpyDeviceToHost, st));
    if (h_pixels) PT_HIP(hipMemcpyAsync(h_pixels, base + op, px * 4, hipMemcpyDeviceToHost, st));
    PT_HIP(hipStreamSynchronize(st));
    s->pool.push_back(sl->frame);   // still counted in `used`; reused or freed by admission
    sl->frame = ptgpu::DevBuf();
    for (size_t i = 0; i < s->active.size(); i++)
        if (s->active[i] == sl) {
            s->active.erase(s->active.begin() + (long)i);
            break;
        }
    s->by_id.erase(it);
    delete sl;
    return PTGPU_OK;
}

int ptgpu_server_stats(ptgpu_server *s, unsigned long long *rays2, size_t *hbm_used, int *scenes_resident,
                       unsigned long long *builds) {
    PT_CHECK(s, PTGPU_ERR_INVALID, "null server");
    if (rays2) {
        PT_HIP(hipMemcpyAsync(rays2, s->rays.p, 16, hipMemcpyDeviceToHost, s->ctx->stream));
        PT_HIP(hipStreamSynchronize(s->ctx->stream));
    }
    if (hbm_used) *hbm_used = s->used;
    if (scenes_resident) *scenes_resident = (int)s->scenes.size();
    if (builds) *builds = s->builds;
    return PTGPU_OK;
}

int ptgpu_server_destroy(ptgpu_server *s) {
    if (!s) return PTGPU_OK;
    if (s->ctx && s->ctx->stream) (void)hipStreamSynchronize(s->ctx->stream);
    for (Slot *sl : s->active) {
        sl->frame.release();
        delete sl;
    }
    for (auto &b : s->pool) b.release();
    for (int c = 0; c < kChannels; c++) {
        if (s->ch[c].stream) (void)hipStreamSynchronize(s->ch[c].stream);
        s->ch[c].release();
        if (s->ch[c].stream) (void)hipStreamDestroy(s->ch[c].stream);
        if (s->done[c]) (void)hipEventDestroy(s->done[c]);
    }
    for (auto &kv : s->scenes) {
        release_scene(kv.second);
        delete kv.second;
    }
    s->rays.release();
    delete s;
    return PTGPU_OK;
}

}  // extern "C"
