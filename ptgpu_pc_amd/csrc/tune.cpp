// tune.cpp -- autotuner for the path's launch and build parameters (no reference counterpart).
//
// Every parameter tuned here changes speed only: images, hits and exported trees are bit-identical
// for every value (the tests check each one).  The tuner times real work on the caller's own scene
// and frame with HIP events on the context's stream and keeps the fastest setting:
//   * LBVH sort digit width (4 or 8 bits): full rebuild time;
//   * single-frame render: workgroups per CU (1..4; 1..2 when frames_per_build is 0, see below),
//     shading-pass threshold `regroup` (8, 16, 32), node format (four-wide binary16, four-wide fp32 or
//     binary);
//   * treelet restructuring (none, 5 or 7 leaves): build + optimise time plus `frames_per_build`
//     renders on the resulting tree, the cost a caller pays per rebuild.  frames_per_build 0 means
//     each rebuild runs on another stream beside the previous frame's render: the cost is the render
//     alone, for a build short enough to hide there (kOverlapBuildShare);
//   * batched renders: workgroups per CU (2..4) and `batch_regroup` (16, 32, 64) on the caller's job list.
// Trial renders write to scratch copies of the caller's frame buffers, so the caller's colours,
// seeds and pixels are untouched.
#include <cstring>
#include <vector>

#include "lbvh.hpp"

using namespace ptgpu;

namespace {

// frames_per_build 0: the largest rebuild (+ optimise) time, as a share of the render, that the
// tuner takes to be hidden behind the previous frame's render
constexpr float kOverlapBuildShare = 0.125f;

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    ~Events() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    int init() {
        PT_HIP(hipEventCreate(&a));
        PT_HIP(hipEventCreate(&b));
        return PTGPU_OK;
    }
};

// Fastest of `trials` runs of fn() (each bracketed by events on the context's stream); prep() runs
// before each trial, outside the timed region.
template <class Prep, class Fn>
int best_ms(ptgpu_ctx *ctx, Events &ev, int trials, Prep prep, Fn fn, float *out) {
    float best = 1e30f;
    for (int t = 0; t < trials; t++) {
        PT_TRY(prep());
        PT_HIP(hipEventRecord(ev.a, ctx->stream));
        PT_TRY(fn());
        PT_HIP(hipEventRecord(ev.b, ctx->stream));
        PT_HIP(hipEventSynchronize(ev.b));
        float ms = 0.f;
        PT_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
        if (ms < best) best = ms;
    }
    *out = best;
    return PTGPU_OK;
}

int rebuild(ptgpu_clbvh *h, int treelet) {
    PT_TRY(ptgpu_clbvh_reset(h, h->n));
    PT_TRY(ptgpu_clbvh_build_radix_tree(h));
    PT_TRY(ptgpu_clbvh_build_bvh_tree(h));
    if (treelet) PT_TRY(ptgpu_clbvh_optimize(h, treelet));
    return PTGPU_OK;
}

// Scratch copies of one frame's buffers: trials render into them, restored before every trial.
struct Frame {
    DevBuf colors0, colors, seeds0, seeds, pixels;
    size_t px = 0;
    int init(ptgpu_ctx *ctx, const ptgpu_render_args &a) {
        px = (size_t)a.width * (size_t)a.height;
        PT_TRY(colors0.ensure(px * 12));
        PT_TRY(colors.ensure(px * 12));
        PT_TRY(seeds0.ensure(px * 8));
        PT_TRY(seeds.ensure(px * 8));
        PT_TRY(pixels.ensure(px * 4));
        PT_HIP(hipMemcpyAsync(colors0.p, a.colors, px * 12, hipMemcpyDeviceToDevice, ctx->stream));
        PT_HIP(hipMemcpyAsync(seeds0.p, a.seeds, px * 8, hipMemcpyDeviceToDevice, ctx->stream));
        return PTGPU_OK;
    }
    int restore(ptgpu_ctx *ctx) {
        PT_HIP(hipMemcpyAsync(colors.p, colors0.p, px * 12, hipMemcpyDeviceToDevice, ctx->stream));
        PT_HIP(hipMemcpyAsync(seeds.p, seeds0.p, px * 8, hipMemcpyDeviceToDevice, ctx->stream));
        return PTGPU_OK;
    }
    ptgpu_render_args args(const ptgpu_render_args &a) const {
        ptgpu_render_args r = a;
        r.colors = colors.as<Vec>();
        r.seeds = seeds.as<unsigned>();
        r.pixels = a.pixels ? pixels.as<unsigned>() : nullptr;
        r.ray_counter = nullptr;
        r.trav_stats = nullptr;
        return r;
    }
    void release() {
        colors0.release(); colors.release(); seeds0.release(); seeds.release(); pixels.release();
    }
};

struct SavedTuning {
    int bpc, regroup, wide, bbpc, bregroup;
    explicit SavedTuning(const ptgpu_ctx *c)
        : bpc(c->blocks_per_cu), regroup(c->regroup), wide(c->wide), bbpc(c->batch_blocks_per_cu),
          bregroup(c->batch_regroup) {}
    void restore(ptgpu_ctx *c) const {
        c->blocks_per_cu = bpc;
        c->regroup = regroup;
        c->wide = wide;
        c->batch_blocks_per_cu = bbpc;
        c->batch_regroup = bregroup;
    }
};

int autotune_frame(ptgpu_ctx *ctx, ptgpu_clbvh *h, const ptgpu_render_args *a, int frames_per_build, int trials,
                   ptgpu_tuning *out, Frame &fr, Events &ev) {
    PT_TRY(ev.init());
    PT_TRY(fr.init(ctx, *a));
    const ptgpu_render_args ra = fr.args(*a);
    auto restore = [&]() { return fr.restore(ctx); };
    auto none = []() -> int { return PTGPU_OK; };
    auto render = [&]() { return ptgpu_render(ctx, h, nullptr, &ra); };

    // 1. sort digit width
    float best_build = 1e30f;
    int best_bits = 8;
    for (int bits : {8, 4}) {
        PT_TRY(ptgpu_clbvh_set_digit_bits(h, bits));
        float ms = 0.f;
        PT_TRY(best_ms(ctx, ev, trials, none, [&]() { return rebuild(h, 0); }, &ms));
        if (ms < best_build) {
            best_build = ms;
            best_bits = bits;
        }
    }
    PT_TRY(ptgpu_clbvh_set_digit_bits(h, best_bits));
    PT_TRY(rebuild(h, 0));

    // 2. single-frame launch shape, on the unoptimised tree
    PT_TRY(restore());
    PT_TRY(render());   // first use after a build: derived views and light list
    float best_render = 1e30f;
    int bb = ctx->blocks_per_cu, br = ctx->regroup, bw = ctx->wide;
    // With frames in flight (frames_per_build 0) the next frame's render starts while this one drains:
    // a launch that asks for more than half of a CU's workgroups leaves it no room (measured: three or
    // four per CU render one frame alone as fast, but frames in flight 2-5% slower at C2).
    const int bpc_max = frames_per_build == 0 ? 2 : 4;
    for (int wide : {2, 1, 0})
        for (int bpc = 1; bpc <= bpc_max; bpc++)
            for (int rg : {8, 16, 32}) {
                ctx->wide = wide;
                ctx->blocks_per_cu = bpc;
                ctx->regroup = rg;
                float ms = 0.f;
                PT_TRY(best_ms(ctx, ev, trials, restore, render, &ms));
                if (ms < best_render) {
                    best_render = ms;
                    bb = bpc;
                    br = rg;
                    bw = wide;
                }
            }
    ctx->blocks_per_cu = bb;
    ctx->regroup = br;
    ctx->wide = bw;

    // 3. treelet size: rebuild (+ optimise) once, then frames_per_build renders on that tree; a
    // restructuring has to beat the plain tree by 1% (about the scatter of trial timings).  A build
    // beside a render gets the device only as the render's waves retire, so it hides only when it
    // is short against the render: at 65K triangles and an 8.2 ms render, a 0.75 ms build +
    // optimise completed inside the render and a 1.7 ms one (treelet 7) held up the next frame.
    float best_cost = 1e30f;
    int best_treelet = 0;
    float best_opt_render = best_render, best_opt_build = best_build;
    for (int t : {0, 5, 7}) {
        float bms = 0.f, rms = 0.f;
        PT_TRY(best_ms(ctx, ev, trials, none, [&]() { return rebuild(h, t); }, &bms));
        PT_TRY(restore());
        PT_TRY(render());
        PT_TRY(best_ms(ctx, ev, trials, restore, render, &rms));
        if (frames_per_build == 0 && t != 0 && bms > kOverlapBuildShare * rms) continue;
        const float cost = frames_per_build == 0 ? rms : bms + (float)frames_per_build * rms;
        if (t == 0 ? cost < best_cost : cost < best_cost * (best_treelet == 0 ? 0.99f : 1.f)) {
            best_cost = cost;
            best_treelet = t;
            best_opt_render = rms;
            best_opt_build = bms;
        }
    }
    // leave the handle holding the chosen tree
    PT_TRY(rebuild(h, best_treelet));
    PT_TRY(restore());
    PT_TRY(render());
    PT_HIP(hipStreamSynchronize(ctx->stream));

    out->blocks_per_cu = bb;
    out->regroup = br;
    out->wide = bw;
    out->batch_blocks_per_cu = ctx->batch_blocks_per_cu;
    out->batch_regroup = ctx->batch_regroup;
    out->digit_bits = best_bits;
    out->treelet = best_treelet;
    out->render_ms = best_opt_render;
    out->build_ms = best_opt_build;
    return PTGPU_OK;
}

}  // namespace

extern "C" {

int ptgpu_autotune(ptgpu_ctx *ctx, ptgpu_clbvh *h, const ptgpu_render_args *args, int frames_per_build, int trials,
                   ptgpu_tuning *out) {
    PT_CHECK(ctx && h && args && out, PTGPU_ERR_INVALID, "null argument");
    PT_CHECK(h->ctx == ctx && h->state >= 1, PTGPU_ERR_INVALID, "handle of another context or not constructed");
    PT_CHECK(h->n >= 2, PTGPU_ERR_INVALID, "a scene of at least two shapes is needed to tune");
    PT_CHECK(args->colors && args->seeds && args->camera && args->shapes, PTGPU_ERR_INVALID, "null buffer");
    PT_CHECK(args->width > 0 && args->height > 0 && args->spp >= 1, PTGPU_ERR_INVALID, "bad frame");
    PT_CHECK(frames_per_build >= 0 && trials >= 1 && trials <= 16, PTGPU_ERR_INVALID,
             "frames_per_build %d / trials %d out of range", frames_per_build, trials);
    PT_HIP(hipSetDevice(ctx->device));
    const SavedTuning saved(ctx);
    const int saved_bits = h->digit_bits;
    Frame fr;
    Events ev;
    std::memset(out, 0, sizeof(*out));
    const int st = autotune_frame(ctx, h, args, frames_per_build, trials, out, fr, ev);
    (void)hipStreamSynchronize(ctx->stream);
    fr.release();
    if (st != PTGPU_OK) {
        saved.restore(ctx);
        h->digit_bits = saved_bits;
    }
    return st;
}

int ptgpu_autotune_batch(ptgpu_ctx *ctx, const ptgpu_render_job *jobs, int njobs, int trials, ptgpu_tuning *out) {
    PT_CHECK(ctx && jobs && out && njobs >= 1, PTGPU_ERR_INVALID, "null argument or empty job list");
    PT_CHECK(trials >= 1 && trials <= 16, PTGPU_ERR_INVALID, "trials %d out of range", trials);
    PT_HIP(hipSetDevice(ctx->device));
    for (int j = 0; j < njobs; j++)
        PT_CHECK(jobs[j].args.colors && jobs[j].args.seeds && jobs[j].args.width > 0 && jobs[j].args.height > 0,
                 PTGPU_ERR_INVALID, "job %d: bad frame", j);
    const SavedTuning saved(ctx);
    std::memset(out, 0, sizeof(*out));
    std::vector<Frame> frames((size_t)njobs);
    std::vector<ptgpu_render_job> trial(jobs, jobs + njobs);
    Events ev;
    int st = ev.init();
    for (int j = 0; j < njobs && st == PTGPU_OK; j++) {
        st = frames[(size_t)j].init(ctx, jobs[j].args);
        if (st == PTGPU_OK) trial[(size_t)j].args = frames[(size_t)j].args(jobs[j].args);
    }
    auto restore = [&]() -> int {
        for (auto &f : frames) PT_TRY(f.restore(ctx));
        return PTGPU_OK;
    };
    auto render = [&]() { return ptgpu_render_batch(ctx, trial.data(), njobs, nullptr); };
    float best = 1e30f;
    int bb = ctx->batch_blocks_per_cu, br = ctx->batch_regroup;
    if (st == PTGPU_OK) st = restore();
    if (st == PTGPU_OK) st = render();   // first use: derived views and light lists
    for (int bpc : {2, 3, 4})
        for (int rg : {16, 32, 64}) {
            if (st != PTGPU_OK) break;
            ctx->batch_blocks_per_cu = bpc;
            ctx->batch_regroup = rg;
            float ms = 0.f;
            st = best_ms(ctx, ev, trials, restore, render, &ms);
            if (st == PTGPU_OK && ms < best) {
                best = ms;
                bb = bpc;
                br = rg;
            }
        }
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &f : frames) f.release();
    if (st != PTGPU_OK) {
        saved.restore(ctx);
        return st;
    }
    ctx->batch_blocks_per_cu = bb;
    ctx->batch_regroup = br;
    out->blocks_per_cu = ctx->blocks_per_cu;
    out->regroup = ctx->regroup;
    out->wide = ctx->wide;
    out->batch_blocks_per_cu = bb;
    out->batch_regroup = br;
    out->digit_bits = 0;
    out->treelet = 0;
    out->render_ms = best;
    out->build_ms = 0.f;
    return PTGPU_OK;
}

}  // extern "C"
