// pt_engine.cpp — the wavefront render engine behind pt_render, pt_render_batch and
// the render server.
//
// A pool of path slots is shared by any number of jobs. A job's work is its samples (every
// sample has sampler streams of its own, so the samples of a pixel are independent): each
// of its path slots runs one sample at a time and then takes the job's next sample, and a
// sample's colour goes to the job's sample buffer; when every sample is done the pixels
// are summed in sample order (PathTracing::Render's accumulation,
// Core/Integrator/PathTracing.cpp:38-198). Every iteration runs one persistent trace launch
// over the three ray queues and one shading launch over the live paths; both read their
// counts from device memory, so an iteration is a fixed sequence of launches that is
// captured once as a hipGraph and replayed several times between host synchronisations.
// Jobs are admitted into free slots between replays (continuous batching) and retired when
// all their samples are done. Once every job is admitted, every sample handed out and few
// paths are live, one path-completion launch runs them to the end; a large single-job
// render runs as two halves of its pixels on two engines, whose launches overlap.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "pt_engine.h"

using namespace pt;

// With PT_RENDER_TWO_STREAMS, images of at least this many pixels render in two halves
// (run_halves).
static constexpr int64_t kHalvesMinPixels = (int64_t)1 << 19;
// Path slots of pt_render's engines: a job runs on up to max(its pixels, this many) slots.
// 8M slots per engine (about 4 GB of path state and queues): on the single large scene a
// half of the frame ran at 1749 / 1887 / 1904 / 1934 / 1935 Mrays/s with 2M / 4M / 6M / 8M /
// 12M slots, as fewer, larger trace launches spend less of their time in the slowest rays.
#ifndef PT_SINGLE_JOB_SLOTS
#define PT_SINGLE_JOB_SLOTS ((int64_t)1 << 23)
#endif
static constexpr int64_t kSingleJobSlots = PT_SINGLE_JOB_SLOTS;
// Most parts (engines and streams) of an image rendered with PT_RENDER_TWO_STREAMS: two,
// or three with PT_RENDER_THREE_STREAMS.
static constexpr int kRenderParts = 3;

namespace pt {

// Entry 0: the length (as float bits) and MicroJitteredSampler's mu = 2.5 / sqrt(length),
// whose unqualified sqrt is the double C function (MicroJitteredSampler.cpp:20-26); entry
// 1 + i: point i of Math::Halton in bases 2 and 3, accumulated digit by digit in float as
// Halton::Initialize does (f /= base, r += f * digit; Math/Halton.cpp:18-37).
std::vector<float2> micro_jitter_table(int length)
{
    std::vector<float2> t((size_t)length + 1);
    t[0] = make_float2(bits_to_float((uint32_t)length), (float)((double)2.5f / std::sqrt((double)(float)length)));
    for (int i = 0; i < length; ++i) {
        float h[2];
        const unsigned bases[2] = {2u, 3u};
        for (int k = 0; k < 2; ++k) {
            unsigned rest = (unsigned)i;
            float f = 1.0f, r = 0.0f;
            for (; rest > 0; rest /= bases[k]) {
                f /= (float)bases[k];
                r += f * (float)(rest % bases[k]);
            }
            h[k] = r;
        }
        t[1 + (size_t)i] = make_float2(h[0], h[1]);
    }
    return t;
}

int check_job(pt_context* ctx, const pt_job_desc& j)
{
    if (!j.scene || !j.bvh || !j.rgb_out) return pt_fail(ctx, PT_ERR_ARG, "job needs a scene, a BVH and an output");
    if (j.bvh->scene != j.scene) return pt_fail(ctx, PT_ERR_ARG, "BVH belongs to another scene");
    const pt_render_desc& d = j.render;
    if (d.width <= 0 || d.height <= 0 || d.spp <= 0 || d.max_bounces < 0 ||
        (d.integrator != PT_INTEGRATOR_MIS && d.integrator != PT_INTEGRATOR_SIMPLE))
        return pt_fail(ctx, PT_ERR_ARG, "bad render settings");
    if ((int64_t)d.width * d.height > (1 << 28)) return pt_fail(ctx, PT_ERR_ARG, "image too large");
    if ((int64_t)d.spp > ((int64_t)1 << 31) - 1) return pt_fail(ctx, PT_ERR_ARG, "too many samples per pixel");
    if (d.sampler_length < 0 || d.sampler_length > PT_MAX_SAMPLER_LENGTH)
        return pt_fail(ctx, PT_ERR_ARG, "sampler length outside [0, PT_MAX_SAMPLER_LENGTH]");
    return PT_OK;
}

int prepare_output(pt_context* ctx, EngineJob& j)
{
    if (j.out_on_device) {
        j.out_d = j.user_out;
        j.own_out = false;
        return PT_OK;
    }
    PT_HIP(ctx, hipMalloc((void**)&j.out_d, 3 * sizeof(float) * (size_t)j.npix));
    j.own_out = true;
    return PT_OK;
}

int finish_output(pt_context* ctx, EngineJob& j, hipStream_t st)
{
    if (j.own_out) {
        PT_HIP(ctx, hipMemcpyAsync(j.user_out, j.out_d, 3 * sizeof(float) * (size_t)j.npix, hipMemcpyDeviceToHost, st));
        PT_HIP(ctx, hipStreamSynchronize(st));
        (void)hipFree(j.out_d);
        j.out_d = nullptr;
        j.own_out = false;
    }
    return PT_OK;
}

EngineJob make_job(const pt_job_desc& d, int64_t id)
{
    EngineJob j;
    j.id = id;
    j.scene = d.scene;
    j.bvh = d.bvh;
    j.desc = d.render;
    j.user_out = d.rgb_out;
    j.out_on_device = d.out_on_device != 0;
    j.npix = (int64_t)d.render.width * d.render.height;
    j.items = j.npix * d.render.spp;
    j.slots = 0; // set at admission (pt_engine::job_slots)
    j.bytes = std::min(j.items, std::max(j.npix, kMinJobSlots)) * kPathBytes + j.npix * 12 + j.items * kSampleBytes;
    j.submitted = std::chrono::steady_clock::now();
    return j;
}

int run_jobs(pt_engine& E, std::vector<EngineJob>& jobs, int iters_per_sync, pt_batch_stats* stats)
{
    pt_context* ctx = E.ctx;
    // Admission order: largest job first (samples W x H x spp), so that the long jobs start
    // early and the small ones fill the path slots around them and at the end; a job that
    // does not fit yet is passed over for later ones that do (first fit). The images do not
    // depend on the order.
    std::vector<EngineJob*> pending;
    pending.reserve(jobs.size());
    for (EngineJob& j : jobs) pending.push_back(&j);
    std::stable_sort(pending.begin(), pending.end(), [](const EngineJob* a, const EngineJob* b) {
        return a->npix * a->desc.spp > b->npix * b->desc.spp;
    });
    std::vector<EngineJob*> running;
    hipEvent_t ev0, ev1;
    PT_HIP(ctx, hipEventCreate(&ev0));
    PT_HIP(ctx, hipEventCreate(&ev1));
    PT_HIP(ctx, hipEventRecord(ev0, E.st));
    const auto t0 = std::chrono::steady_clock::now();
    int64_t replays = 0;
    // PT_ADMIT_TAILS: the next job waits until every running job has handed out its samples
    auto admit_open = [&]() {
        if (E.tune.admission != PT_ADMIT_TAILS) return true;
        for (const EngineJob* r : running)
            if (!E.job_exhausted(*r)) return false;
        return true;
    };
    while (!pending.empty() || !running.empty()) {
        for (size_t i = 0; i < pending.size() && (int64_t)running.size() < E.max_jobs && admit_open();) {
            EngineJob& j = *pending[i];
            j.slots = E.job_slots(j);
            if (E.largest_free() < j.slots || !E.width_ok(pt_engine::job_width(j))) {
                ++i;
                continue;
            }
            int rc = prepare_output(ctx, j);
            if (rc == PT_OK) rc = E.start(j);
            if (rc != PT_OK) return rc;
            running.push_back(&j);
            pending.erase(pending.begin() + i);
        }
        if (running.empty()) return pt_fail(ctx, PT_ERR_BUDGET, "job larger than the engine's path capacity");
        int rc = E.run(iters_per_sync);
        if (rc != PT_OK) return rc;
        ++replays;
        // every job admitted, every sample handed out and few paths left: the remaining
        // paths in one launch
        bool exhausted = pending.empty();
        for (const EngineJob* j : running) exhausted = exhausted && E.job_exhausted(*j);
        if (exhausted && E.live_h > 0 && E.live_h <= E.finish_threshold) {
            rc = E.finish_live();
            if (rc != PT_OK) return rc;
            for (const EngineJob* j : running)
                if (!E.job_done(*j)) return pt_fail(ctx, PT_ERR_STATE, "path completion left samples unfinished");
        }
        for (size_t i = 0; i < running.size();) {
            EngineJob& j = *running[i];
            if (E.job_done(j)) {
                rc = E.resolve(j);
                if (rc == PT_OK) rc = finish_output(ctx, j, E.st);
                if (rc != PT_OK) return rc;
                E.retire(j);
                running.erase(running.begin() + i);
            } else {
                ++i;
            }
        }
    }
    PT_HIP(ctx, hipEventRecord(ev1, E.st));
    PT_HIP(ctx, hipEventSynchronize(ev1));
    float ms = 0.0f;
    PT_HIP(ctx, hipEventElapsedTime(&ms, ev0, ev1));
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (stats) {
        stats->queries = E.queries();
        stats->iterations = E.iterations;
        stats->elapsed_ms = ms;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->syncs = replays;
    }
    return PT_OK;
}

// pt_render_batch with solo jobs: the jobs of at least solo_min samples render one at a
// time, largest first, on the single-job engine S (the single-job kernels, the scene as
// kernel arguments, S's stream), while the other jobs share the multi-job engine E as in
// run_jobs. Both engines' graph replays are queued before either is waited for, so the GPU
// overlaps the two streams' launches (one engine's kernels fill the CUs that the other's
// trace launch leaves idle while it waits for its slowest rays). Every job's image is the
// same as in one piece.
int run_jobs_solo(pt_engine& E, pt_engine& S, std::vector<EngineJob>& jobs, int64_t solo_min, int iters_per_sync,
                  pt_batch_stats* stats)
{
    pt_context* ctx = E.ctx;
    std::vector<EngineJob*> solo_q, pending;
    for (EngineJob& j : jobs) (j.items >= solo_min ? solo_q : pending).push_back(&j);
    auto by_size = [](const EngineJob* a, const EngineJob* b) {
        return a->npix * a->desc.spp > b->npix * b->desc.spp;
    };
    std::stable_sort(solo_q.begin(), solo_q.end(), by_size);
    std::stable_sort(pending.begin(), pending.end(), by_size);
    // the engines' streams must not start before the context's stream work is done
    PT_HIP(ctx, hipDeviceSynchronize());
    S.iterations = 0;
    PT_HIP(ctx, hipMemsetAsync(S.queries_d, 0, sizeof(unsigned long long), S.st));
    const auto t0 = std::chrono::steady_clock::now();
    size_t next_solo = 0;
    EngineJob* solo = nullptr;
    std::vector<EngineJob*> running;
    int64_t replays = 0;
    while (solo || next_solo < solo_q.size() || !pending.empty() || !running.empty()) {
        if (!solo && next_solo < solo_q.size()) {
            EngineJob& j = *solo_q[next_solo++];
            j.slots = S.job_slots(j);
            int rc = prepare_output(ctx, j);
            if (rc == PT_OK) rc = S.start(j);
            if (rc != PT_OK) return rc;
            S.live_h = (int)j.slots;
            solo = &j;
        }
        for (size_t i = 0; i < pending.size() && (int64_t)running.size() < E.max_jobs;) {
            EngineJob& j = *pending[i];
            j.slots = E.job_slots(j);
            if (E.largest_free() < j.slots || !E.width_ok(pt_engine::job_width(j))) {
                ++i;
                continue;
            }
            int rc = prepare_output(ctx, j);
            if (rc == PT_OK) rc = E.start(j);
            if (rc != PT_OK) return rc;
            E.live_h += (int)j.slots;
            running.push_back(&j);
            pending.erase(pending.begin() + i);
        }
        if (!solo && running.empty()) return pt_fail(ctx, PT_ERR_BUDGET, "job larger than the engine's path capacity");
        // a solo job's last live paths, and the shared engine's once every job is admitted
        // and exhausted, complete in one launch
        const bool s_last = solo && S.job_exhausted(*solo) && S.live_h > 0 && S.live_h <= S.finish_threshold;
        bool e_exhausted = pending.empty();
        for (const EngineJob* j : running) e_exhausted = e_exhausted && E.job_exhausted(*j);
        const bool e_last = !running.empty() && e_exhausted && E.live_h > 0 && E.live_h <= E.finish_threshold;
        int rc = PT_OK;
        if (solo) rc = s_last ? S.launch_finish() : S.launch(iters_per_sync);
        if (rc == PT_OK && !running.empty()) rc = e_last ? E.launch_finish() : E.launch(iters_per_sync);
        if (rc != PT_OK) return rc;
        ++replays;
        if (solo) {
            rc = S.wait();
            if (rc != PT_OK) return rc;
            if (S.job_done(*solo)) {
                rc = S.resolve(*solo);
                if (rc == PT_OK) rc = finish_output(ctx, *solo, S.st);
                if (rc != PT_OK) return rc;
                S.retire(*solo);
                solo = nullptr;
            } else if (s_last || S.live_h == 0) {
                return pt_fail(ctx, PT_ERR_STATE, "solo job ended with samples unfinished");
            }
        }
        if (!running.empty()) {
            rc = E.wait();
            if (rc != PT_OK) return rc;
            if (e_last)
                for (const EngineJob* j : running)
                    if (!E.job_done(*j)) return pt_fail(ctx, PT_ERR_STATE, "path completion left samples unfinished");
            for (size_t i = 0; i < running.size();) {
                EngineJob& j = *running[i];
                if (E.job_done(j)) {
                    rc = E.resolve(j);
                    if (rc == PT_OK) rc = finish_output(ctx, j, E.st);
                    if (rc != PT_OK) return rc;
                    E.retire(j);
                    running.erase(running.begin() + i);
                } else {
                    ++i;
                }
            }
        }
    }
    PT_HIP(ctx, hipStreamSynchronize(S.st));
    PT_HIP(ctx, hipStreamSynchronize(E.st));
    const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->queries = E.queries() + S.queries();
        stats->iterations = E.iterations + S.iterations;
        stats->elapsed_ms = wall;
        stats->wall_ms = wall;
        stats->syncs = replays;
    }
    return PT_OK;
}

// A large single-job render in two halves of its pixels, one per engine and stream, whose
// launches the GPU overlaps: one half's kernels fill the CUs that the other's leave idle
// at the end of each launch (the trace launches wait for their slowest rays). Each half is
// the same job restricted to a pixel range, so every pixel is computed as in one piece.
int run_halves(pt_context* ctx, pt_engine* const* E, int n, const pt_job_desc& jd, int iters_per_sync,
               pt_batch_stats* stats)
{
    EngineJob full = make_job(jd, 0);
    int rc = prepare_output(ctx, full);
    if (rc != PT_OK) return rc;
    EngineJob h[kRenderParts];
    for (int i = 0; i < n; ++i) {
        h[i] = full;
        h[i].pixel_base = full.npix * i / n;
        h[i].npix = full.npix * (i + 1) / n - h[i].pixel_base;
        h[i].own_out = false; // the image belongs to `full`
    }
    // the engines' streams must not start before the output exists (ctx stream work done)
    PT_HIP(ctx, hipDeviceSynchronize());
    const auto t0 = std::chrono::steady_clock::now();
    bool done[kRenderParts] = {};
    for (int i = 0; i < n; ++i) {
        E[i]->iterations = 0;
        PT_HIP(ctx, hipMemsetAsync(E[i]->queries_d, 0, sizeof(unsigned long long), E[i]->st));
        h[i].items = h[i].npix * h[i].desc.spp;
        h[i].slots = 0;
        rc = E[i]->start(h[i]);
        if (rc != PT_OK) return rc;
        E[i]->live_h = (int)h[i].slots;
    }
    auto all_done = [&] {
        for (int i = 0; i < n; ++i)
            if (!done[i]) return false;
        return true;
    };
    while (!all_done()) {
        // the parts end together: their completion launches start in the same round, once
        // every sample of every part is handed out and the live paths of all fit one
        // completion launch
        int live = 0;
        bool exhausted = true;
        for (int i = 0; i < n; ++i) {
            if (done[i]) continue;
            live += E[i]->live_h;
            exhausted = exhausted && E[i]->job_exhausted(h[i]);
        }
        const bool last = exhausted && live > 0 && live <= E[0]->finish_threshold;
        for (int i = 0; i < n; ++i) {
            if (done[i]) continue;
            rc = last ? E[i]->launch_finish() : E[i]->launch(iters_per_sync);
            if (rc != PT_OK) return rc;
        }
        for (int i = 0; i < n; ++i) {
            if (done[i]) continue;
            rc = E[i]->wait();
            if (rc != PT_OK) return rc;
            done[i] = E[i]->job_done(h[i]);
            if (!done[i] && (last || E[i]->live_h == 0))
                return pt_fail(ctx, PT_ERR_STATE, "render ended with samples unfinished");
            if (done[i]) {
                rc = E[i]->resolve(h[i]);
                if (rc != PT_OK) return rc;
            }
        }
    }
    for (int i = 0; i < n; ++i) PT_HIP(ctx, hipStreamSynchronize(E[i]->st));
    const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; i < n; ++i) E[i]->retire(h[i]);
    rc = finish_output(ctx, full, E[0]->st);
    if (rc != PT_OK) return rc;
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->queries = 0;
        stats->iterations = 0;
        for (int i = 0; i < n; ++i) {
            stats->queries += E[i]->queries();
            stats->iterations = std::max(stats->iterations, E[i]->iterations);
        }
        stats->elapsed_ms = wall;
        stats->wall_ms = wall;
        stats->syncs = 0;
    }
    return PT_OK;
}

} // namespace pt

extern "C" {

int pt_engine_create(pt_context* ctx, int64_t max_paths, int32_t max_jobs, int32_t multi_job, pt_engine** out)
{
    if (!ctx || !out || max_paths <= 0 || max_paths > ((int64_t)1 << 30) || max_jobs <= 0)
        return pt_fail(ctx, PT_ERR_ARG, "bad engine size");
    *out = nullptr;
    pt_engine* e = new pt_engine();
    const int rc = e->init(ctx, max_paths, max_jobs, multi_job != 0);
    if (rc != PT_OK) {
        delete e;
        return rc;
    }
    *out = e;
    return PT_OK;
}

int pt_engine_destroy(pt_engine* e)
{
    if (!e) return PT_ERR_ARG;
    delete e;
    return PT_OK;
}

int pt_render_aov(pt_scene* s, pt_bvh* b, int32_t aov, const pt_render_desc* d, float* rgb_out,
                  int32_t out_on_device, pt_render_stats* stats)
{
    if (!s || !b || !d || !rgb_out) return PT_ERR_ARG;
    pt_context* ctx = s->ctx;
    if (aov != PT_AOV_DENOISING_ALBEDO && aov != PT_AOV_DENOISING_NORMAL && aov != PT_AOV_UV)
        return pt_fail(ctx, PT_ERR_ARG, "unknown AOV");
    pt_job_desc jd;
    jd.scene = s;
    jd.bvh = b;
    jd.render = *d;
    jd.render.integrator = PT_INTEGRATOR_MIS; // not read by the AOV passes
    jd.rgb_out = rgb_out;
    jd.out_on_device = out_on_device;
    int rc = check_job(ctx, jd);
    if (rc != PT_OK) return rc;
    PT_HIP(ctx, hipSetDevice(ctx->device));
    EngineJob j = make_job(jd, 0);
    rc = prepare_output(ctx, j);
    if (rc != PT_OK) return rc;
    unsigned long long* q_d = nullptr;
    hipError_t e = hipMalloc((void**)&q_d, sizeof(unsigned long long));
    const auto t0 = std::chrono::steady_clock::now();
    if (e == hipSuccess) e = hipMemsetAsync(q_d, 0, sizeof(unsigned long long), ctx->stream);
    if (e == hipSuccess)
        e = launch_render_aov(pt_engine::scene_view(s, b), d->width, d->height, d->spp, d->max_bounces, d->seed,
                              aov == PT_AOV_UV ? 2 : (aov == PT_AOV_DENOISING_NORMAL ? 1 : 0), b->view.nodes8 ? b->view.width : 2,
                              j.out_d, q_d, ctx->stream);
    unsigned long long q = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&q, q_d, sizeof q, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (q_d) (void)hipFree(q_d);
    if (e != hipSuccess) {
        if (j.own_out) (void)hipFree(j.out_d);
        return pt_fail(ctx, e == hipErrorOutOfMemory ? PT_ERR_OOM : PT_ERR_HIP, hipGetErrorString(e));
    }
    rc = finish_output(ctx, j, ctx->stream);
    if (rc != PT_OK) return rc;
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->queries = (int64_t)q;
        stats->iterations = 1;
        stats->elapsed_ms = wall;
    }
    return PT_OK;
}

int pt_render(pt_scene* s, pt_bvh* b, const pt_render_desc* d, float* rgb_out, int32_t out_on_device,
              pt_render_stats* stats)
{
    if (!s || !b || !d || !rgb_out) return PT_ERR_ARG;
    pt_context* ctx = s->ctx;
    pt_job_desc jd;
    jd.scene = s;
    jd.bvh = b;
    jd.render = *d;
    jd.rgb_out = rgb_out;
    jd.out_on_device = out_on_device;
    int rc = check_job(ctx, jd);
    if (rc != PT_OK) return rc;
    const int64_t npix = (int64_t)d->width * d->height;
    // path slots of a single-job engine: every sample, up to max(pixels, kSingleJobSlots)
    auto slots_for = [&](int64_t n) { return std::min(n * d->spp, std::max(n, kSingleJobSlots)); };
    // On request, large images render in two halves on two engines (streams), whose
    // launches overlap.
    if ((d->flags & PT_RENDER_TWO_STREAMS) && npix >= kHalvesMinPixels &&
        !(d->flags & (PT_RENDER_PLAIN_LAUNCHES | PT_RENDER_TIME_KERNELS))) {
        const int n = (d->flags & PT_RENDER_THREE_STREAMS) ? 3 : 2;
        const int64_t cap = slots_for(npix - npix * (n - 1) / n);
        pt_engine** slots[kRenderParts];
        pt_engine* E2[kRenderParts];
        slots[0] = &ctx->engine;
        for (int i = 1; i < n; ++i) slots[i] = &ctx->engine2[i - 1];
        for (int i = 0; i < n; ++i) {
            if (!*slots[i] || (*slots[i])->capacity < cap) {
                delete *slots[i];
                *slots[i] = nullptr;
                rc = pt_engine_create(ctx, cap, 1, 0, slots[i]);
                if (rc != PT_OK) return rc;
                (*slots[i])->iters_per_sync = 8;
                (*slots[i])->min_job_slots = kSingleJobSlots;
            }
            E2[i] = *slots[i];
        }
        const bool log = (d->flags & PT_RENDER_LAUNCH_LOG) != 0;
        for (int i = 0; i < n; ++i) {
            E2[i]->use_trace_tuning(b->trace_tune);
            rc = E2[i]->set_logging(log, (d->flags & PT_RENDER_TRACE_STATS) != 0);
            if (rc != PT_OK) return rc;
        }
        pt_batch_stats bs;
        rc = run_halves(ctx, E2, n, jd, E2[0]->iters_per_sync, &bs);
        if (rc != PT_OK) return rc;
        if (stats) {
            std::memset(stats, 0, sizeof *stats);
            stats->queries = bs.queries;
            stats->iterations = bs.iterations;
            stats->elapsed_ms = bs.elapsed_ms;
        }
        if (log) {
            for (int i = 0; i < n; ++i) {
                if (stats) rc = E2[i]->read_log(stats);
                if (rc == PT_OK) rc = E2[i]->set_logging(false, false);
                if (rc != PT_OK) return rc;
            }
        }
        return PT_OK;
    }
    // the context keeps a single-job engine, grown on demand
    if (!ctx->engine || ctx->engine->capacity < slots_for(npix)) {
        delete ctx->engine;
        ctx->engine = nullptr;
        rc = pt_engine_create(ctx, slots_for(npix), 1, 0, &ctx->engine);
        if (rc != PT_OK) return rc;
        ctx->engine->iters_per_sync = 8;
        ctx->engine->min_job_slots = kSingleJobSlots;
    }
    pt_engine& E = *ctx->engine;
    E.use_trace_tuning(b->trace_tune);
    E.iterations = 0;
    PT_HIP(ctx, hipMemsetAsync(E.queries_d, 0, sizeof(unsigned long long), E.st));
    std::vector<EngineJob> jobs(1, make_job(jd, 0));
    if (d->flags & PT_RENDER_PLAIN_LAUNCHES) {
        EngineJob& j = jobs[0];
        rc = prepare_output(ctx, j);
        if (rc == PT_OK) rc = E.start(j);
        if (rc != PT_OK) return rc;
        const auto t0 = std::chrono::steady_clock::now();
        bool live = true;
        while (live) {
            rc = E.run_plain(&live);
            if (rc != PT_OK) return rc;
        }
        rc = E.resolve(j);
        if (rc == PT_OK) rc = finish_output(ctx, j, E.st);
        const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        E.retire(j);
        if (rc != PT_OK) return rc;
        if (stats) {
            std::memset(stats, 0, sizeof *stats);
            stats->queries = E.queries();
            stats->iterations = E.iterations;
            stats->elapsed_ms = wall;
        }
        return PT_OK;
    }
    if (d->flags & PT_RENDER_TIME_KERNELS) {
        EngineJob& j = jobs[0];
        rc = prepare_output(ctx, j);
        if (rc == PT_OK) rc = E.start(j);
        if (rc != PT_OK) return rc;
        int64_t items[4] = {0, 0, 0, 0}, launches[4] = {0, 0, 0, 0};
        double ms[4] = {0, 0, 0, 0};
        unsigned long long* stats_d = nullptr;
        unsigned long long stats_h[6] = {0, 0, 0, 0, 0, 0};
        if (d->flags & PT_RENDER_TRACE_STATS) {
            PT_HIP(ctx, hipMalloc((void**)&stats_d, sizeof stats_h));
            PT_HIP(ctx, hipMemcpyAsync(stats_d, stats_h, sizeof stats_h, hipMemcpyHostToDevice, E.st));
        }
        const auto t0 = std::chrono::steady_clock::now();
        bool live = true;
        while (live) {
            rc = E.run_timed(items, launches, ms, &live, stats_d);
            if (rc != PT_OK) {
                if (stats_d) (void)hipFree(stats_d);
                return rc;
            }
        }
        if (stats_d) {
            const hipError_t e = hipMemcpy(stats_h, stats_d, sizeof stats_h, hipMemcpyDeviceToHost);
            (void)hipFree(stats_d);
            PT_HIP(ctx, e);
        }
        rc = E.resolve(j);
        if (rc == PT_OK) rc = finish_output(ctx, j, E.st);
        const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        E.retire(j);
        if (rc != PT_OK) return rc;
        if (stats) {
            stats->queries = E.queries();
            stats->iterations = E.iterations;
            stats->elapsed_ms = wall;
            for (int k = 0; k < 4; ++k) {
                stats->kernel_items[k] = items[k];
                stats->kernel_launches[k] = launches[k];
                stats->kernel_ms[k] = ms[k];
                stats->kernel_nodes[k] = k < 3 ? (int64_t)stats_h[2 * k] : 0;
                stats->kernel_tris[k] = k < 3 ? (int64_t)stats_h[2 * k + 1] : 0;
            }
        }
        return PT_OK;
    }
    const bool log = (d->flags & PT_RENDER_LAUNCH_LOG) != 0;
    rc = E.set_logging(log, (d->flags & PT_RENDER_TRACE_STATS) != 0);
    if (rc != PT_OK) return rc;
    pt_batch_stats bs;
    rc = run_jobs(E, jobs, E.iters_per_sync, &bs);
    if (rc != PT_OK) return rc;
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->queries = bs.queries;
        stats->iterations = bs.iterations;
        stats->elapsed_ms = bs.elapsed_ms;
    }
    if (log) {
        if (stats) rc = E.read_log(stats);
        if (rc == PT_OK) rc = E.set_logging(false, false);
        if (rc != PT_OK) return rc;
    }
    return PT_OK;
}

int pt_render_batch(pt_engine* E, int32_t n, const pt_job_desc* jobs, int32_t iterations_per_sync,
                    pt_batch_stats* stats)
{
    if (!E || n < 0 || (n > 0 && !jobs)) return PT_ERR_ARG;
    if (!E->multi) return pt_fail(E->ctx, PT_ERR_STATE, "batch rendering needs a multi-job engine");
    std::vector<EngineJob> js;
    js.reserve(n);
    for (int i = 0; i < n; ++i) {
        const int rc = check_job(E->ctx, jobs[i]);
        if (rc != PT_OK) return rc;
        if (jobs[i].scene->ctx != E->ctx) return pt_fail(E->ctx, PT_ERR_ARG, "job scene on another context");
        js.push_back(make_job(jobs[i], i));

    }
    E->iterations = 0;
    PT_HIP(E->ctx, hipMemsetAsync(E->queries_d, 0, sizeof(unsigned long long), E->st));
    const int iters = iterations_per_sync > 0 ? iterations_per_sync : E->iters_per_sync;
    if (E->tune.solo_log2_samples > 0) {
        const int64_t solo_min = (int64_t)1 << E->tune.solo_log2_samples;
        bool any_solo = false;
        for (const EngineJob& j : js) any_solo = any_solo || j.items >= solo_min;
        if (any_solo) {
            pt_context* ctx = E->ctx;
            if (!ctx->engine || ctx->engine->capacity < kSingleJobSlots) {
                delete ctx->engine;
                ctx->engine = nullptr;
                const int rc = pt_engine_create(ctx, kSingleJobSlots, 1, 0, &ctx->engine);
                if (rc != PT_OK) return rc;
                ctx->engine->min_job_slots = kSingleJobSlots;
            }
            ctx->engine->iters_per_sync = 8;
            ctx->engine->use_trace_tuning(E->tune);
            return run_jobs_solo(*E, *ctx->engine, js, solo_min, iters, stats);
        }
    }
    return run_jobs(*E, js, iters, stats);
}

int pt_engine_set_tuning(pt_engine* E, const pt_tuning* t)
{
    if (!E || !t) return PT_ERR_ARG;
    if (t->iterations_per_sync < 1 || t->iterations_per_sync > 256 || t->trace_blocks_per_cu < 0 ||
        t->trace_blocks_per_cu > 8 || t->trace_chunk < 64 || t->trace_chunk > 4096 || t->shade_blocks_per_cu < 1 ||
        t->shade_blocks_per_cu > 32 || (t->admission != PT_ADMIT_FIT && t->admission != PT_ADMIT_TAILS) ||
        (t->solo_log2_samples != 0 && (t->solo_log2_samples < 10 || t->solo_log2_samples > 40)) ||
        t->threads_per_ray < 0 || t->threads_per_ray > 2 || t->leaf_batch < -1 || t->leaf_batch > 4)
        return pt_fail(E->ctx, PT_ERR_ARG, "tuning out of range");
    E->iters_per_sync = t->iterations_per_sync;
    E->tune.trace_blocks_per_cu = t->trace_blocks_per_cu;
    E->tune.trace_chunk = t->trace_chunk;
    E->tune.shade_blocks_per_cu = t->shade_blocks_per_cu;
    E->tune.admission = t->admission;
    E->tune.solo_log2_samples = t->solo_log2_samples;
    E->tune.threads_per_ray = t->threads_per_ray == 0 ? 1 : t->threads_per_ray;
    E->tune.leaf_batch = t->leaf_batch == 0 ? pt::LaunchTuning{}.leaf_batch : t->leaf_batch;
    E->drop_graphs(); // the launch shapes are captured in the graphs
    return PT_OK;
}

int pt_engine_get_tuning(pt_engine* E, pt_tuning* t)
{
    if (!E || !t) return PT_ERR_ARG;
    t->iterations_per_sync = E->iters_per_sync;
    t->trace_blocks_per_cu = E->tune.trace_blocks_per_cu;
    t->trace_chunk = E->tune.trace_chunk;
    t->shade_blocks_per_cu = E->tune.shade_blocks_per_cu;
    t->admission = E->tune.admission;
    t->solo_log2_samples = E->tune.solo_log2_samples;
    t->threads_per_ray = E->tune.threads_per_ray;
    t->leaf_batch = E->tune.leaf_batch;
    return PT_OK;
}

} // extern "C"
