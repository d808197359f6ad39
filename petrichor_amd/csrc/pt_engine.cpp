// pt_engine.cpp — the wavefront render engine behind pt_render, pt_render_batch and
// the render server.
//
// A pool of path slots is shared by any number of jobs (one path per pixel: the
// samples of a pixel run one after another, as in PathTracing::Render,
// Core/Integrator/PathTracing.cpp:38-196, so a pixel's sampler stream is consumed in the
// reference's order). Every iteration runs one persistent trace launch over the three
// ray queues and one shading launch over the live paths; both read their counts from
// device memory, so an iteration is a fixed sequence of launches that is captured once
// as a hipGraph and replayed several times between host synchronisations. Jobs are
// admitted into free slots between replays (continuous batching) and retired when
// all their pixels are done.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "pt_internal.h"

using namespace pt;

namespace {

// Bytes of device memory one path slot uses (path state, queues, hits, live lists).
constexpr int64_t kPathBytes = 3 * 16 + 6 * 16 + 16 + 8 + 6 * 16 + 4 * 4 + 2 * 4 + 3 * (32 + 4);

struct Buffers {
    std::vector<void*> ptrs;
    template <typename T>
    hipError_t alloc(T** p, size_t count)
    {
        *p = nullptr;
        hipError_t e = hipMalloc((void**)p, (count ? count : 1) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
    void release()
    {
        for (void* p : ptrs) (void)hipFree(p);
        ptrs.clear();
    }
    ~Buffers() { release(); }
};

struct EngineJob {
    int64_t id = 0;
    pt_scene* scene = nullptr;
    pt_bvh* bvh = nullptr;
    pt_render_desc desc{};
    float* user_out = nullptr; // caller's buffer
    bool out_on_device = false;
    float* out_d = nullptr;    // device image (== user_out when on device)
    bool own_out = false;
    int64_t base = 0, npix = 0;
    int slot = -1;             // job table index
    int64_t bytes = 0;         // working memory charged to the budget
    std::chrono::steady_clock::time_point submitted, started;
};

} // namespace

struct pt_engine {
    pt_context* ctx = nullptr;
    hipStream_t st = nullptr;
    int64_t capacity = 0;
    int max_jobs = 0;
    bool multi = true;
    Buffers buf;
    PathView ps{};
    float *hit_t = nullptr, *sh_t = nullptr;
    int *hit_k = nullptr, *sh_k = nullptr;
    int* act[2] = {nullptr, nullptr};
    int* counters = nullptr; // [4*set + kind] kind 0 ext, 1 any, 2 closest, 3 live; [8] trace cursor
    float4* q_rays[3] = {nullptr, nullptr, nullptr};
    int32_t* q_path[3] = {nullptr, nullptr, nullptr};
    JobDev* jobs_d = nullptr;
    BvhView* views_d = nullptr;
    int* done_d = nullptr;
    unsigned long long* queries_d = nullptr;
    std::vector<JobDev> jobs_h;
    std::vector<int> done_h;
    std::vector<char> slot_used;
    std::map<int64_t, int64_t> free_ranges; // first path slot -> length
    int set = 1;                            // counter set the next iteration reads
    // single-job mode: the job's views are kernel arguments (captured in the graphs)
    SceneView sc1{};
    JobView job1{};
    int integrator1 = 0;
    hipGraphExec_t graphs[2] = {nullptr, nullptr};
    int64_t iterations = 0;

    Queue queue(int s, int kind) const
    {
        Queue q;
        q.rays = q_rays[kind];
        q.path = q_path[kind];
        q.count = counters + 4 * s + kind;
        return q;
    }

    void drop_graphs()
    {
        for (auto& g : graphs)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
    }

    ~pt_engine()
    {
        drop_graphs();
        if (st) (void)hipStreamDestroy(st);
    }

    int init(pt_context* c, int64_t cap, int jobs, bool multi_job)
    {
        ctx = c;
        capacity = cap;
        max_jobs = jobs;
        multi = multi_job;
        PT_HIP(ctx, hipSetDevice(ctx->device));
        PT_HIP(ctx, hipStreamCreate(&st));
        (void)cu_count();
        const size_t P = (size_t)cap;
#define ALLOC(p, n) PT_HIP(ctx, buf.alloc(&(p), (n)))
        ALLOC(ps.rng, 3 * P);
        ALLOC(ps.sum, P);
        ALLOC(ps.color, P);
        ALLOC(ps.ray_o, P);
        ALLOC(ps.ray_d, P);
        ALLOC(ps.thr, P);
        ALLOC(ps.meta, P);
        ALLOC(ps.pixel, P);
        ALLOC(ps.job, P);
        ALLOC(ps.pend_a, P);
        ALLOC(ps.pend_b, P);
        ALLOC(ps.pend_c, P);
        ALLOC(ps.pend_d, P);
        ALLOC(ps.pend_o, P);
        ALLOC(ps.pend_dir, P);
        ALLOC(hit_t, P);
        ALLOC(hit_k, P);
        ALLOC(sh_t, P);
        ALLOC(sh_k, P);
        ALLOC(act[0], P);
        ALLOC(act[1], P);
        ALLOC(counters, 12);
        for (int k = 0; k < 3; ++k) {
            ALLOC(q_rays[k], 2 * P);
            ALLOC(q_path[k], P);
        }
        ALLOC(jobs_d, (size_t)jobs);
        ALLOC(views_d, (size_t)jobs);
        ALLOC(done_d, (size_t)jobs);
        ALLOC(queries_d, 1);
#undef ALLOC
        jobs_h.assign(jobs, JobDev{});
        done_h.assign(jobs, 0);
        slot_used.assign(jobs, 0);
        free_ranges.clear();
        free_ranges[0] = cap;
        PT_HIP(ctx, launch_zero(counters, 12, st));
        PT_HIP(ctx, hipMemsetAsync(queries_d, 0, sizeof(unsigned long long), st));
        PT_HIP(ctx, hipMemsetAsync(done_d, 0, sizeof(int) * (size_t)jobs, st));
        PT_HIP(ctx, hipStreamSynchronize(st));
        set = 1;
        return PT_OK;
    }

    // First-fit allocation of a contiguous range of path slots.
    bool alloc_range(int64_t n, int64_t* base)
    {
        for (auto it = free_ranges.begin(); it != free_ranges.end(); ++it) {
            if (it->second >= n) {
                *base = it->first;
                const int64_t rest = it->second - n;
                const int64_t start = it->first + n;
                free_ranges.erase(it);
                if (rest > 0) free_ranges[start] = rest;
                return true;
            }
        }
        return false;
    }

    void free_range(int64_t base, int64_t n)
    {
        auto it = free_ranges.emplace(base, n).first;
        auto next = std::next(it);
        if (next != free_ranges.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_ranges.erase(next);
        }
        if (it != free_ranges.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_ranges.erase(it);
            }
        }
    }

    int64_t largest_free() const
    {
        int64_t m = 0;
        for (const auto& r : free_ranges) m = std::max(m, r.second);
        return m;
    }

    static SceneView scene_view(pt_scene* s, pt_bvh* b)
    {
        SceneView sc;
        sc.bvh = b->view;
        sc.tri_mat = b->tri_mat_d;
        sc.tri_mat_shading = b->tri_mat_shading_d;
        sc.tri_smooth = b->tri_smooth_d;
        sc.tri_nrm = b->tri_nrm_d;
        sc.mats = s->mats_d;
        sc.lights = s->lights_d;
        sc.num_lights = s->num_lights;
        sc.cam = s->cam;
        sc.env = s->env;
        return sc;
    }

    // Starts a job: path slots, job table entry, set-up kernel appending its paths and
    // camera rays to the queues the next iteration reads.
    int start(EngineJob& j)
    {
        int slot = -1;
        for (int i = 0; i < max_jobs; ++i)
            if (!slot_used[i]) {
                slot = i;
                break;
            }
        if (slot < 0) return pt_fail(ctx, PT_ERR_BUDGET, "job table full");
        if (!alloc_range(j.npix, &j.base)) return pt_fail(ctx, PT_ERR_BUDGET, "no free path slots");
        slot_used[slot] = 1;
        j.slot = slot;
        JobDev jd;
        jd.sc = scene_view(j.scene, j.bvh);
        JobView& v = jd.job;
        v.width = j.desc.width;
        v.height = j.desc.height;
        v.spp = j.desc.spp;
        v.max_bounces = j.desc.max_bounces;
        v.integrator = j.desc.integrator;
        v.seed = j.desc.seed;
        v.out = j.out_d;
        v.done = done_d + slot;
        v.path_base = (int)j.base;
        v.job_index = slot;
        jobs_h[slot] = jd;
        done_h[slot] = 0;
        PT_HIP(ctx, hipMemsetAsync(done_d + slot, 0, sizeof(int), st));
        PT_HIP(ctx, hipMemcpyAsync(jobs_d + slot, &jd, sizeof(JobDev), hipMemcpyHostToDevice, st));
        PT_HIP(ctx, hipMemcpyAsync(views_d + slot, &jd.sc.bvh, sizeof(BvhView), hipMemcpyHostToDevice, st));
        if (!multi) {
            sc1 = jd.sc;
            job1 = v;
            integrator1 = v.integrator;
            drop_graphs();
        }
        PT_HIP(ctx, launch_paths_init(jd.sc, v, ps, queue(set, 0), act[set], counters + 4 * set + 3, st));
        j.started = std::chrono::steady_clock::now();
        return PT_OK;
    }

    void retire(EngineJob& j)
    {
        free_range(j.base, j.npix);
        slot_used[j.slot] = 0;
    }

    // One wavefront iteration reading counter set r and writing set 1 - r.
    int enqueue_iteration(int r)
    {
        const int w = 1 - r;
        TraceJob tj;
        tj.num_seg = 3;
        float* outs_t[3] = {hit_t, nullptr, sh_t};
        int* outs_k[3] = {hit_k, sh_k, sh_k};
        for (int k = 0; k < 3; ++k) {
            TraceSegment& g = tj.seg[k];
            g.rays = q_rays[k];
            g.path = q_path[k];
            g.count = counters + 4 * r + k;
            g.any = k == 1;
            g.map_prim = 0;
            g.hit_flag = 0;
            g.out_t = outs_t[k];
            g.out_k = outs_k[k];
        }
        tj.views = multi ? views_d : nullptr;
        tj.job_of_path = multi ? ps.job : nullptr;
        tj.queries = queries_d;
        PT_HIP(ctx, launch_zero(counters + 8, 1, st));
        PT_HIP(ctx, launch_trace_job(sc1.bvh, tj, counters + 8, nullptr, st));
        PT_HIP(ctx, launch_zero(counters + 4 * w, 4, st));
        ShadeArgs a;
        a.ps = ps;
        a.active_in = act[r];
        a.active_count = counters + 4 * r + 3;
        a.active_out = act[w];
        a.active_out_count = counters + 4 * w + 3;
        a.ext = queue(w, 0);
        a.sh_any = queue(w, 1);
        a.sh_closest = queue(w, 2);
        a.hit_t = hit_t;
        a.hit_k = hit_k;
        a.sh_t = sh_t;
        a.sh_k = sh_k;
        PT_HIP(ctx, launch_shade(integrator1, sc1, job1, multi ? jobs_d : nullptr, a, st));
        return PT_OK;
    }

    int build_graphs()
    {
        for (int r = 0; r < 2; ++r) {
            if (graphs[r]) continue;
            hipGraph_t g = nullptr;
            PT_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue_iteration(r);
            hipError_t e = hipStreamEndCapture(st, &g);
            if (rc != PT_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            PT_HIP(ctx, e);
            e = hipGraphInstantiate(&graphs[r], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            PT_HIP(ctx, e);
        }
        return PT_OK;
    }

    // k iterations as graph replays, then a host synchronisation that refreshes the
    // per-job finished-pixel counts.
    int run(int k)
    {
        PT_HIP(ctx, hipSetDevice(ctx->device));
        int rc = build_graphs();
        if (rc != PT_OK) return rc;
        for (int i = 0; i < k; ++i) {
            PT_HIP(ctx, hipGraphLaunch(graphs[set], st));
            set = 1 - set;
            ++iterations;
        }
        PT_HIP(ctx, hipMemcpyAsync(done_h.data(), done_d, sizeof(int) * (size_t)max_jobs, hipMemcpyDeviceToHost, st));
        PT_HIP(ctx, hipStreamSynchronize(st));
        return PT_OK;
    }

    // One iteration with every launch timed by HIP events (kernel kinds: extension
    // trace, environment-shadow trace, area-shadow trace, shading).
    int run_timed(int64_t items[4], int64_t launches[4], double ms[4], bool* any_live)
    {
        PT_HIP(ctx, hipSetDevice(ctx->device));
        int h[4];
        PT_HIP(ctx, hipMemcpyAsync(h, counters + 4 * set, sizeof h, hipMemcpyDeviceToHost, st));
        PT_HIP(ctx, hipStreamSynchronize(st));
        *any_live = h[3] > 0;
        if (!*any_live) return PT_OK;
        const int r = set, w = 1 - r;
        hipEvent_t ev[8];
        for (auto& e : ev) PT_HIP(ctx, hipEventCreate(&e));
        float* outs_t[3] = {hit_t, nullptr, sh_t};
        int* outs_k[3] = {hit_k, sh_k, sh_k};
        for (int k = 0; k < 3; ++k) {
            TraceJob tj;
            tj.num_seg = 1;
            TraceSegment& g = tj.seg[0];
            g.rays = q_rays[k];
            g.path = q_path[k];
            g.count = counters + 4 * r + k;
            g.any = k == 1;
            g.map_prim = 0;
            g.hit_flag = 0;
            g.out_t = outs_t[k];
            g.out_k = outs_k[k];
            tj.views = multi ? views_d : nullptr;
            tj.job_of_path = multi ? ps.job : nullptr;
            tj.queries = queries_d;
            PT_HIP(ctx, launch_zero(counters + 8, 1, st));
            PT_HIP(ctx, hipEventRecord(ev[2 * k], st));
            PT_HIP(ctx, launch_trace_job(sc1.bvh, tj, counters + 8, nullptr, st));
            PT_HIP(ctx, hipEventRecord(ev[2 * k + 1], st));
        }
        PT_HIP(ctx, launch_zero(counters + 4 * w, 4, st));
        ShadeArgs a;
        a.ps = ps;
        a.active_in = act[r];
        a.active_count = counters + 4 * r + 3;
        a.active_out = act[w];
        a.active_out_count = counters + 4 * w + 3;
        a.ext = queue(w, 0);
        a.sh_any = queue(w, 1);
        a.sh_closest = queue(w, 2);
        a.hit_t = hit_t;
        a.hit_k = hit_k;
        a.sh_t = sh_t;
        a.sh_k = sh_k;
        PT_HIP(ctx, hipEventRecord(ev[6], st));
        PT_HIP(ctx, launch_shade(integrator1, sc1, job1, multi ? jobs_d : nullptr, a, st));
        PT_HIP(ctx, hipEventRecord(ev[7], st));
        PT_HIP(ctx, hipMemcpyAsync(done_h.data(), done_d, sizeof(int) * (size_t)max_jobs, hipMemcpyDeviceToHost, st));
        PT_HIP(ctx, hipStreamSynchronize(st));
        for (int k = 0; k < 4; ++k) {
            if (h[k] <= 0) continue;
            float t = 0.0f;
            PT_HIP(ctx, hipEventElapsedTime(&t, ev[2 * k], ev[2 * k + 1]));
            items[k] += h[k];
            launches[k] += 1;
            ms[k] += t;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        set = w;
        ++iterations;
        return PT_OK;
    }

    int64_t queries()
    {
        unsigned long long q = 0;
        if (hipMemcpy(&q, queries_d, sizeof q, hipMemcpyDeviceToHost) != hipSuccess) return -1;
        return (int64_t)q;
    }
};

namespace {

int check_job(pt_context* ctx, const pt_job_desc& j)
{
    if (!j.scene || !j.bvh || !j.rgb_out) return pt_fail(ctx, PT_ERR_ARG, "job needs a scene, a BVH and an output");
    if (j.bvh->scene != j.scene) return pt_fail(ctx, PT_ERR_ARG, "BVH belongs to another scene");
    const pt_render_desc& d = j.render;
    if (d.width <= 0 || d.height <= 0 || d.spp <= 0 || d.max_bounces < 0 ||
        (d.integrator != PT_INTEGRATOR_MIS && d.integrator != PT_INTEGRATOR_SIMPLE))
        return pt_fail(ctx, PT_ERR_ARG, "bad render settings");
    if ((int64_t)d.width * d.height > (1 << 28)) return pt_fail(ctx, PT_ERR_ARG, "image too large");
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
    j.bytes = j.npix * (kPathBytes + 12);
    j.submitted = std::chrono::steady_clock::now();
    return j;
}

// Runs a set of jobs to completion on an engine, admitting them in order as path
// slots free up.
int run_jobs(pt_engine& E, std::vector<EngineJob>& jobs, int iters_per_sync, pt_batch_stats* stats)
{
    pt_context* ctx = E.ctx;
    size_t next = 0;
    std::vector<EngineJob*> running;
    hipEvent_t ev0, ev1;
    PT_HIP(ctx, hipEventCreate(&ev0));
    PT_HIP(ctx, hipEventCreate(&ev1));
    PT_HIP(ctx, hipEventRecord(ev0, E.st));
    const auto t0 = std::chrono::steady_clock::now();
    int64_t replays = 0;
    while (next < jobs.size() || !running.empty()) {
        while (next < jobs.size() && (int64_t)running.size() < E.max_jobs && E.largest_free() >= jobs[next].npix) {
            EngineJob& j = jobs[next];
            int rc = prepare_output(ctx, j);
            if (rc == PT_OK) rc = E.start(j);
            if (rc != PT_OK) return rc;
            running.push_back(&j);
            ++next;
        }
        if (running.empty()) return pt_fail(ctx, PT_ERR_BUDGET, "job larger than the engine's path capacity");
        int rc = E.run(iters_per_sync);
        if (rc != PT_OK) return rc;
        ++replays;
        for (size_t i = 0; i < running.size();) {
            EngineJob& j = *running[i];
            if (E.done_h[j.slot] >= j.npix) {
                rc = finish_output(ctx, j, E.st);
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

} // namespace

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
    // the context keeps a single-job engine, grown on demand
    if (!ctx->engine || ctx->engine->capacity < npix) {
        delete ctx->engine;
        ctx->engine = nullptr;
        rc = pt_engine_create(ctx, npix, 1, 0, &ctx->engine);
        if (rc != PT_OK) return rc;
    }
    pt_engine& E = *ctx->engine;
    E.iterations = 0;
    PT_HIP(ctx, hipMemsetAsync(E.queries_d, 0, sizeof(unsigned long long), E.st));
    std::vector<EngineJob> jobs(1, make_job(jd, 0));
    if (d->flags & PT_RENDER_TIME_KERNELS) {
        EngineJob& j = jobs[0];
        rc = prepare_output(ctx, j);
        if (rc == PT_OK) rc = E.start(j);
        if (rc != PT_OK) return rc;
        int64_t items[4] = {0, 0, 0, 0}, launches[4] = {0, 0, 0, 0};
        double ms[4] = {0, 0, 0, 0};
        const auto t0 = std::chrono::steady_clock::now();
        bool live = true;
        while (live) {
            rc = E.run_timed(items, launches, ms, &live);
            if (rc != PT_OK) return rc;
        }
        const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        rc = finish_output(ctx, j, E.st);
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
            }
        }
        return PT_OK;
    }
    pt_batch_stats bs;
    rc = run_jobs(E, jobs, 16, &bs);
    if (rc != PT_OK) return rc;
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->queries = bs.queries;
        stats->iterations = bs.iterations;
        stats->elapsed_ms = bs.elapsed_ms;
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
        if (js.back().npix > E->capacity) return pt_fail(E->ctx, PT_ERR_BUDGET, "job larger than the engine");
    }
    E->iterations = 0;
    PT_HIP(E->ctx, hipMemsetAsync(E->queries_d, 0, sizeof(unsigned long long), E->st));
    return run_jobs(*E, js, iterations_per_sync > 0 ? iterations_per_sync : 8, stats);
}

} // extern "C"
