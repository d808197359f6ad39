// pt_engine.h — the wavefront render engine shared by pt_render, pt_render_batch and
// the render server: a pool of path slots, the ray queues, the job table and the
// graph-captured iteration.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <map>
#include <vector>

#include "pt_internal.h"

namespace pt {

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
    int64_t base = 0, npix = 0; // path slots; the job renders npix pixels from pixel_base
    int64_t pixel_base = 0;
    int slot = -1;             // job table index
    int64_t bytes = 0;         // working memory charged to the budget
    std::chrono::steady_clock::time_point submitted, started;
};

} // namespace pt

struct pt_engine {
    pt_context* ctx = nullptr;
    hipStream_t st = nullptr;
    int64_t capacity = 0;
    int max_jobs = 0;
    bool multi = true;
    pt::Buffers buf;
    pt::PathView ps{};
    float *hit_t = nullptr, *sh_t = nullptr;
    int *hit_k = nullptr, *sh_k = nullptr;
    int* act[2] = {nullptr, nullptr};
    // [4*set + kind] kind 0 ext, 1 any, 2 closest, 3 live; the trace cursors from
    // PT_CURSOR_STRIDE on
    int* counters = nullptr;
    float4* q_rays[3] = {nullptr, nullptr, nullptr};
    int32_t* q_path[3] = {nullptr, nullptr, nullptr};
    pt::JobDev* jobs_d = nullptr;
    pt::BvhView* views_d = nullptr;
    int* done_d = nullptr;
    unsigned long long* queries_d = nullptr;
    std::vector<pt::JobDev> jobs_h;
    std::vector<int> done_h;
    std::vector<char> slot_used;
    std::map<int64_t, int64_t> free_ranges; // first path slot -> length
    int set = 1;                            // counter set the next iteration reads
    // single-job mode: the job's views are kernel arguments (captured in the graphs)
    pt::SceneView sc1{};
    pt::JobView job1{};
    int integrator1 = 0;
    hipGraphExec_t graphs[2] = {nullptr, nullptr};
    int64_t iterations = 0;
    int width = 0;             // BVH width of the jobs a multi-job engine runs (0: none yet)
    pt::LaunchTuning tune;     // launch shapes (captured in the graphs)
    int iters_per_sync = 8;    // default graph replays between host synchronisations
    // Once every job is admitted and at most this many paths are live at a
    // synchronisation, the rest runs as one path-completion launch (pt::launch_finish_paths).
    int finish_threshold = 1 << 17;
    int live_h = 0;            // live paths at the last synchronisation
    // pinned host copies of the finished-pixel counts and the live count, written by
    // asynchronous copies (so that queueing them does not wait for the stream)
    int* stage_h = nullptr;

    pt::Queue queue(int s, int kind) const
    {
        pt::Queue q;
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
        if (stage_h) (void)hipHostFree(stage_h);
    }

    int init(pt_context* c, int64_t cap, int jobs, bool multi_job)
    {
        ctx = c;
        capacity = cap;
        max_jobs = jobs;
        multi = multi_job;
        PT_HIP(ctx, hipSetDevice(ctx->device));
        PT_HIP(ctx, hipStreamCreate(&st));
        (void)pt::cu_count();
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
        ALLOC(counters, PT_CURSOR_STRIDE + PT_CURSOR_INTS);
        for (int k = 0; k < 3; ++k) {
            ALLOC(q_rays[k], 2 * P);
            ALLOC(q_path[k], P);
        }
        ALLOC(jobs_d, (size_t)jobs);
        ALLOC(views_d, (size_t)jobs);
        ALLOC(done_d, (size_t)jobs);
        ALLOC(queries_d, 1);
#undef ALLOC
        PT_HIP(ctx, hipHostMalloc((void**)&stage_h, sizeof(int) * ((size_t)jobs + 1)));
        jobs_h.assign(jobs, pt::JobDev{});
        done_h.assign(jobs, 0);
        slot_used.assign(jobs, 0);
        free_ranges.clear();
        free_ranges[0] = cap;
        PT_HIP(ctx, pt::launch_zero(counters, 8, st));
        PT_HIP(ctx, pt::launch_zero(counters + PT_CURSOR_STRIDE, PT_TRACE_PARTS, st, PT_CURSOR_STRIDE));
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

    static pt::SceneView scene_view(pt_scene* s, pt_bvh* b)
    {
        pt::SceneView sc;
        sc.bvh = b->view;
        sc.tri_mat = b->tri_mat_d;
        sc.tri_mat_shading = b->tri_mat_shading_d;
        sc.tri_smooth = b->tri_smooth_d;
        sc.tri_nrm = b->tri_nrm_d;
        sc.tri_uv = b->tri_uv_d;
        sc.textures = s->textures_d;
        sc.mat_maps = s->mat_maps_d;
        sc.mats = s->mats_d;
        sc.lights = s->lights_d;
        sc.num_lights = s->num_lights;
        sc.cam = s->cam;
        sc.env = s->env;
        return sc;
    }

    // Starts a job: path slots, job table entry, set-up kernel appending its paths and
    // camera rays to the queues the next iteration reads.
    static int job_width(const pt::EngineJob& j) { return j.bvh->view.nodes8 ? 8 : 2; }

    // A multi-job engine runs BVHs of one width at a time (one trace kernel per launch).
    bool width_ok(int w) const
    {
        if (!multi) return true;
        for (int i = 0; i < max_jobs; ++i)
            if (slot_used[i]) return w == width;
        return true;
    }

    int start(pt::EngineJob& j)
    {
        const int jw = job_width(j);
        if (!width_ok(jw)) return pt_fail(ctx, PT_ERR_UNSUPPORTED, "jobs of another BVH width are running");
        if (multi && jw != width) {
            width = jw;
            drop_graphs(); // the trace kernel is captured in the graphs
        }
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
        pt::JobDev jd;
        jd.sc = scene_view(j.scene, j.bvh);
        pt::JobView& v = jd.job;
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
        v.pixel_base = (int)j.pixel_base;
        v.num_pixels = (int)j.npix;
        jobs_h[slot] = jd;
        done_h[slot] = 0;
        PT_HIP(ctx, hipMemsetAsync(done_d + slot, 0, sizeof(int), st));
        PT_HIP(ctx, hipMemcpyAsync(jobs_d + slot, &jd, sizeof(pt::JobDev), hipMemcpyHostToDevice, st));
        PT_HIP(ctx, hipMemcpyAsync(views_d + slot, &jd.sc.bvh, sizeof(pt::BvhView), hipMemcpyHostToDevice, st));
        if (!multi) {
            sc1 = jd.sc;
            job1 = v;
            integrator1 = v.integrator;
            drop_graphs();
        }
        PT_HIP(ctx, pt::launch_paths_init(jd.sc, v, ps, queue(set, 0), act[set], counters + 4 * set + 3, st));
        j.started = std::chrono::steady_clock::now();
        return PT_OK;
    }

    void retire(pt::EngineJob& j)
    {
        free_range(j.base, j.npix);
        slot_used[j.slot] = 0;
    }

    // One wavefront iteration reading counter set r and writing set 1 - r.
    int enqueue_iteration(int r)
    {
        const int w = 1 - r;
        pt::TraceJob tj;
        tj.num_seg = 3;
        float* outs_t[3] = {hit_t, nullptr, sh_t};
        int* outs_k[3] = {hit_k, sh_k, sh_k};
        for (int k = 0; k < 3; ++k) {
            pt::TraceSegment& g = tj.seg[k];
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
        tj.wide = multi && width == 8;
        tj.job_of_path = multi ? ps.job : nullptr;
        tj.queries = queries_d;
        PT_HIP(ctx, pt::launch_zero(counters + PT_CURSOR_STRIDE, PT_TRACE_PARTS, st, PT_CURSOR_STRIDE));
        PT_HIP(ctx, pt::launch_trace_job(sc1.bvh, tj, counters + PT_CURSOR_STRIDE, nullptr, st, tune));
        PT_HIP(ctx, pt::launch_zero(counters + 4 * w, 4, st));
        pt::ShadeArgs a;
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
        PT_HIP(ctx, pt::launch_shade(integrator1, sc1, job1, multi ? jobs_d : nullptr, a, st, tune));
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
        const int rc = launch(k);
        return rc != PT_OK ? rc : wait();
    }

    // k graph replays and the copies of the counts, queued without waiting.
    int launch(int k)
    {
        PT_HIP(ctx, hipSetDevice(ctx->device));
        int rc = build_graphs();
        if (rc != PT_OK) return rc;
        for (int i = 0; i < k; ++i) {
            PT_HIP(ctx, hipGraphLaunch(graphs[set], st));
            set = 1 - set;
            ++iterations;
        }
        return queue_counts();
    }

    int queue_counts()
    {
        PT_HIP(ctx, hipMemcpyAsync(stage_h, done_d, sizeof(int) * (size_t)max_jobs, hipMemcpyDeviceToHost, st));
        PT_HIP(ctx, hipMemcpyAsync(stage_h + max_jobs, counters + 4 * set + 3, sizeof(int), hipMemcpyDeviceToHost,
                                   st));
        return PT_OK;
    }

    // Waits for the queued work and takes the counts.
    int wait()
    {
        PT_HIP(ctx, hipStreamSynchronize(st));
        for (int i = 0; i < max_jobs; ++i) done_h[i] = stage_h[i];
        live_h = stage_h[max_jobs];
        return PT_OK;
    }

    // Traces the queued rays and then completes every live path in one launch, and
    // synchronises as run() does.
    int finish_live()
    {
        const int rc = launch_finish();
        return rc != PT_OK ? rc : wait();
    }

    int launch_finish()
    {
        PT_HIP(ctx, hipSetDevice(ctx->device));
        const int r = set;
        pt::TraceJob tj;
        tj.num_seg = 3;
        float* outs_t[3] = {hit_t, nullptr, sh_t};
        int* outs_k[3] = {hit_k, sh_k, sh_k};
        for (int k = 0; k < 3; ++k) {
            pt::TraceSegment& g = tj.seg[k];
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
        tj.wide = multi && width == 8;
        tj.job_of_path = multi ? ps.job : nullptr;
        tj.queries = queries_d;
        PT_HIP(ctx, pt::launch_zero(counters + PT_CURSOR_STRIDE, PT_TRACE_PARTS, st, PT_CURSOR_STRIDE));
        PT_HIP(ctx, pt::launch_trace_job(sc1.bvh, tj, counters + PT_CURSOR_STRIDE, nullptr, st, tune));
        pt::ShadeArgs a{};
        a.ps = ps;
        a.active_in = act[r];
        a.active_count = counters + 4 * r + 3;
        a.hit_t = hit_t;
        a.hit_k = hit_k;
        a.sh_t = sh_t;
        a.sh_k = sh_k;
        const pt::HitBuffers h{hit_t, hit_k, sh_t, sh_k};
        const bool wide = multi ? width == 8 : sc1.bvh.nodes8 != nullptr;
        PT_HIP(ctx, pt::launch_finish_paths(integrator1, sc1, job1, multi ? jobs_d : nullptr, wide, a, h, queries_d,
                                            st));
        // no live path and no queued ray is left: the counters of this set start the next job
        PT_HIP(ctx, pt::launch_zero(counters + 4 * r, 4, st));
        ++iterations;
        return queue_counts();
    }

    // One iteration with every launch timed by HIP events (kernel kinds: extension
    // trace, environment-shadow trace, area-shadow trace, shading).
    // stats_d (optional): 2 device counters per trace kind (node visits, triangle tests)
    int run_timed(int64_t items[4], int64_t launches[4], double ms[4], bool* any_live,
                  unsigned long long* stats_d = nullptr)
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
            pt::TraceJob tj;
            tj.num_seg = 1;
            pt::TraceSegment& g = tj.seg[0];
            g.rays = q_rays[k];
            g.path = q_path[k];
            g.count = counters + 4 * r + k;
            g.any = k == 1;
            g.map_prim = 0;
            g.hit_flag = 0;
            g.out_t = outs_t[k];
            g.out_k = outs_k[k];
            tj.views = multi ? views_d : nullptr;
            tj.wide = multi && width == 8;
            tj.job_of_path = multi ? ps.job : nullptr;
            tj.queries = queries_d;
            PT_HIP(ctx, pt::launch_zero(counters + PT_CURSOR_STRIDE, PT_TRACE_PARTS, st, PT_CURSOR_STRIDE));
            PT_HIP(ctx, hipEventRecord(ev[2 * k], st));
            PT_HIP(ctx, pt::launch_trace_job(sc1.bvh, tj, counters + PT_CURSOR_STRIDE, stats_d ? stats_d + 2 * k : nullptr, st, tune));
            PT_HIP(ctx, hipEventRecord(ev[2 * k + 1], st));
        }
        PT_HIP(ctx, pt::launch_zero(counters + 4 * w, 4, st));
        pt::ShadeArgs a;
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
        PT_HIP(ctx, pt::launch_shade(integrator1, sc1, job1, multi ? jobs_d : nullptr, a, st, tune));
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

    // One iteration as plain launches: the host reads the queue counts, launches a trace
    // for each non-empty queue and the shading, and waits (no graph, one synchronisation
    // per iteration; the baseline of the batched and served forms).
    int run_plain(bool* any_live)
    {
        PT_HIP(ctx, hipSetDevice(ctx->device));
        int h[4];
        PT_HIP(ctx, hipMemcpyAsync(h, counters + 4 * set, sizeof h, hipMemcpyDeviceToHost, st));
        PT_HIP(ctx, hipStreamSynchronize(st));
        *any_live = h[3] > 0;
        if (!*any_live) return PT_OK;
        const int r = set, w = 1 - r;
        float* outs_t[3] = {hit_t, nullptr, sh_t};
        int* outs_k[3] = {hit_k, sh_k, sh_k};
        for (int k = 0; k < 3; ++k) {
            if (h[k] <= 0) continue;
            pt::TraceJob tj;
            tj.num_seg = 1;
            pt::TraceSegment& g = tj.seg[0];
            g.rays = q_rays[k];
            g.path = q_path[k];
            g.count = counters + 4 * r + k;
            g.any = k == 1;
            g.map_prim = 0;
            g.hit_flag = 0;
            g.out_t = outs_t[k];
            g.out_k = outs_k[k];
            tj.views = multi ? views_d : nullptr;
            tj.wide = multi && width == 8;
            tj.job_of_path = multi ? ps.job : nullptr;
            tj.queries = queries_d;
            PT_HIP(ctx, pt::launch_zero(counters + PT_CURSOR_STRIDE, PT_TRACE_PARTS, st, PT_CURSOR_STRIDE));
            PT_HIP(ctx, pt::launch_trace_job(sc1.bvh, tj, counters + PT_CURSOR_STRIDE, nullptr, st, tune));
        }
        PT_HIP(ctx, pt::launch_zero(counters + 4 * w, 4, st));
        pt::ShadeArgs a;
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
        PT_HIP(ctx, pt::launch_shade(integrator1, sc1, job1, multi ? jobs_d : nullptr, a, st, tune));
        PT_HIP(ctx, hipMemcpyAsync(done_h.data(), done_d, sizeof(int) * (size_t)max_jobs, hipMemcpyDeviceToHost, st));
        PT_HIP(ctx, hipStreamSynchronize(st));
        set = w;
        ++iterations;
        return PT_OK;
    }

    int64_t queries()
    {
        unsigned long long q = 0;
        if (hipMemcpyAsync(&q, queries_d, sizeof q, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
        if (hipStreamSynchronize(st) != hipSuccess) return -1;
        return (int64_t)q;
    }
};

namespace pt {

int check_job(pt_context* ctx, const pt_job_desc& j);
int prepare_output(pt_context* ctx, EngineJob& j);
int finish_output(pt_context* ctx, EngineJob& j, hipStream_t st);
EngineJob make_job(const pt_job_desc& d, int64_t id);
// Runs a set of jobs to completion on an engine, admitting them in order as path slots
// free up.
int run_jobs(pt_engine& E, std::vector<EngineJob>& jobs, int iters_per_sync, pt_batch_stats* stats);

} // namespace pt
