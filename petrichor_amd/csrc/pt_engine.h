// pt_engine.h — the wavefront render engine shared by pt_render, pt_render_batch and
// the render server: a pool of path slots, the ray queues, the job table and the
// graph-captured iteration.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <vector>

#include "pt_internal.h"

namespace pt {

// Bytes of device memory one path slot uses (path state, queues, hits, live lists).
constexpr int64_t kPathBytes = 3 * 16 + 5 * 16 + 16 + 8 + 6 * 16 + 4 * 4 + 2 * 4 + 3 * (32 + 4);
// Bytes per sample of a job's sample buffer.
constexpr int64_t kSampleBytes = 12;
// Path slots a job gets at least (when it has that many samples): its samples, not only
// its pixels, run in parallel.
constexpr int64_t kMinJobSlots = (int64_t)1 << 20;

// The micro-jittered 2D sampler's table for a sequence length (1 + length entries).
std::vector<float2> micro_jitter_table(int length);

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
    int64_t items = 0;         // samples: npix * spp
    int64_t slots = 0;         // path slots it runs on (set at admission)
    float* samples_d = nullptr; // its sample buffer (kSampleBytes per sample), from the engine's pool
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
    unsigned long long* done_d = nullptr;   // samples finished, per job slot
    unsigned long long* cursor_d = nullptr; // next item handed out, per job slot
    unsigned long long* queries_d = nullptr;
    std::vector<pt::JobDev> jobs_h;
    std::vector<unsigned long long> done_h, cursor_h;
    // sample buffers of retired jobs, kept for the next jobs (bytes -> buffer)
    std::multimap<size_t, void*> spare_samples;
    std::vector<void*> sample_bufs;
    bool clear_samples = false;     // fill new sample buffers with the preview's "not done" pattern
    bool keep_spare_samples = true; // keep retired jobs' sample buffers for later jobs (else free them)
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
    bool ext = false;          // a multi-job engine's launches take the extended walk (a job's BVH is deep or has spheres)
    pt::LaunchTuning tune;     // launch shapes (captured in the graphs)
    int iters_per_sync = 8;    // default graph replays between host synchronisations
    // Once every job is admitted and at most this many paths are live at a
    // synchronisation, the rest runs as one path-completion launch (pt::launch_finish_paths).
    int finish_threshold = 1 << 17;
    int live_h = 0;            // live paths at the last synchronisation
    // pinned host copies of the per-job sample counts and cursors and of the live count,
    // written by asynchronous copies (so that queueing them does not wait for the stream)
    unsigned long long* stage_h = nullptr;
    int* stage_live_h = nullptr;
    // Launch log (PT_RENDER_LAUNCH_LOG): every trace launch of the graphs records its
    // device-clock span and ray counts (log_counts: also its node visits and triangle
    // tests); the completion launches are timed with events and their rays counted from
    // the query counter (finish_q: before and after each, in stream order).
    bool logging = false, log_counts = false;
    pt::LaunchRec* log_d = nullptr;
    int* log_n_d = nullptr;
    unsigned long long* finish_q_d = nullptr; // 2 per completion launch
    std::vector<hipEvent_t> finish_ev;        // 2 per completion launch
    int finish_n = 0;

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
        for (hipEvent_t e : finish_ev) (void)hipEventDestroy(e);
        if (st) (void)hipStreamSynchronize(st);
        for (void* b : sample_bufs) (void)hipFree(b);
        for (auto& t : halton_d) (void)hipFree(t.second);
        if (st) (void)hipStreamDestroy(st);
        if (stage_h) (void)hipHostFree(stage_h);
        if (stage_live_h) (void)hipHostFree(stage_live_h);
    }

    // Path slots of a job: all its samples, up to max(its pixels, min_job_slots), within the
    // engine's capacity; in a multi-job engine at most half of it, so that jobs arriving
    // later find slots while a large one runs.
    int64_t min_job_slots = pt::kMinJobSlots;
    int64_t job_slots(const pt::EngineJob& j) const
    {
        const int64_t want = std::max(j.npix, min_job_slots);
        const int64_t cap = multi ? std::max<int64_t>(1, capacity / 2) : capacity;
        return std::max<int64_t>(1, std::min(std::min(j.items, want), cap));
    }

    static constexpr int kFinishLogCap = 64;

    // Turns the launch log on (clearing it) or off; the graphs are captured again.
    // The trace launch shapes of t (blocks per CU, chunk, lanes per ray, phased steps); the
    // graphs are captured again when they change.
    void use_trace_tuning(const pt::LaunchTuning& t)
    {
        if (tune.trace_blocks_per_cu == t.trace_blocks_per_cu && tune.trace_chunk == t.trace_chunk &&
            tune.threads_per_ray == t.threads_per_ray && tune.leaf_batch == t.leaf_batch)
            return;
        tune.trace_blocks_per_cu = t.trace_blocks_per_cu;
        tune.trace_chunk = t.trace_chunk;
        tune.threads_per_ray = t.threads_per_ray;
        tune.leaf_batch = t.leaf_batch;
        drop_graphs();
    }

    int set_logging(bool on, bool counts)
    {
        PT_HIP(ctx, hipSetDevice(ctx->device));
        if (on && !log_d) {
            PT_HIP(ctx, buf.alloc(&log_d, PT_LAUNCH_LOG_CAP));
            PT_HIP(ctx, buf.alloc(&log_n_d, 1));
            PT_HIP(ctx, buf.alloc(&finish_q_d, 2 * kFinishLogCap));
            finish_ev.resize(2 * kFinishLogCap);
            for (hipEvent_t& e : finish_ev) PT_HIP(ctx, hipEventCreate(&e));
        }
        if (on) {
            std::vector<pt::LaunchRec> init(PT_LAUNCH_LOG_CAP);
            for (pt::LaunchRec& r : init) {
                std::memset(&r, 0, sizeof r);
                r.t0 = ~0ull;
            }
            PT_HIP(ctx, hipMemcpyAsync(log_d, init.data(), init.size() * sizeof(pt::LaunchRec),
                                       hipMemcpyHostToDevice, st));
            PT_HIP(ctx, pt::launch_set(log_n_d, 0, st));
            PT_HIP(ctx, hipStreamSynchronize(st));
            finish_n = 0;
        }
        if (on != logging || counts != log_counts) drop_graphs();
        logging = on;
        log_counts = on && counts;
        return PT_OK;
    }

    void log_job(pt::TraceJob& tj) const
    {
        tj.launch_log = logging ? log_d : nullptr;
        tj.launch_index = logging ? log_n_d : nullptr;
        tj.log_counts = log_counts ? 1 : 0;
    }

    // Sums of the log: trace launches, their rays, device-clock milliseconds, node visits,
    // triangle tests; completion launches, their rays and event milliseconds.
    int read_log(pt_render_stats* s)
    {
        if (!logging) return PT_OK;
        PT_HIP(ctx, hipStreamSynchronize(st));
        int n = 0;
        PT_HIP(ctx, hipMemcpy(&n, log_n_d, sizeof n, hipMemcpyDeviceToHost));
        std::vector<pt::LaunchRec> recs((size_t)n);
        if (n > 0) PT_HIP(ctx, hipMemcpy(recs.data(), log_d, (size_t)n * sizeof(pt::LaunchRec), hipMemcpyDeviceToHost));
        int khz = 0;
        PT_HIP(ctx, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
        for (const pt::LaunchRec& r : recs) {
            const int64_t items = (int64_t)r.items[0] + r.items[1] + r.items[2];
            if (items == 0 || r.t1 < r.t0) continue; // an empty launch returns before logging
            s->log_launches += 1;
            s->log_items += items;
            s->log_trace_ms += khz > 0 ? (double)(r.t1 - r.t0) / (double)khz : 0.0;
            s->log_nodes += (int64_t)r.nodes;
            s->log_tris += (int64_t)r.tris;
        }
        std::vector<unsigned long long> q((size_t)2 * finish_n);
        if (finish_n > 0)
            PT_HIP(ctx, hipMemcpy(q.data(), finish_q_d, q.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        for (int i = 0; i < finish_n; ++i) {
            float ms = 0.0f;
            PT_HIP(ctx, hipEventElapsedTime(&ms, finish_ev[2 * i], finish_ev[2 * i + 1]));
            s->finish_launches += 1;
            s->finish_rays += (int64_t)(q[2 * i + 1] - q[2 * i]);
            s->finish_ms += ms;
        }
        return PT_OK;
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
        ALLOC(cursor_d, (size_t)jobs);
        ALLOC(queries_d, 1);
#undef ALLOC
        PT_HIP(ctx, hipHostMalloc((void**)&stage_h, sizeof(unsigned long long) * 2 * (size_t)jobs));
        PT_HIP(ctx, hipHostMalloc((void**)&stage_live_h, sizeof(int)));
        jobs_h.assign(jobs, pt::JobDev{});
        done_h.assign(jobs, 0);
        cursor_h.assign(jobs, 0);
        slot_used.assign(jobs, 0);
        free_ranges.clear();
        free_ranges[0] = cap;
        PT_HIP(ctx, pt::launch_zero(counters, 8, st));
        PT_HIP(ctx, pt::launch_zero(counters + PT_CURSOR_STRIDE, PT_TRACE_PARTS, st, PT_CURSOR_STRIDE));
        PT_HIP(ctx, hipMemsetAsync(queries_d, 0, sizeof(unsigned long long), st));
        PT_HIP(ctx, hipMemsetAsync(done_d, 0, sizeof(unsigned long long) * (size_t)jobs, st));
        PT_HIP(ctx, hipMemsetAsync(cursor_d, 0, sizeof(unsigned long long) * (size_t)jobs, st));
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
        sc.mat_nmaps = s->mat_nmaps_d;
        sc.map_strength = s->map_strength_d;
        sc.mats = s->mats_d;
        sc.lights = s->lights_d;
        sc.light_kinds = s->light_kinds_d;
        sc.num_lights = s->num_lights;
        sc.cam = s->cam;
        sc.env = s->env;
        return sc;
    }

    // Starts a job: path slots, job table entry, set-up kernel appending its paths and
    // camera rays to the queues the next iteration reads.
    static int job_width_of(const pt_bvh* b) { return b->view.nodes8 ? b->view.width : 2; }
    static int job_width(const pt::EngineJob& j) { return job_width_of(j.bvh); }

    // A multi-job engine runs BVHs of one width at a time (one trace kernel per launch).
    bool width_ok(int w) const
    {
        if (!multi) return true;
        for (int i = 0; i < max_jobs; ++i)
            if (slot_used[i]) return w == width;
        return true;
    }

    // A sample buffer of at least `bytes` from the spare ones, or a new one.
    int take_samples(size_t bytes, float** out)
    {
        auto it = spare_samples.lower_bound(bytes);
        if (it != spare_samples.end() && it->first <= 2 * bytes) {
            *out = (float*)it->second;
            spare_samples.erase(it);
            return PT_OK;
        }
        void* b = nullptr;
        hipError_t e = hipMalloc(&b, bytes);
        if (e != hipSuccess && !spare_samples.empty()) {
            // give back the spare buffers and try once more
            PT_HIP(ctx, hipStreamSynchronize(st));
            for (auto& kv : spare_samples) {
                (void)hipFree(kv.second);
                sample_bufs.erase(std::find(sample_bufs.begin(), sample_bufs.end(), kv.second));
            }
            spare_samples.clear();
            e = hipMalloc(&b, bytes);
        }
        PT_HIP(ctx, e);
        sample_bufs.push_back(b);
        *out = (float*)b;
        return PT_OK;
    }
    static size_t samples_bytes(const pt::EngineJob& j) { return (size_t)pt::kSampleBytes * (size_t)std::max<int64_t>(j.items, 1); }

    // The micro-jittered 2D sampler's table for a sequence length (pt::Samplers), made once
    // per engine and length: entry 0 is (length as bits, mu), then the Halton pairs.
    std::map<int, float2*> halton_d;
    int halton_table(int length, const float2** out)
    {
        auto it = halton_d.find(length);
        if (it == halton_d.end()) {
            const std::vector<float2> t = pt::micro_jitter_table(length);
            float2* d = nullptr;
            PT_HIP(ctx, hipMalloc((void**)&d, t.size() * sizeof(float2)));
            const hipError_t e = hipMemcpyAsync(d, t.data(), t.size() * sizeof(float2), hipMemcpyHostToDevice, st);
            if (e == hipSuccess) (void)hipStreamSynchronize(st);
            if (e != hipSuccess) {
                (void)hipFree(d);
                return pt_fail(ctx, PT_ERR_HIP, std::string("sampler table: ") + hipGetErrorString(e));
            }
            it = halton_d.emplace(length, d).first;
        }
        *out = it->second;
        return PT_OK;
    }

    int start(pt::EngineJob& j)
    {
        const int jw = job_width(j);
        if (!width_ok(jw)) return pt_fail(ctx, PT_ERR_UNSUPPORTED, "jobs of another BVH width are running");
        if (multi && jw != width) {
            width = jw;
            drop_graphs(); // the trace kernel is captured in the graphs
        }
        if (multi) {
            // the extended walk from the first job that needs it until the engine runs empty
            const bool je = j.bvh->view.parents != nullptr || j.bvh->view.kinds != nullptr;
            bool running = false;
            for (int i = 0; i < max_jobs; ++i) running |= slot_used[i] != 0;
            const bool want = running ? (ext || je) : je;
            if (want != ext) {
                ext = want;
                drop_graphs();
            }
        }
        int slot = -1;
        for (int i = 0; i < max_jobs; ++i)
            if (!slot_used[i]) {
                slot = i;
                break;
            }
        if (slot < 0) return pt_fail(ctx, PT_ERR_BUDGET, "job table full");
        const float2* mj = nullptr;
        if (j.desc.sampler_length > 0) {
            if (const int rc = halton_table(j.desc.sampler_length, &mj)) return rc;
        }
        if (j.slots <= 0) j.slots = job_slots(j);
        if (!alloc_range(j.slots, &j.base)) return pt_fail(ctx, PT_ERR_BUDGET, "no free path slots");
        int rc = take_samples(samples_bytes(j), &j.samples_d);
        if (rc != PT_OK) {
            free_range(j.base, j.slots);
            return rc;
        }
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
        v.cursor = cursor_d + slot;
        v.samples = j.samples_d;
        v.items = (unsigned long long)j.items;
        v.path_base = (int)j.base;
        v.num_slots = (int)j.slots;
        v.job_index = slot;
        v.pixel_base = (int)j.pixel_base;
        v.num_pixels = (int)j.npix;
        v.mj = mj;
        jobs_h[slot] = jd;
        done_h[slot] = 0;
        // k_paths_init hands out the first items, one per slot
        cursor_h[slot] = (unsigned long long)std::min(j.slots, j.items);
        PT_HIP(ctx, hipMemcpyAsync(jobs_d + slot, &jobs_h[slot], sizeof(pt::JobDev), hipMemcpyHostToDevice, st));
        PT_HIP(ctx, hipMemcpyAsync(views_d + slot, &jobs_h[slot].sc.bvh, sizeof(pt::BvhView), hipMemcpyHostToDevice, st));
        if (!multi) {
            sc1 = jd.sc;
            job1 = v;
            integrator1 = v.integrator;
            drop_graphs();
        }
        if (clear_samples) PT_HIP(ctx, pt::launch_samples_clear(v, st));
        // counters[8..9]: the reservation of the job's first entries (free words of the
        // counters' first 256-byte line)
        PT_HIP(ctx, pt::launch_paths_init(jd.sc, v, ps, queue(set, 0), act[set], counters + 4 * set + 3, counters + 8,
                                          st));
        j.started = std::chrono::steady_clock::now();
        return PT_OK;
    }

    bool job_done(const pt::EngineJob& j) const { return done_h[j.slot] >= (unsigned long long)j.items; }
    bool job_exhausted(const pt::EngineJob& j) const { return cursor_h[j.slot] >= (unsigned long long)j.items; }

    // The image of a job whose samples are all done (in stream order after them).
    int resolve(const pt::EngineJob& j)
    {
        PT_HIP(ctx, pt::launch_resolve(jobs_h[j.slot].job, st));
        return PT_OK;
    }

    void retire(pt::EngineJob& j)
    {
        free_range(j.base, j.slots);
        slot_used[j.slot] = 0;
        if (j.samples_d) {
            if (keep_spare_samples) {
                // reused by a later job only after the work queued so far (stream order)
                spare_samples.emplace(samples_bytes(j), j.samples_d);
            } else {
                (void)hipStreamSynchronize(st);
                (void)hipFree(j.samples_d);
                sample_bufs.erase(std::find(sample_bufs.begin(), sample_bufs.end(), (void*)j.samples_d));
            }
            j.samples_d = nullptr;
        }
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
        tj.width = multi ? width : 2;
        tj.ext = multi && ext;
        tj.job_of_path = multi ? ps.job : nullptr;
        tj.queries = queries_d;
        log_job(tj);
        // set w's counts are read by nothing before shading appends to them
        PT_HIP(ctx, pt::launch_zero_iteration(counters + PT_CURSOR_STRIDE, PT_TRACE_PARTS, PT_CURSOR_STRIDE,
                                              counters + 4 * w, st));
        PT_HIP(ctx, pt::launch_trace_job(sc1.bvh, tj, counters + PT_CURSOR_STRIDE, nullptr, st, tune));
        if (logging) PT_HIP(ctx, pt::launch_log_tick(counters + 4 * r, log_d, log_n_d, st));
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
        const size_t nb = sizeof(unsigned long long) * (size_t)max_jobs;
        PT_HIP(ctx, hipMemcpyAsync(stage_h, done_d, nb, hipMemcpyDeviceToHost, st));
        PT_HIP(ctx, hipMemcpyAsync(stage_h + max_jobs, cursor_d, nb, hipMemcpyDeviceToHost, st));
        PT_HIP(ctx, hipMemcpyAsync(stage_live_h, counters + 4 * set + 3, sizeof(int), hipMemcpyDeviceToHost, st));
        return PT_OK;
    }

    // Waits for the queued work and takes the counts.
    int wait()
    {
        PT_HIP(ctx, hipStreamSynchronize(st));
        for (int i = 0; i < max_jobs; ++i) {
            done_h[i] = stage_h[i];
            cursor_h[i] = stage_h[max_jobs + i];
        }
        live_h = *stage_live_h;
        return PT_OK;
    }

    // Counts of the jobs, synchronously (the timed and plain loops).
    int read_counts()
    {
        const int rc = queue_counts();
        return rc != PT_OK ? rc : wait();
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
        tj.width = multi ? width : 2;
        tj.ext = multi && ext;
        tj.job_of_path = multi ? ps.job : nullptr;
        tj.queries = queries_d;
        log_job(tj);
        PT_HIP(ctx, pt::launch_zero(counters + PT_CURSOR_STRIDE, PT_TRACE_PARTS, st, PT_CURSOR_STRIDE));
        PT_HIP(ctx, pt::launch_trace_job(sc1.bvh, tj, counters + PT_CURSOR_STRIDE, nullptr, st, tune));
        if (logging) PT_HIP(ctx, pt::launch_log_tick(counters + 4 * r, log_d, log_n_d, st));
        pt::ShadeArgs a{};
        a.ps = ps;
        a.active_in = act[r];
        a.active_count = counters + 4 * r + 3;
        a.hit_t = hit_t;
        a.hit_k = hit_k;
        a.sh_t = sh_t;
        a.sh_k = sh_k;
        const pt::HitBuffers h{hit_t, hit_k, sh_t, sh_k};
        const int walk_width = multi ? width : (sc1.bvh.nodes8 != nullptr ? sc1.bvh.width : 2);
        const bool log_finish = logging && finish_n < kFinishLogCap;
        if (log_finish) {
            PT_HIP(ctx, pt::launch_copy_u64(queries_d, finish_q_d + 2 * finish_n, st));
            PT_HIP(ctx, hipEventRecord(finish_ev[2 * finish_n], st));
        }
        PT_HIP(ctx, pt::launch_finish_paths(integrator1, sc1, job1, multi ? jobs_d : nullptr, walk_width, a, h, queries_d,
                                            st));
        if (log_finish) {
            PT_HIP(ctx, hipEventRecord(finish_ev[2 * finish_n + 1], st));
            PT_HIP(ctx, pt::launch_copy_u64(queries_d, finish_q_d + 2 * finish_n + 1, st));
            ++finish_n;
        }
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
            tj.width = multi ? width : 2;
        tj.ext = multi && ext;
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
            tj.width = multi ? width : 2;
        tj.ext = multi && ext;
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
