// pt_server.cpp — the render server: resident scenes and BVHs, a request queue and a
// host thread that drives one multi-job engine with continuous admission.
//
// Requests are admitted between graph replays whenever the job table, a contiguous
// range of path slots and the device-memory budget allow (first fit over the head of
// the queue; a request that has waited more than kHeadWaitMs blocks later ones, so
// large requests are not starved). Each running request owns a device image that the
// engine writes; when its last pixel finishes the image is copied to host memory and
// the request is done. Scenes are charged to the budget when added; idle scenes (no
// queued or running request) are evicted least-recently-used first when a new scene
// or a request image does not fit.
//
// Optionally, requests of at least 2^solo_log2_samples samples render one at a time on a
// second, single-job engine S (the single-job kernels, with the scene as kernel arguments)
// while the others share the multi-job engine E; each round queues both engines' graph
// replays before waiting for either, so their launches overlap on the GPU. A request's
// image does not depend on the engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "pt_engine.h"

using namespace pt;

namespace {

using Clock = std::chrono::steady_clock;

constexpr double kHeadWaitMs = 1000.0;
constexpr int kLookahead = 32;

thread_local std::string t_err;

int sfail(int code, const std::string& msg)
{
    t_err = msg;
    return code;
}

double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct SScene {
    pt_scene* scene = nullptr;
    pt_bvh* bvh = nullptr;
    int64_t bytes = 0;
    int users = 0; // queued and running requests
    bool removing = false;
    uint64_t last_use = 0;
};

struct SRequest {
    int64_t id = 0;
    int64_t scene_id = 0;
    pt_render_desc desc{};
    int state = PT_REQ_QUEUED;
    int64_t npix = 0;
    int64_t bytes = 0; // device image, plus the preview buffer once one was asked for
    EngineJob job;
    bool on_solo = false; // runs on the single-job engine S
    float* img_d = nullptr;
    float* prev_d = nullptr;
    std::vector<float> image;   // final image (host)
    std::vector<float> preview; // last progressive estimate (host)
    float progress = 0.0f;
    bool preview_wanted = false;
    uint64_t preview_gen = 0;
    std::string err;
    Clock::time_point submitted, started, finished;
};

} // namespace

struct pt_server {
    pt_context* up = nullptr;  // scene uploads and BVH builds, under up_mu
    pt_context* run = nullptr; // the request loop's context
    pt_engine* E = nullptr;
    pt_engine* S = nullptr;     // single-job engine of the solo requests, or null
    SRequest* solo = nullptr;   // the request running on S
    int64_t solo_min = 0;       // samples from which a request runs on S (0: never)
    int iters = 8;
    std::mutex mu; // guards everything below
    std::mutex up_mu;
    std::condition_variable cv_work, cv_done;
    std::map<int64_t, SScene> scenes;
    std::map<int64_t, std::unique_ptr<SRequest>> reqs;
    std::deque<SRequest*> queue;
    std::vector<SRequest*> running;
    int64_t budget = 0, used = 0;
    int64_t pool = 0; // path pool and job tables, charged at creation
    int64_t next_id = 1;
    uint64_t use_clock = 0;
    bool stop = false;
    pt_server_stats st{};
    unsigned long long* samples_d = nullptr;
    std::thread worker;

    // Frees idle scenes, least recently used first, until `need` more bytes fit; frees
    // nothing when even all idle scenes would not make room.
    bool make_room(int64_t need)
    {
        if (used + need <= budget) return true;
        int64_t idle = 0;
        for (const auto& kv : scenes)
            if (kv.second.users == 0) idle += kv.second.bytes;
        if (used - idle + need > budget) return false;
        while (used + need > budget) {
            auto victim = scenes.end();
            for (auto it = scenes.begin(); it != scenes.end(); ++it)
                if (it->second.users == 0 && (victim == scenes.end() || it->second.last_use < victim->second.last_use))
                    victim = it;
            if (victim == scenes.end()) return false;
            drop_scene(victim);
            st.evicted_scenes += 1;
        }
        return true;
    }

    void drop_scene(std::map<int64_t, SScene>::iterator it)
    {
        if (it->second.bvh) pt_bvh_destroy(it->second.bvh);
        if (it->second.scene) pt_scene_destroy(it->second.scene);
        used -= it->second.bytes;
        scenes.erase(it);
    }

    void release_scene_user(int64_t scene_id)
    {
        auto it = scenes.find(scene_id);
        if (it == scenes.end()) return;
        it->second.users -= 1;
        if (it->second.users == 0 && it->second.removing) drop_scene(it);
    }

    void free_request_memory(SRequest* r)
    {
        if (r->img_d) (void)hipFree(r->img_d);
        if (r->prev_d) (void)hipFree(r->prev_d);
        r->img_d = r->prev_d = nullptr;
        used -= r->bytes;
        r->bytes = 0;
    }

    void fail_request(SRequest* r, const std::string& why)
    {
        r->state = PT_REQ_FAILED;
        r->err = why;
        r->finished = Clock::now();
        free_request_memory(r);
        release_scene_user(r->scene_id);
    }

    // Starts queued requests that fit (under mu).
    void admit()
    {
        int scanned = 0;
        const bool head_only = !queue.empty() && ms_between(queue.front()->submitted, Clock::now()) > kHeadWaitMs;
        for (auto it = queue.begin(); it != queue.end() && scanned < kLookahead; ++scanned) {
            SRequest* r = *it;
            // the image and the request's sample buffer (the engine allocates the latter)
            const int64_t items = r->npix * r->desc.spp;
            const int64_t bytes = 12 * r->npix + kSampleBytes * items;
            EngineJob probe;
            probe.npix = r->npix;
            probe.items = items;
            const bool to_solo = S && items >= solo_min;
            const bool fits = to_solo ? (solo == nullptr && make_room(bytes))
                                      : ((int64_t)running.size() < E->max_jobs && E->largest_free() >= E->job_slots(probe) &&
                                         E->width_ok(pt_engine::job_width_of(scenes.at(r->scene_id).bvh)) &&
                                         make_room(bytes));
            if (!fits) {
                if (running.empty() && !solo && it == queue.begin() && !make_room(bytes)) {
                    // nothing will free memory for it: no request runs and no scene is idle
                    fail_request(r, "request does not fit the memory budget next to the resident scenes");
                    it = queue.erase(it);
                    cv_done.notify_all();
                    continue;
                }
                if (head_only) break;
                ++it;
                continue;
            }
            hipError_t e = hipMalloc((void**)&r->img_d, 12 * (size_t)r->npix);
            if (e != hipSuccess) {
                r->img_d = nullptr;
                fail_request(r, std::string("image allocation: ") + hipGetErrorString(e));
                it = queue.erase(it);
                cv_done.notify_all();
                continue;
            }
            used += bytes;
            r->bytes = bytes;
            SScene& sc = scenes.at(r->scene_id);
            sc.last_use = ++use_clock;
            pt_job_desc jd;
            jd.scene = sc.scene;
            jd.bvh = sc.bvh;
            jd.render = r->desc;
            jd.rgb_out = r->img_d;
            jd.out_on_device = 1;
            r->job = make_job(jd, r->id);
            pt_engine* eng = to_solo ? S : E;
            if (to_solo) r->job.slots = S->job_slots(r->job);
            int rc = prepare_output(run, r->job);
            if (rc == PT_OK) rc = eng->start(r->job);
            if (rc != PT_OK) {
                fail_request(r, run->err);
                it = queue.erase(it);
                cv_done.notify_all();
                continue;
            }
            r->state = PT_REQ_RUNNING;
            r->started = Clock::now();
            r->on_solo = to_solo;
            if (to_solo) {
                solo = r;
                S->live_h = (int)r->job.slots;
            } else {
                running.push_back(r);
            }
            it = queue.erase(it);
            if (head_only) break;
        }
        st.running = (int64_t)running.size() + (solo ? 1 : 0);
        st.queued = (int64_t)queue.size();
    }

    // Progressive estimate of a running request into its host preview (under mu).
    int make_preview(SRequest* r)
    {
        pt_engine* E = r->on_solo ? this->S : this->E;
        if (!r->prev_d) {
            PT_HIP(run, hipMalloc((void**)&r->prev_d, 12 * (size_t)r->npix));
            used += 12 * r->npix;
            r->bytes += 12 * r->npix;
        }
        PT_HIP(run, hipMemsetAsync(samples_d, 0, sizeof(unsigned long long), E->st));
        PT_HIP(run, launch_preview(E->jobs_h[r->job.slot].job, E->ps, r->prev_d, samples_d, E->st));
        r->preview.resize(3 * (size_t)r->npix);
        unsigned long long samples = 0;
        PT_HIP(run, hipMemcpyAsync(r->preview.data(), r->prev_d, 12 * (size_t)r->npix, hipMemcpyDeviceToHost, E->st));
        PT_HIP(run, hipMemcpyAsync(&samples, samples_d, sizeof samples, hipMemcpyDeviceToHost, E->st));
        PT_HIP(run, hipStreamSynchronize(E->st));
        r->progress = (float)((double)samples / ((double)r->npix * r->desc.spp));
        return PT_OK;
    }

    // A finished request's image to host memory; the request leaves its engine (under mu).
    void complete(SRequest* r, pt_engine* eng)
    {
        r->image.resize(3 * (size_t)r->npix);
        hipError_t e = eng->resolve(r->job) == PT_OK ? hipSuccess : hipErrorUnknown;
        if (e == hipSuccess)
            e = hipMemcpyAsync(r->image.data(), r->img_d, 12 * (size_t)r->npix, hipMemcpyDeviceToHost, eng->st);
        if (e == hipSuccess) e = hipStreamSynchronize(eng->st);
        eng->retire(r->job);
        if (e != hipSuccess) {
            fail_request(r, std::string("image copy: ") + hipGetErrorString(e));
        } else {
            r->state = PT_REQ_DONE;
            r->progress = 1.0f;
            r->finished = Clock::now();
            free_request_memory(r);
            release_scene_user(r->scene_id);
            st.completed += 1;
        }
    }

    void serve_preview(SRequest* r, bool* notify)
    {
        if (!r->preview_wanted) return;
        if (make_preview(r) != PT_OK) r->progress = -1.0f;
        r->preview_wanted = false;
        r->preview_gen += 1;
        *notify = true;
    }

    void loop()
    {
        (void)hipSetDevice(run->device);
        std::unique_lock<std::mutex> lk(mu);
        while (!stop) {
            admit();
            if (running.empty() && !solo) {
                cv_work.wait(lk, [&] { return stop || !queue.empty(); });
                continue;
            }
            const auto t0 = Clock::now();
            // only this thread changes `running` and `solo`
            const bool e_run = !running.empty();
            SRequest* const sr = solo;
            // the solo request's last live paths complete in one launch
            const bool s_last = sr && S->job_exhausted(sr->job) && S->live_h > 0 && S->live_h <= S->finish_threshold;
            lk.unlock();
            int rc = PT_OK;
            if (sr) rc = s_last ? S->launch_finish() : S->launch(iters);
            if (rc == PT_OK && e_run) rc = E->launch(iters);
            if (rc == PT_OK && sr) rc = S->wait();
            if (rc == PT_OK && e_run) rc = E->wait();
            const int64_t q = rc == PT_OK ? E->queries() + (S ? S->queries() : 0) : -1;
            lk.lock();
            st.busy_ms += ms_between(t0, Clock::now());
            st.syncs += 1;
            st.iterations = E->iterations + (S ? S->iterations : 0);
            if (q >= 0) st.queries = q;
            if (rc != PT_OK || q < 0) {
                const std::string why = "request loop: " + run->err;
                for (SRequest* r : running) {
                    E->retire(r->job);
                    fail_request(r, why);
                }
                running.clear();
                if (solo) {
                    S->retire(solo->job);
                    fail_request(solo, why);
                    solo = nullptr;
                }
                cv_done.notify_all();
                continue;
            }
            bool notify = false;
            if (sr) {
                if (S->job_done(sr->job)) {
                    complete(sr, S);
                    solo = nullptr;
                    notify = true;
                } else if (s_last || S->live_h == 0) {
                    S->retire(sr->job);
                    fail_request(sr, "request loop: path completion left samples unfinished");
                    solo = nullptr;
                    notify = true;
                } else {
                    serve_preview(sr, &notify);
                }
            }
            for (size_t i = 0; i < running.size();) {
                SRequest* r = running[i];
                if (E->job_done(r->job)) {
                    complete(r, E);
                    running.erase(running.begin() + i);
                    notify = true;
                    continue;
                }
                serve_preview(r, &notify);
                ++i;
            }
            st.running = (int64_t)running.size() + (solo ? 1 : 0);
            if (notify) cv_done.notify_all();
        }
    }
};

namespace {

SRequest* find_request(pt_server* s, int64_t id)
{
    auto it = s->reqs.find(id);
    return it == s->reqs.end() ? nullptr : it->second.get();
}

void fill_info(const SRequest* r, pt_request_info* info)
{
    info->state = r->state;
    info->progress = r->progress;
    info->queue_ms = r->state == PT_REQ_QUEUED ? ms_between(r->submitted, Clock::now())
                                               : ms_between(r->submitted, r->started);
    info->latency_ms = r->state == PT_REQ_DONE ? ms_between(r->submitted, r->finished) : 0.0;
}

} // namespace

extern "C" {

const char* pt_server_last_error(void) { return t_err.c_str(); }

int pt_server_create(const pt_server_desc* d, pt_server** out)
{
    if (!d || !out) return sfail(PT_ERR_ARG, "null argument");
    *out = nullptr;
    const int64_t max_paths = d->max_paths > 0 ? d->max_paths : ((int64_t)4 << 20);
    const int max_jobs = d->max_jobs > 0 ? d->max_jobs : 64;
    if (d->hbm_budget_bytes <= 0) return sfail(PT_ERR_ARG, "the server needs a device-memory budget");
    if (d->solo_log2_samples != 0 && (d->solo_log2_samples < 10 || d->solo_log2_samples > 40))
        return sfail(PT_ERR_ARG, "solo_log2_samples must be 0 or 10..40");
    const int64_t solo_paths = d->solo_log2_samples ? (d->solo_paths > 0 ? d->solo_paths : ((int64_t)8 << 20)) : 0;
    const int64_t pool = (max_paths + solo_paths) * kPathBytes + (int64_t)max_jobs * 512;
    if (pool > d->hbm_budget_bytes) return sfail(PT_ERR_BUDGET, "the path pool alone exceeds the budget");
    auto s = std::make_unique<pt_server>();
    s->budget = d->hbm_budget_bytes;
    s->iters = d->iterations_per_sync > 0 ? d->iterations_per_sync : 8;
    if (pt_ctx_create(d->device, &s->up) != PT_OK || pt_ctx_create(d->device, &s->run) != PT_OK) {
        if (s->up) pt_ctx_destroy(s->up);
        if (s->run) pt_ctx_destroy(s->run);
        return sfail(PT_ERR_HIP, "context creation failed");
    }
    int rc = pt_engine_create(s->run, max_paths, max_jobs, 1, &s->E);
    if (rc == PT_OK) {
        s->E->clear_samples = true;     // progressive previews read the sample buffers
        s->E->keep_spare_samples = false; // their memory is charged to the budget per request
    }
    if (rc == PT_OK && solo_paths > 0) {
        rc = pt_engine_create(s->run, solo_paths, 1, 0, &s->S);
        if (rc == PT_OK) {
            s->S->clear_samples = true;
            s->S->keep_spare_samples = false;
            s->S->min_job_slots = solo_paths;
            s->solo_min = (int64_t)1 << d->solo_log2_samples;
        }
    }
    if (rc == PT_OK && hipMalloc((void**)&s->samples_d, sizeof(unsigned long long)) != hipSuccess) rc = PT_ERR_OOM;
    if (rc != PT_OK) {
        const std::string why = s->run->err;
        if (s->S) pt_engine_destroy(s->S);
        if (s->E) pt_engine_destroy(s->E);
        pt_ctx_destroy(s->up);
        pt_ctx_destroy(s->run);
        return sfail(rc, "engine: " + why);
    }
    s->used = pool;
    s->pool = pool;
    s->st.budget_bytes = s->budget;
    pt_server* raw = s.release();
    raw->worker = std::thread([raw] { raw->loop(); });
    *out = raw;
    return PT_OK;
}

int pt_server_destroy(pt_server* s)
{
    if (!s) return sfail(PT_ERR_ARG, "null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->run->device);
    (void)hipStreamSynchronize(s->E->st);
    if (s->S) (void)hipStreamSynchronize(s->S->st);
    for (SRequest* r : s->running) s->E->retire(r->job);
    if (s->solo) s->S->retire(s->solo->job);
    for (auto& kv : s->reqs) s->free_request_memory(kv.second.get());
    s->reqs.clear();
    while (!s->scenes.empty()) s->drop_scene(s->scenes.begin());
    if (s->samples_d) (void)hipFree(s->samples_d);
    if (s->S) pt_engine_destroy(s->S);
    pt_engine_destroy(s->E);
    pt_ctx_destroy(s->up);
    pt_ctx_destroy(s->run);
    delete s;
    return PT_OK;
}

int pt_server_add_scene(pt_server* s, const pt_scene_desc* d, const pt_bvh_desc* bd, int64_t* scene_id)
{
    if (!s || !d || !scene_id) return sfail(PT_ERR_ARG, "null argument");
    if (d->num_triangles < 0) return sfail(PT_ERR_ARG, "bad triangle count");
    // upper bound of the scene's and the BVH's device bytes, reserved while they are built
    const int64_t estimate = ((int64_t)d->num_triangles + std::max(d->num_spheres, 0)) * 320 +
                             (int64_t)d->num_materials * 64 + (int64_t)d->num_lights * 48 +
                             (d->env_rgb ? (int64_t)d->env_width * d->env_height * 12 + 8 * (int64_t)d->env_height : 0) +
                             (d->uvs ? ((int64_t)d->num_triangles + std::max(d->num_spheres, 0)) * 48 : 0) +
                             pt_scene_map_bytes(d) + (1 << 20);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return sfail(PT_ERR_STATE, "server stopped");
        if (!s->make_room(estimate)) return sfail(PT_ERR_BUDGET, "scene does not fit the memory budget");
        s->used += estimate;
    }
    pt_scene* sc = nullptr;
    pt_bvh* bvh = nullptr;
    int rc;
    std::string why;
    {
        std::lock_guard<std::mutex> lk(s->up_mu);
        rc = pt_scene_create(s->up, d, &sc);
        if (rc == PT_OK) rc = pt_bvh_build(sc, bd, &bvh);
        if (rc != PT_OK) why = s->up->err;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->used -= estimate;
    if (rc != PT_OK) {
        if (sc) pt_scene_destroy(sc);
        return sfail(rc, why);
    }
    SScene e;
    e.scene = sc;
    e.bvh = bvh;
    e.bytes = sc->device_bytes + bvh->device_bytes;
    e.last_use = ++s->use_clock;
    s->used += e.bytes;
    const int64_t id = s->next_id++;
    s->scenes[id] = e;
    *scene_id = id;
    s->st.resident_scenes = (int64_t)s->scenes.size();
    return PT_OK;
}

int pt_server_remove_scene(pt_server* s, int64_t scene_id)
{
    if (!s) return sfail(PT_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->scenes.find(scene_id);
    if (it == s->scenes.end() || it->second.removing) return sfail(PT_ERR_STATE, "unknown or evicted scene");
    if (it->second.users > 0) it->second.removing = true;
    else s->drop_scene(it);
    s->st.resident_scenes = (int64_t)s->scenes.size();
    return PT_OK;
}

int pt_server_scene_resident(pt_server* s, int64_t scene_id)
{
    if (!s) return 0;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->scenes.find(scene_id);
    return it != s->scenes.end() && !it->second.removing ? 1 : 0;
}

int pt_server_submit(pt_server* s, int64_t scene_id, const pt_render_desc* d, int64_t* request_id)
{
    if (!s || !d || !request_id) return sfail(PT_ERR_ARG, "null argument");
    if (d->width <= 0 || d->height <= 0 || d->spp <= 0 || d->max_bounces < 0 ||
        (d->integrator != PT_INTEGRATOR_MIS && d->integrator != PT_INTEGRATOR_SIMPLE) ||
        d->sampler_length < 0 || d->sampler_length > PT_MAX_SAMPLER_LENGTH)
        return sfail(PT_ERR_ARG, "bad render settings");
    const int64_t npix = (int64_t)d->width * d->height;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return sfail(PT_ERR_STATE, "server stopped");
    auto it = s->scenes.find(scene_id);
    if (it == s->scenes.end() || it->second.removing) return sfail(PT_ERR_STATE, "unknown or evicted scene");
    if (s->pool + it->second.bytes + 12 * npix + kSampleBytes * npix * d->spp > s->budget)
        return sfail(PT_ERR_BUDGET, "request image and samples do not fit the budget next to its scene");
    auto r = std::make_unique<SRequest>();
    r->id = s->next_id++;
    r->scene_id = scene_id;
    r->desc = *d;
    r->desc.flags = 0;
    r->npix = npix;
    r->submitted = Clock::now();
    it->second.users += 1;
    it->second.last_use = ++s->use_clock;
    *request_id = r->id;
    s->queue.push_back(r.get());
    s->reqs[r->id] = std::move(r);
    s->st.submitted += 1;
    s->st.queued = (int64_t)s->queue.size();
    s->cv_work.notify_one();
    return PT_OK;
}

int pt_server_poll(pt_server* s, int64_t id, pt_request_info* info)
{
    if (!s || !info) return sfail(PT_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    SRequest* r = find_request(s, id);
    if (!r) return sfail(PT_ERR_ARG, "unknown request");
    fill_info(r, info);
    if (r->state == PT_REQ_FAILED) t_err = r->err;
    return PT_OK;
}

int pt_server_wait(pt_server* s, int64_t id, int32_t timeout_ms, pt_request_info* info)
{
    if (!s) return sfail(PT_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    SRequest* r = find_request(s, id);
    if (!r) return sfail(PT_ERR_ARG, "unknown request");
    auto finished = [&] { return s->stop || r->state == PT_REQ_DONE || r->state == PT_REQ_FAILED; };
    if (timeout_ms < 0) s->cv_done.wait(lk, finished);
    else s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), finished);
    if (info) fill_info(r, info);
    if (r->state == PT_REQ_FAILED) t_err = r->err;
    return PT_OK;
}

int pt_server_preview(pt_server* s, int64_t id, float* rgb_out, float* progress)
{
    if (!s || !rgb_out) return sfail(PT_ERR_ARG, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    SRequest* r = find_request(s, id);
    if (!r) return sfail(PT_ERR_ARG, "unknown request");
    if (r->state == PT_REQ_RUNNING) {
        const uint64_t gen = r->preview_gen;
        r->preview_wanted = true;
        s->cv_done.wait(lk, [&] { return s->stop || r->preview_gen != gen || r->state != PT_REQ_RUNNING; });
    }
    const size_t n = 3 * (size_t)r->npix;
    switch (r->state) {
    case PT_REQ_QUEUED:
        std::fill(rgb_out, rgb_out + n, 0.0f);
        if (progress) *progress = 0.0f;
        return PT_OK;
    case PT_REQ_RUNNING:
        if (r->progress < 0.0f || r->preview.size() != n) return sfail(PT_ERR_HIP, "preview failed");
        std::memcpy(rgb_out, r->preview.data(), n * sizeof(float));
        if (progress) *progress = r->progress;
        return PT_OK;
    case PT_REQ_DONE:
        std::memcpy(rgb_out, r->image.data(), n * sizeof(float));
        if (progress) *progress = 1.0f;
        return PT_OK;
    default:
        return sfail(PT_ERR_STATE, "request failed: " + r->err);
    }
}

int pt_server_fetch(pt_server* s, int64_t id, float* rgb_out)
{
    if (!s || !rgb_out) return sfail(PT_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    SRequest* r = find_request(s, id);
    if (!r) return sfail(PT_ERR_ARG, "unknown request");
    if (r->state == PT_REQ_FAILED) {
        const std::string why = r->err;
        s->reqs.erase(id);
        return sfail(PT_ERR_STATE, "request failed: " + why);
    }
    if (r->state != PT_REQ_DONE) return sfail(PT_ERR_STATE, "request not done");
    std::memcpy(rgb_out, r->image.data(), r->image.size() * sizeof(float));
    s->reqs.erase(id);
    return PT_OK;
}

int pt_server_get_stats(pt_server* s, pt_server_stats* out)
{
    if (!s || !out) return sfail(PT_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->resident_scenes = (int64_t)s->scenes.size();
    out->bytes_in_use = s->used;
    out->budget_bytes = s->budget;
    out->queued = (int64_t)s->queue.size();
    out->running = (int64_t)s->running.size() + (s->solo ? 1 : 0);
    return PT_OK;
}

} // extern "C"
