// serve.cpp — resident scenes, batched jobs and the serving loop of the C-ABI.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <thread>

#include "api_internal.h"

using namespace dfr;

#define CTX_GUARD(ctx)                                  \
    if (!(ctx)) return DFR_E_NULL;                      \
    std::lock_guard<std::mutex> _lock((ctx)->mutex);     \
    hipSetDevice((ctx)->device)

namespace dfr {

static int validate_jobs(Ctx* c, const dfr_job* jobs, int n) {
    for (int j = 0; j < n; j++) {
        const dfr_job& J = jobs[j];
        if (J.scene < 0 || J.scene >= (int)c->scenes.size() || !c->scenes[J.scene])
            return c->fail(DFR_E_ARGUMENT, "job scene is not resident");
        if (J.width < 1 || J.height < 1 || J.spp < 1) return c->fail(DFR_E_ARGUMENT, "job size must be positive");
    }
    return DFR_OK;
}

static DevJob to_dev(const dfr_job& J, int64_t film_offset, int64_t work_begin) {
    DevJob d{};
    d.seed = J.seed;
    d.width = J.width; d.height = J.height;
    d.spp = J.spp; d.sample_base = 0;
    d.max_bounce = J.maximum_bounce; d.min_rr = J.minimum_bounce_rr;
    d.algorithm = DFR_ALGO_PATH_TRACING; d.ao_samples = 0; d.max_distance = 0.0f;
    d.film_offset = film_offset;
    d.work_begin = work_begin;
    d.work_count = (int64_t)J.width * J.height * J.spp;
    std::memcpy(d.cam, J.camera, sizeof(d.cam));
    d.cam[DFR_CAMERA_RES_X] = (float)J.width;
    d.cam[DFR_CAMERA_RES_Y] = (float)J.height;
    d.scene = J.scene;
    return d;
}

// Renders jobs with the film laid out job after job; `films` receive the
// per-job means.
int batch_render_impl(Ctx* c, const dfr_job* jobs, int n, int mode, float* const* films) {
    int rc = validate_jobs(c, jobs, n);
    if (rc) return rc;
    std::vector<DevJob> dj(n);
    int64_t pixels = 0, work = 0;
    for (int j = 0; j < n; j++) {
        dj[j] = to_dev(jobs[j], pixels, work);
        pixels += (int64_t)jobs[j].width * jobs[j].height;
        work += dj[j].work_count;
    }
    DeviceBuffer film;
    if (film.alloc(std::max<int64_t>(1, pixels) * sizeof(float4)) != hipSuccess) return c->fail(DFR_E_MEMORY, "batch film");
    hipMemsetAsync(film.ptr, 0, pixels * sizeof(float4), c->stream);
    dfr_stats st{};
    hipEventRecord(c->ev0, c->stream);
    if (mode == DFR_BATCH_PLAIN) {
        // the plain per-job kernel: one launch per job per pass, jobs in order
        if (c->counters.alloc(128) != hipSuccess) return c->fail(DFR_E_MEMORY, "counters");
        hipMemsetAsync(c->counters.ptr, 0, 64, c->stream);
        int64_t max_threads = 0;
        for (int j = 0; j < n; j++) max_threads = std::max<int64_t>(max_threads, ((int64_t)jobs[j].width * jobs[j].height + 255) / 256 * 256);
        if ((rc = c->ensure_overflow(max_threads))) return rc;
        if (c->job_buf.alloc(std::max<size_t>(1, (size_t)n * jobs[0].spp) * sizeof(DevJob)) != hipSuccess)
            return c->fail(DFR_E_MEMORY, "job table");
        std::vector<DevJob> passes;
        for (int j = 0; j < n; j++)
            for (int s = 0; s < jobs[j].spp; s++) {
                DevJob p = dj[j];
                p.spp = 1; p.sample_base = s;
                passes.push_back(p);
            }
        if (c->job_buf.alloc(passes.size() * sizeof(DevJob)) != hipSuccess) return c->fail(DFR_E_MEMORY, "job table");
        hipMemcpyAsync(c->job_buf.ptr, passes.data(), passes.size() * sizeof(DevJob), hipMemcpyHostToDevice, c->stream);
        size_t k = 0;
        for (int j = 0; j < n; j++)
            for (int s = 0; s < jobs[j].spp; s++, k++)
                launch_megakernel((const DevScene*)c->scene_table.ptr, (const DevJob*)c->job_buf.ptr + k,
                                  (float4*)film.ptr, (int*)c->overflow.ptr, jobs[j].width * jobs[j].height, 256,
                                  c->any_ext_features(), (int64_t*)c->counters.ptr, c->stream);
        hipEventRecord(c->ev1, c->stream);
        hipError_t e = hipEventSynchronize(c->ev1);
        if (e != hipSuccess) return c->hip_fail(e, "plain batch");
        int64_t cnt[2];
        hipMemcpy(cnt, c->counters.ptr, sizeof(cnt), hipMemcpyDeviceToHost);
        st.rays_closest = cnt[0]; st.rays_any = cnt[1];
    } else {
        if ((rc = c->upload_jobs(dj))) return rc;
        Counters out{};
        if ((rc = c->run_wavefront(n, work, (float4*)film.ptr, &out))) return rc;
        hipEventRecord(c->ev1, c->stream);
        hipError_t e = hipEventSynchronize(c->ev1);
        if (e != hipSuccess) return c->hip_fail(e, "packed batch");
        st.rays_closest = out.rays_closest; st.rays_any = out.rays_any; st.bounces = out.bounces;
        st.extend_seconds = c->prof_stats.extend_seconds; st.shade_seconds = c->prof_stats.shade_seconds;
        st.shadow_seconds = c->prof_stats.shadow_seconds; st.extend_launches = c->prof_stats.extend_launches;
        st.extend_rays = c->prof_stats.extend_rays; st.extend_nodes = c->prof_stats.extend_nodes;
        st.extend_tris = c->prof_stats.extend_tris; st.shadow_rays = c->prof_stats.shadow_rays;
        st.shadow_nodes = c->prof_stats.shadow_nodes; st.shadow_tris = c->prof_stats.shadow_tris;
    }
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    st.seconds = ms * 1e-3;
    st.path_samples = work;
    c->stats = st;
    if (films) {
        std::vector<float4> host(pixels);
        hipError_t e = hipMemcpy(host.data(), film.ptr, pixels * sizeof(float4), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return c->hip_fail(e, "batch readback");
        for (int j = 0; j < n; j++) {
            if (!films[j]) continue;
            int64_t np = (int64_t)jobs[j].width * jobs[j].height;
            for (int64_t i = 0; i < np; i++) {
                const float4& f = host[dj[j].film_offset + i];
                float inv = f.w > 0.0f ? 1.0f / f.w : 0.0f;
                films[j][3 * i] = f.x * inv; films[j][3 * i + 1] = f.y * inv; films[j][3 * i + 2] = f.z * inv;
            }
        }
    }
    return DFR_OK;
}


// ---------------------------------------------------------------------------
// Autotuner: times the packed launch on a reduced copy of the job list (spp
// capped) for every (block, pool capacity, sort keys) candidate and keeps the
// fastest on the context.
// ---------------------------------------------------------------------------
int autotune_impl(Ctx* c, const dfr_job* jobs, int n, int32_t* out) {
    int rc = validate_jobs(c, jobs, n);
    if (rc) return rc;
    std::vector<dfr_job> probe(jobs, jobs + n);
    int64_t work = 0;
    for (auto& j : probe) { j.spp = std::min(j.spp, 4); work += (int64_t)j.width * j.height * j.spp; }
    const int blocks[] = {128, 256};
    const int caps[] = {1 << 20, 1 << 22, 1 << 24, 1 << 25, 1 << 26};
    const int sorts[] = {0, 1};
    WavefrontConfig best = c->wf_params;
    double best_t = 1e30;
    for (int b : blocks)
        for (int cap : caps)
            for (int so : sorts) {
                c->wf_params.block = b; c->wf_params.capacity = cap; c->wf_params.sort_keys = so;
                // warm the allocation, then time
                if ((rc = batch_render_impl(c, probe.data(), n, DFR_BATCH_PACKED, nullptr))) return rc;
                if ((rc = batch_render_impl(c, probe.data(), n, DFR_BATCH_PACKED, nullptr))) return rc;
                double t = c->stats.seconds;
                if (t < best_t) { best_t = t; best = c->wf_params; }
            }
    c->wf_params = best;
    if (out) { out[0] = best.block; out[1] = best.capacity; out[2] = best.sort_keys; }
    return DFR_OK;
}

// ---------------------------------------------------------------------------
// Serving loop with continuous batching.  A worker thread owns the context's
// wavefront engine: it admits requests while their films fit the HBM budget
// (and the active cap), appends them to the device job table (the generate
// kernel pulls work from every admitted job in admission order), runs
// wavefront iterations, and retires a request once all of its paths have
// finished (per-job counters on the device), copying its film out.
// ---------------------------------------------------------------------------
struct Arena {  // first-fit allocator over a film arena (units: float4)
    std::map<int64_t, int64_t> free_;  // offset -> size
    int64_t used = 0, peak = 0;
    void init(int64_t n) { free_.clear(); free_[0] = n; used = 0; peak = 0; }
    int64_t alloc(int64_t n) {
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                int64_t off = it->first, rest = it->second - n;
                free_.erase(it);
                if (rest > 0) free_[off + n] = rest;
                used += n; peak = std::max(peak, used);
                return off;
            }
        return -1;
    }
    void release(int64_t off, int64_t n) {
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) { it->second += nx->second; free_.erase(nx); }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) { pv->second += it->second; free_.erase(it); }
        }
    }
};

struct Request {
    dfr_job job;
    int state = 0;  // 0 pending, 1 active, 2 done, -1 rejected
    int slot = -1;
    int64_t film_off = 0, pixels = 0, work = 0;
    std::chrono::steady_clock::time_point t_submit, t_done;
    std::vector<float> film;
};

struct Server {
    Ctx* ctx = nullptr;
    dfr_server_opts opts{};
    std::thread worker;
    std::mutex m;
    std::condition_variable cv, done_cv;
    bool stopping = false;
    std::deque<int64_t> pending;
    std::vector<std::unique_ptr<Request>> reqs;
    std::vector<int64_t> active;
    // device state
    DeviceBuffer arena_mem, jobs_mem, done_mem;
    Arena arena;
    int job_capacity = 1 << 16;
    int n_slots = 0;
    int slots_scene = -2;  // the scene of every admitted request (-2: none yet, -1: more than one)
    int64_t work_total = 0;
    int64_t arena_pixels = 0;
    dfr_server_stats stats{};
    std::string error;
    std::chrono::steady_clock::time_point t_busy0;
    bool busy = false;
};

static int64_t film_pixels(const dfr_job& j) { return (int64_t)j.width * j.height; }

static void server_loop(Server* S) {
    Ctx* c = S->ctx;
    std::lock_guard<std::mutex> ctx_lock(c->mutex);  // the server owns the context while it runs
    hipSetDevice(c->device);
    if (c->ensure_wavefront(S->opts.pool_paths) || c->ensure_overflow(wavefront_stack_threads(S->opts.pool_paths))) {
        std::lock_guard<std::mutex> l(S->m);
        S->error = c->error;
        S->stopping = true;
        S->done_cv.notify_all();
        return;
    }
    c->wf.overflow = (int*)c->overflow.ptr;
    c->wf.job_done = (int*)S->done_mem.ptr;
    Counters zero{};
    hipMemcpyAsync(c->wf.ctr, &zero, sizeof(Counters), hipMemcpyHostToDevice, c->stream);
    hipMemsetAsync(S->done_mem.ptr, 0, (size_t)S->job_capacity * sizeof(int), c->stream);
    WavefrontParams p{c->wf_params.block, c->wf_params.sort_keys, 0, c->any_area_lights()};
    p.ext_features = c->any_ext_features();
    std::vector<int> done_host;
    Counters host{};
    for (;;) {
        {
            std::unique_lock<std::mutex> l(S->m);
            // admission: FIFO while the film fits the budget and the active cap
            while (!S->pending.empty()) {
                Request& r = *S->reqs[S->pending.front()];
                if ((int)S->active.size() >= S->opts.max_active) break;
                if (S->n_slots >= S->job_capacity) { r.state = -1; S->stats.rejected++; S->pending.pop_front(); S->done_cv.notify_all(); continue; }
                if (r.pixels > S->arena_pixels) { r.state = -1; S->stats.rejected++; S->pending.pop_front(); S->done_cv.notify_all(); continue; }
                int64_t off = S->arena.alloc(r.pixels);
                if (off < 0) break;  // wait for films to be released
                r.film_off = off;
                r.slot = S->n_slots++;
                r.state = 1;
                S->slots_scene = S->slots_scene == -2 || S->slots_scene == r.job.scene ? r.job.scene : -1;
                DevJob dj = to_dev(r.job, off, S->work_total);
                S->work_total += r.work;
                hipMemcpyAsync((DevJob*)S->jobs_mem.ptr + r.slot, &dj, sizeof(DevJob), hipMemcpyHostToDevice, c->stream);
                hipMemsetAsync((float4*)S->arena_mem.ptr + off, 0, r.pixels * sizeof(float4), c->stream);
                S->active.push_back(S->pending.front());
                S->pending.pop_front();
                S->stats.peak_active = std::max<int32_t>(S->stats.peak_active, (int32_t)S->active.size());
                S->stats.peak_bytes = std::max<int64_t>(S->stats.peak_bytes, S->arena.used * (int64_t)sizeof(float4) +
                                                                                 (int64_t)wavefront_bytes(S->opts.pool_paths));
            }
            if (S->active.empty()) {
                if (S->busy) {
                    S->stats.busy_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - S->t_busy0).count();
                    S->busy = false;
                }
                if (S->stopping) break;
                S->cv.wait(l, [&] { return S->stopping || !S->pending.empty(); });
                continue;
            }
            if (!S->busy) { S->busy = true; S->t_busy0 = std::chrono::steady_clock::now(); }
        }
        // publish the new work total, run iterations (scene-uniform traversal
        // while every admitted request uses one scene)
        hipMemcpyAsync(&c->wf.ctr->work_total, &S->work_total, sizeof(int64_t), hipMemcpyHostToDevice, c->stream);
        p.uscene = S->slots_scene >= 0 ? (const DevScene*)c->scene_table.ptr + S->slots_scene : nullptr;
        for (int k = 0; k < 2; k++)
            wavefront_iteration((const DevScene*)c->scene_table.ptr, (const DevJob*)S->jobs_mem.ptr, S->n_slots,
                                (float4*)S->arena_mem.ptr, c->wf, p, nullptr, c->stream);
        // retire finished requests
        std::vector<int64_t> act;
        {
            std::lock_guard<std::mutex> l(S->m);
            act = S->active;
        }
        int lo = S->n_slots, hi = 0;
        for (int64_t id : act) { lo = std::min(lo, S->reqs[id]->slot); hi = std::max(hi, S->reqs[id]->slot + 1); }
        done_host.resize(std::max(0, hi - lo));
        hipMemcpyAsync(done_host.data(), (int*)S->done_mem.ptr + lo, (hi - lo) * sizeof(int), hipMemcpyDeviceToHost, c->stream);
        hipMemcpyAsync(&host, c->wf.ctr, sizeof(Counters), hipMemcpyDeviceToHost, c->stream);
        if (hipStreamSynchronize(c->stream) != hipSuccess) {
            std::lock_guard<std::mutex> l(S->m);
            S->error = "device error in the serving loop";
            S->stopping = true;
            for (int64_t id : S->active) S->reqs[id]->state = -1;
            S->active.clear();
            S->done_cv.notify_all();
            break;
        }
        std::vector<int64_t> finished;
        for (int64_t id : act) {
            Request& r = *S->reqs[id];
            if (done_host[r.slot - lo] >= r.work) finished.push_back(id);
        }
        for (int64_t id : finished) {
            Request& r = *S->reqs[id];
            std::vector<float4> f(r.pixels);
            hipMemcpy(f.data(), (float4*)S->arena_mem.ptr + r.film_off, r.pixels * sizeof(float4), hipMemcpyDeviceToHost);
            r.film.resize(3 * r.pixels);
            for (int64_t i = 0; i < r.pixels; i++) {
                float inv = f[i].w > 0.0f ? 1.0f / f[i].w : 0.0f;
                r.film[3 * i] = f[i].x * inv; r.film[3 * i + 1] = f[i].y * inv; r.film[3 * i + 2] = f[i].z * inv;
            }
            std::lock_guard<std::mutex> l(S->m);
            S->arena.release(r.film_off, r.pixels);
            r.t_done = std::chrono::steady_clock::now();
            r.state = 2;
            S->stats.completed++;
            S->stats.path_samples += r.work;
            S->active.erase(std::find(S->active.begin(), S->active.end(), id));
        }
        if (!finished.empty()) {
            std::lock_guard<std::mutex> l(S->m);
            S->done_cv.notify_all();
        }
    }
    c->wf.job_done = nullptr;
}

Server* server_create(Ctx* c, const dfr_server_opts& o) {
    auto S = std::make_unique<Server>();
    S->ctx = c;
    S->opts = o;
    if (S->opts.max_active <= 0) S->opts.max_active = 256;
    if (S->opts.pool_paths <= 0) S->opts.pool_paths = c->wf_params.capacity;
    if (S->opts.pool_paths > MAX_POOL_PATHS) { c->error = "pool_paths above 2^26"; return nullptr; }
    std::lock_guard<std::mutex> lock(c->mutex);
    hipSetDevice(c->device);
    if (c->scenes.empty()) { c->error = "no resident scene"; return nullptr; }
    int64_t budget = o.hbm_budget_bytes > 0 ? o.hbm_budget_bytes : (int64_t)8 << 30;
    int64_t film_bytes = budget - (int64_t)wavefront_bytes(S->opts.pool_paths);
    if (film_bytes < (int64_t)sizeof(float4)) { c->error = "budget below the wavefront pool"; return nullptr; }
    S->arena_pixels = film_bytes / (int64_t)sizeof(float4);
    if (S->arena_mem.alloc((size_t)S->arena_pixels * sizeof(float4)) != hipSuccess ||
        S->jobs_mem.alloc((size_t)S->job_capacity * sizeof(DevJob)) != hipSuccess ||
        S->done_mem.alloc((size_t)S->job_capacity * sizeof(int)) != hipSuccess) {
        c->error = "server allocation";
        return nullptr;
    }
    S->arena.init(S->arena_pixels);
    Server* raw = S.release();
    raw->worker = std::thread(server_loop, raw);
    return raw;
}

void server_destroy(Server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> l(S->m);
        S->stopping = true;
        // drop requests that never started
        for (int64_t id : S->pending) S->reqs[id]->state = -1;
        S->pending.clear();
    }
    S->cv.notify_all();
    if (S->worker.joinable()) S->worker.join();
    {
        std::lock_guard<std::mutex> lock(S->ctx->mutex);
        hipSetDevice(S->ctx->device);
        S->arena_mem.release(); S->jobs_mem.release(); S->done_mem.release();
    }
    delete S;
}

int64_t server_submit(Server* S, const dfr_job& job) {
    if (job.width < 1 || job.height < 1 || job.spp < 1) return DFR_E_ARGUMENT;
    if (job.scene < 0 || job.scene >= (int)S->ctx->scenes.size() || !S->ctx->scenes[job.scene]) return DFR_E_ARGUMENT;
    auto r = std::make_unique<Request>();
    r->job = job;
    r->pixels = film_pixels(job);
    r->work = r->pixels * job.spp;
    r->t_submit = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> l(S->m);
    if (S->stopping) return DFR_E_STATE;
    int64_t id = (int64_t)S->reqs.size();
    S->reqs.push_back(std::move(r));
    S->pending.push_back(id);
    S->cv.notify_all();
    return id;
}

int server_poll(Server* S, int64_t id) {
    std::lock_guard<std::mutex> l(S->m);
    if (id < 0 || id >= (int64_t)S->reqs.size()) return DFR_E_ARGUMENT;
    int st = S->reqs[id]->state;
    return st == 2 ? 1 : (st < 0 ? DFR_E_MEMORY : 0);
}

int server_wait(Server* S, int64_t id, float* film) {
    std::unique_lock<std::mutex> l(S->m);
    if (id < 0 || id >= (int64_t)S->reqs.size()) return DFR_E_ARGUMENT;
    Request& r = *S->reqs[id];
    S->done_cv.wait(l, [&] { return r.state == 2 || r.state < 0; });
    if (r.state < 0) return DFR_E_MEMORY;
    if (film) std::memcpy(film, r.film.data(), r.film.size() * sizeof(float));
    return DFR_OK;
}

int server_latency(Server* S, int64_t id, double* sec) {
    std::lock_guard<std::mutex> l(S->m);
    if (id < 0 || id >= (int64_t)S->reqs.size()) return DFR_E_ARGUMENT;
    Request& r = *S->reqs[id];
    if (r.state != 2) return DFR_E_STATE;
    *sec = std::chrono::duration<double>(r.t_done - r.t_submit).count();
    return DFR_OK;
}

int server_stats(Server* S, dfr_server_stats* st) {
    std::lock_guard<std::mutex> l(S->m);
    *st = S->stats;
    if (S->busy) st->busy_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - S->t_busy0).count();
    return DFR_OK;
}

}  // namespace dfr

extern "C" {

int32_t dfr_scene_add(dfr_ctx* h, const dfr_scene_desc* scene) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    SceneHost host;
    int rc = copy_scene_desc(scene, host, c->error);
    if (rc) return rc;
    auto cs = std::make_unique<CompiledScene>();
    rc = compile_scene(host, *cs, c->light_pdf_mode, c->stream, c->error);
    if (rc) return rc;
    int id = -1;
    for (size_t i = 0; i < c->scenes.size(); i++)
        if (!c->scenes[i] && (int)i != c->main_scene) { id = (int)i; break; }
    if (id < 0) { id = (int)c->scenes.size(); c->scenes.push_back(nullptr); }
    c->scenes[id] = std::move(cs);
    if ((rc = c->sync_scene_table())) return rc;
    return id;
}

int32_t dfr_scene_remove(dfr_ctx* h, int32_t scene) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (scene < 0 || scene >= (int)c->scenes.size() || !c->scenes[scene] || scene == c->main_scene)
        return c->fail(DFR_E_ARGUMENT, "no such resident scene");
    hipStreamSynchronize(c->stream);
    c->scenes[scene].reset();
    return c->sync_scene_table();
}

int32_t dfr_batch_render(dfr_ctx* h, const dfr_job* jobs, int32_t n, int32_t mode, float* const* films) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (!jobs) return c->fail(DFR_E_NULL, "jobs == null");
    if (n < 1) return c->fail(DFR_E_ARGUMENT, "n_jobs must be >= 1");
    if (mode != DFR_BATCH_PLAIN && mode != DFR_BATCH_PACKED) return c->fail(DFR_E_ARGUMENT, "unknown batch mode");
    return batch_render_impl(c, jobs, n, mode, films);
}

int32_t dfr_autotune(dfr_ctx* h, const dfr_job* jobs, int32_t n, int32_t* out) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (!jobs) return c->fail(DFR_E_NULL, "jobs == null");
    if (n < 1) return c->fail(DFR_E_ARGUMENT, "n_jobs must be >= 1");
    return autotune_impl(c, jobs, n, out);
}

int64_t dfr_wavefront_bytes(int32_t pool_paths) {
    return pool_paths > 0 ? (int64_t)wavefront_bytes(pool_paths) : DFR_E_ARGUMENT;
}

dfr_server* dfr_server_create(dfr_ctx* h, const dfr_server_opts* opts) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    if (!c || !opts) return nullptr;
    return reinterpret_cast<dfr_server*>(server_create(c, *opts));
}

void dfr_server_destroy(dfr_server* s) { server_destroy(reinterpret_cast<Server*>(s)); }

int64_t dfr_server_submit(dfr_server* s, const dfr_job* job) {
    if (!s || !job) return DFR_E_NULL;
    return server_submit(reinterpret_cast<Server*>(s), *job);
}

int32_t dfr_server_poll(dfr_server* s, int64_t r) {
    if (!s) return DFR_E_NULL;
    return server_poll(reinterpret_cast<Server*>(s), r);
}

int32_t dfr_server_wait(dfr_server* s, int64_t r, float* film) {
    if (!s) return DFR_E_NULL;
    return server_wait(reinterpret_cast<Server*>(s), r, film);
}

int32_t dfr_server_latency(dfr_server* s, int64_t r, double* sec) {
    if (!s || !sec) return DFR_E_NULL;
    return server_latency(reinterpret_cast<Server*>(s), r, sec);
}

int32_t dfr_server_get_stats(dfr_server* s, dfr_server_stats* st) {
    if (!s || !st) return DFR_E_NULL;
    return server_stats(reinterpret_cast<Server*>(s), st);
}

}  // extern "C"
