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
    const WavefrontConfig original = c->wf_params;
    WavefrontConfig best = original;
    double best_t = 1e30;
    for (int b : blocks)
        for (int cap : caps)
            for (int so : sorts) {
                c->wf_params.block = b; c->wf_params.capacity = cap; c->wf_params.sort_keys = so;
                // warm the allocation, then time; a candidate whose buffers do
                // not fit is skipped
                rc = batch_render_impl(c, probe.data(), n, DFR_BATCH_PACKED, nullptr);
                if (rc == DFR_E_MEMORY) continue;
                if (!rc) rc = batch_render_impl(c, probe.data(), n, DFR_BATCH_PACKED, nullptr);
                if (rc) { c->wf_params = best_t < 1e30 ? best : original; return rc; }
                double t = c->stats.seconds;
                if (t < best_t) { best_t = t; best = c->wf_params; }
            }
    if (best_t >= 1e30) { c->wf_params = original; return c->fail(DFR_E_MEMORY, "no autotune candidate fits"); }
    c->wf_params = best;
    if (out) { out[0] = best.block; out[1] = best.capacity; out[2] = best.sort_keys; }
    return DFR_OK;
}


// ---------------------------------------------------------------------------
// Serving loop with continuous batching.  A worker thread owns a wavefront
// engine of its own (path pool, queues, traversal overflow, stream).  Each
// round it admits pending requests while their films fit the HBM budget and
// a job slot is free (slots of retired requests are reused), publishes the
// new work, and enqueues ROUND_ITERATIONS wavefront iterations followed by
// asynchronous copies of the queue counters and the per-slot finished-path
// counts.  While the device runs that round the worker reads the previous
// round's counts: a request whose paths have all finished has its film
// copied out asynchronously, and once that copy is done its film area and
// slot are released.  The context's mutex is held only while a round is
// enqueued (the scene table must not change under it), so the context's
// other calls interleave with serving.
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
    int state = 0;  // 0 pending, 1 active, 3 copying its film out, 2 done, -1 rejected
    int slot = -1;
    int64_t film_off = 0, pixels = 0, work = 0, work_begin = 0;
    int admitted_round = -1;
    float4* staging = nullptr;  // pinned host copy of the film while state 3
    hipEvent_t copied = nullptr;
    std::chrono::steady_clock::time_point t_submit, t_done;
    std::vector<float> film;
};

constexpr int ROUND_ITERATIONS = 4;

struct Server {
    Ctx* ctx = nullptr;
    dfr_server_opts opts{};
    std::thread worker;
    std::mutex m;
    std::condition_variable cv, done_cv;
    bool stopping = false;
    std::deque<int64_t> pending;
    std::vector<std::unique_ptr<Request>> reqs;
    std::vector<int64_t> active;     // admitted and not yet copied out, in admission order
    std::deque<int64_t> copying;     // film copies in flight, in the order they were enqueued
    std::thread finisher;            // completes copied films off the worker thread
    std::condition_variable copy_cv;
    bool worker_done = false;
    uint64_t submits = 0, releases = 0;  // bumped by submissions and by freed films (wake the idle worker)
    std::multimap<size_t, float4*> staging_free;  // pinned film buffers for reuse, by size
    std::map<float4*, size_t> staging_sizes;      // every pinned film buffer and its size
    // device state, owned by the server
    int device = 0;
    hipStream_t stream = nullptr;
    WavefrontBuffers wf{};
    DeviceBuffer wf_mem[17], overflow, arena_mem, jobs_mem, done_mem, order_mem, table_mem;
    int overflow_depth = 0;
    uint64_t table_gen = ~0ull;
    Arena arena;
    int slot_capacity = 1 << 16;
    int slot_hw = 0;                 // slots handed out so far (the job table's used length)
    std::vector<int> free_slots;
    int64_t work_total = 0;
    int64_t work_next_seen = 0;      // the device's work_next at the last counter read
    int64_t arena_pixels = 0;
    // pinned host mirrors: two rounds in flight at most
    int* done_host[2] = {nullptr, nullptr};
    Counters* ctr_host[2] = {nullptr, nullptr};
    int64_t* total_host[2] = {nullptr, nullptr};
    int done_len[2] = {0, 0};
    hipEvent_t round_ev[2] = {nullptr, nullptr};
    std::vector<DevJob> job_upload[2];
    std::vector<GenEntry> order_upload[2];
    dfr_server_stats stats{};
    std::string error;
    std::chrono::steady_clock::time_point t_busy0;
    bool busy = false;
};

static int64_t film_pixels(const dfr_job& j) { return (int64_t)j.width * j.height; }

// Marks every unfinished request failed (device error), under S->m.
static void fail_all(Server* S, const std::string& msg) {
    S->error = msg;
    S->stopping = true;
    for (int64_t id : S->active) S->reqs[id]->state = -1;
    for (int64_t id : S->copying) S->reqs[id]->state = -1;
    for (int64_t id : S->pending) S->reqs[id]->state = -1;
    S->active.clear(); S->pending.clear();
    S->done_cv.notify_all();
}

// A pinned staging buffer of at least `n` float4 for a film copy.
static float4* take_staging(Server* S, size_t n) {
    {
        std::lock_guard<std::mutex> l(S->m);
        auto it = S->staging_free.lower_bound(n);
        if (it != S->staging_free.end()) {
            float4* p = it->second;
            S->staging_free.erase(it);
            return p;
        }
    }
    size_t cap = 1;
    while (cap < n) cap <<= 1;  // power-of-two sizes keep the pool reusable
    float4* p = nullptr;
    if (hipHostMalloc((void**)&p, cap * sizeof(float4), hipHostMallocDefault) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> l(S->m);
    S->staging_sizes[p] = cap;
    return p;
}

// Completes copied films off the worker thread: waits for each copy in
// order, turns (sum, count) into means, then frees the film area, the slot
// and the staging buffer.
static void finisher_loop(Server* S) {
    hipSetDevice(S->device);
    for (;;) {
        int64_t id;
        {
            std::unique_lock<std::mutex> l(S->m);
            S->copy_cv.wait(l, [&] { return !S->copying.empty() || S->worker_done; });
            if (S->copying.empty()) break;
            id = S->copying.front();
            S->copying.pop_front();
        }
        Request& r = *S->reqs[id];
        const hipError_t e = hipEventSynchronize(r.copied);
        if (e == hipSuccess) {
            std::vector<float> film(3 * r.pixels);
            for (int64_t i = 0; i < r.pixels; i++) {
                const float4 f = r.staging[i];
                const float inv = f.w > 0.0f ? 1.0f / f.w : 0.0f;
                film[3 * i] = f.x * inv; film[3 * i + 1] = f.y * inv; film[3 * i + 2] = f.z * inv;
            }
            r.film.swap(film);
        }
        hipEventDestroy(r.copied);
        std::lock_guard<std::mutex> l(S->m);
        r.copied = nullptr;
        S->staging_free.emplace(S->staging_sizes[r.staging], r.staging);
        r.staging = nullptr;
        S->arena.release(r.film_off, r.pixels);
        S->free_slots.push_back(r.slot);
        S->releases++;
        S->cv.notify_all();
        r.t_done = std::chrono::steady_clock::now();
        if (r.state == 3) {
            r.state = e == hipSuccess ? 2 : -1;
            if (e == hipSuccess) { S->stats.completed++; S->stats.path_samples += r.work; }
        }
        S->done_cv.notify_all();
    }
}

static void server_loop(Server* S) {
    Ctx* c = S->ctx;
    hipSetDevice(S->device);
    WavefrontParams p{};
    {
        std::lock_guard<std::mutex> cl(c->mutex);
        p.block = c->wf_params.block;
        p.sort_keys = c->wf_params.sort_keys;
    }
    hipMemsetAsync(S->done_mem.ptr, 0, (size_t)S->slot_capacity * sizeof(int), S->stream);
    Counters zero{};
    hipMemcpyAsync(S->wf.ctr, &zero, sizeof(Counters), hipMemcpyHostToDevice, S->stream);
    hipStreamSynchronize(S->stream);
    int round = 0;
    bool prev_pending = false;  // the previous round's counters are on their way
    for (;;) {
        const int cur = round & 1, prev = cur ^ 1;
        // ---- admission --------------------------------------------------
        std::vector<int64_t> admitted;
        int scene_uniform = -2;  // the scene of every active request (-1: more than one)
        {
            std::unique_lock<std::mutex> l(S->m);
            while (!S->pending.empty()) {
                Request& r = *S->reqs[S->pending.front()];
                if ((int)S->active.size() >= S->opts.max_active) break;
                if (r.pixels > S->arena_pixels) {
                    r.state = -1; S->stats.rejected++; S->pending.pop_front(); S->done_cv.notify_all(); continue;
                }
                if (S->free_slots.empty() && S->slot_hw >= S->slot_capacity) break;  // wait for a retirement
                const int64_t off = S->arena.alloc(r.pixels);
                if (off < 0) break;  // wait for films to be released
                int slot;
                if (!S->free_slots.empty()) { slot = S->free_slots.back(); S->free_slots.pop_back(); }
                else slot = S->slot_hw++;
                r.film_off = off;
                r.slot = slot;
                r.state = 1;
                r.work_begin = S->work_total;
                r.admitted_round = round;
                S->work_total += r.work;
                admitted.push_back(S->pending.front());
                S->active.push_back(S->pending.front());
                S->pending.pop_front();
                S->stats.peak_active = std::max<int32_t>(S->stats.peak_active, (int32_t)S->active.size());
                S->stats.peak_bytes = std::max<int64_t>(S->stats.peak_bytes, S->arena.used * (int64_t)sizeof(float4) +
                                                                                 (int64_t)wavefront_bytes(S->opts.pool_paths));
            }
            for (int64_t id : S->active) {
                const int sc = S->reqs[id]->job.scene;
                scene_uniform = scene_uniform == -2 || scene_uniform == sc ? sc : -1;
            }
            if (S->active.empty() && !prev_pending) {
                if (S->busy) {
                    S->stats.busy_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - S->t_busy0).count();
                    S->busy = false;
                }
                if (S->stopping) break;
                // nothing runs: wait for a new request, or for a copied-out
                // film to free the room the pending ones need
                const uint64_t sub0 = S->submits, rel0 = S->releases;
                S->cv.wait(l, [&] { return S->stopping || S->submits != sub0 || S->releases != rel0; });
                continue;
            }
            if (!S->busy) { S->busy = true; S->t_busy0 = std::chrono::steady_clock::now(); }
        }
        const bool run = !S->active.empty();
        if (run) {
            // ---- enqueue one round under the context's mutex --------------
            std::lock_guard<std::mutex> cl(c->mutex);
            std::string err;
            // the server's copy of the scene table (pointers into resident scenes)
            if (S->table_gen != c->scene_gen) {
                std::vector<DevScene> table(std::max<size_t>(1, c->scenes.size()));
                for (size_t i = 0; i < c->scenes.size(); i++) if (c->scenes[i]) table[i] = c->scenes[i]->dev;
                if (S->table_mem.bytes < table.size() * sizeof(DevScene)) {
                    hipStreamSynchronize(S->stream);  // kernels in flight still read the old table
                    if (S->table_mem.alloc(table.size() * sizeof(DevScene)) != hipSuccess) err = "server scene table";
                }
                if (err.empty()) {
                    // pageable source: the copy has completed or been staged when the call returns
                    hipMemcpy(S->table_mem.ptr, table.data(), table.size() * sizeof(DevScene), hipMemcpyHostToDevice);
                    S->table_gen = c->scene_gen;
                }
            }
            // traversal overflow for the deepest admitted scene
            int depth = 1;
            for (int64_t id : S->active) depth = std::max(depth, c->scenes[S->reqs[id]->job.scene]->max_depth);
            if (err.empty() && depth > S->overflow_depth) {
                hipStreamSynchronize(S->stream);
                const size_t need = (size_t)overflow_entries_per_thread(depth) *
                                    (size_t)wavefront_stack_threads(S->opts.pool_paths) * sizeof(int);
                if (S->overflow.alloc(need) != hipSuccess) err = "server traversal overflow";
                else { S->overflow_depth = depth; S->wf.overflow = (int*)S->overflow.ptr; }
            }
            if (!err.empty()) {
                std::lock_guard<std::mutex> l(S->m);
                fail_all(S, err);
                break;
            }
            // new jobs: records, cleared films and finished-path counts
            std::vector<DevJob>& up = S->job_upload[cur];
            up.clear();
            for (int64_t id : admitted) {
                Request& r = *S->reqs[id];
                up.push_back(to_dev(r.job, r.film_off, r.work_begin));
            }
            for (size_t k = 0; k < admitted.size(); k++) {
                Request& r = *S->reqs[admitted[k]];
                hipMemcpyAsync((DevJob*)S->jobs_mem.ptr + r.slot, &up[k], sizeof(DevJob), hipMemcpyHostToDevice, S->stream);
                hipMemsetAsync((float4*)S->arena_mem.ptr + r.film_off, 0, r.pixels * sizeof(float4), S->stream);
                hipMemsetAsync((int*)S->done_mem.ptr + r.slot, 0, sizeof(int), S->stream);
            }
            // generation order: the active requests that may still have work
            // to generate, in admission (work_begin) order
            std::vector<GenEntry>& ord = S->order_upload[cur];
            ord.clear();
            for (int64_t id : S->active) {
                const Request& r = *S->reqs[id];
                if (r.work_begin + r.work > S->work_next_seen) ord.push_back(GenEntry{r.work_begin, r.slot, 0});
            }
            if (ord.empty()) ord.push_back(GenEntry{0, 0, 0});  // nothing left to generate
            hipMemcpyAsync(S->order_mem.ptr, ord.data(), ord.size() * sizeof(GenEntry), hipMemcpyHostToDevice, S->stream);
            *S->total_host[cur] = S->work_total;
            hipMemcpyAsync(&S->wf.ctr->work_total, S->total_host[cur], sizeof(int64_t), hipMemcpyHostToDevice, S->stream);
            p.closest_shadow = c->any_area_lights();
            p.ext_features = c->any_ext_features();
            p.items = c->any_items();
            p.uscene = scene_uniform >= 0 ? (const DevScene*)S->table_mem.ptr + scene_uniform : nullptr;
            p.order = (const GenEntry*)S->order_mem.ptr;
            p.n_order = (int)ord.size();
            for (int k = 0; k < ROUND_ITERATIONS; k++)
                wavefront_iteration((const DevScene*)S->table_mem.ptr, (const DevJob*)S->jobs_mem.ptr, S->slot_hw,
                                    (float4*)S->arena_mem.ptr, S->wf, p, nullptr, S->stream);
            S->done_len[cur] = S->slot_hw;
            hipMemcpyAsync(S->done_host[cur], S->done_mem.ptr, (size_t)S->slot_hw * sizeof(int), hipMemcpyDeviceToHost,
                           S->stream);
            hipMemcpyAsync(S->ctr_host[cur], S->wf.ctr, sizeof(Counters), hipMemcpyDeviceToHost, S->stream);
            hipEventRecord(S->round_ev[cur], S->stream);
        }
        // ---- the previous round's results, while this round runs ---------
        if (prev_pending) {
            if (hipEventSynchronize(S->round_ev[prev]) != hipSuccess) {
                std::lock_guard<std::mutex> l(S->m);
                fail_all(S, "device error in the serving loop");
                break;
            }
            S->work_next_seen = S->ctr_host[prev]->work_next;
            std::vector<int64_t> still;
            for (int64_t id : S->active) {
                Request& r = *S->reqs[id];
                const bool seen = r.admitted_round <= round - 1 && r.slot < S->done_len[prev];
                if (!seen || S->done_host[prev][r.slot] < r.work) { still.push_back(id); continue; }
                // every path finished: copy the film out behind the work in flight
                r.staging = take_staging(S, (size_t)r.pixels);
                if (!r.staging || hipEventCreateWithFlags(&r.copied, hipEventDisableTiming) != hipSuccess) {
                    std::lock_guard<std::mutex> l(S->m);
                    fail_all(S, "film staging allocation");
                    S->worker_done = true;
                    S->copy_cv.notify_all();
                    return;
                }
                hipMemcpyAsync(r.staging, (float4*)S->arena_mem.ptr + r.film_off, r.pixels * sizeof(float4),
                               hipMemcpyDeviceToHost, S->stream);
                hipEventRecord(r.copied, S->stream);
                std::lock_guard<std::mutex> l(S->m);
                r.state = 3;
                S->copying.push_back(id);
                S->copy_cv.notify_one();
            }
            std::lock_guard<std::mutex> l(S->m);
            S->active.swap(still);
        }
        prev_pending = run;
        round++;
    }
    std::lock_guard<std::mutex> l(S->m);
    S->worker_done = true;
    S->copy_cv.notify_all();
}

Server* server_create(Ctx* c, const dfr_server_opts& o) {
    auto S = std::make_unique<Server>();
    S->ctx = c;
    S->opts = o;
    if (S->opts.max_active <= 0) S->opts.max_active = 256;
    if (S->opts.pool_paths <= 0) S->opts.pool_paths = c->wf_params.capacity;
    if (S->opts.pool_paths > MAX_POOL_PATHS) { c->error = "pool_paths above 2^26"; return nullptr; }
    if (o.slot_capacity < 0) { c->error = "slot_capacity must be >= 0"; return nullptr; }
    S->slot_capacity = o.slot_capacity > 0 ? o.slot_capacity : 1 << 16;
    std::lock_guard<std::mutex> lock(c->mutex);
    hipSetDevice(c->device);
    S->device = c->device;
    if (c->scenes.empty()) { c->error = "no resident scene"; return nullptr; }
    int64_t budget = o.hbm_budget_bytes > 0 ? o.hbm_budget_bytes : (int64_t)8 << 30;
    int64_t film_bytes = budget - (int64_t)wavefront_bytes(S->opts.pool_paths);
    if (film_bytes < (int64_t)sizeof(float4)) { c->error = "budget below the wavefront pool"; return nullptr; }
    S->arena_pixels = film_bytes / (int64_t)sizeof(float4);
    bool ok = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && alloc_wavefront_buffers(S->wf, S->wf_mem, S->opts.pool_paths);
    ok = ok && S->arena_mem.alloc((size_t)S->arena_pixels * sizeof(float4)) == hipSuccess;
    ok = ok && S->jobs_mem.alloc((size_t)S->slot_capacity * sizeof(DevJob)) == hipSuccess;
    ok = ok && S->done_mem.alloc((size_t)S->slot_capacity * sizeof(int)) == hipSuccess;
    ok = ok && S->order_mem.alloc((size_t)(S->opts.max_active + 1) * sizeof(GenEntry)) == hipSuccess;
    for (int k = 0; k < 2 && ok; k++) {
        ok = hipHostMalloc((void**)&S->done_host[k], (size_t)S->slot_capacity * sizeof(int), hipHostMallocDefault) == hipSuccess &&
             hipHostMalloc((void**)&S->ctr_host[k], sizeof(Counters), hipHostMallocDefault) == hipSuccess &&
             hipHostMalloc((void**)&S->total_host[k], sizeof(int64_t), hipHostMallocDefault) == hipSuccess &&
             hipEventCreateWithFlags(&S->round_ev[k], hipEventDisableTiming) == hipSuccess;
    }
    auto cleanup = [&](Server* s) {
        for (int k = 0; k < 2; k++) {
            if (s->done_host[k]) hipHostFree(s->done_host[k]);
            if (s->ctr_host[k]) hipHostFree(s->ctr_host[k]);
            if (s->total_host[k]) hipHostFree(s->total_host[k]);
            if (s->round_ev[k]) hipEventDestroy(s->round_ev[k]);
        }
        if (s->stream) hipStreamDestroy(s->stream);
    };
    if (!ok) { cleanup(S.get()); c->error = "server allocation"; return nullptr; }
    S->wf.job_done = (int*)S->done_mem.ptr;
    S->arena.init(S->arena_pixels);
    Server* raw = S.release();
    raw->worker = std::thread(server_loop, raw);
    raw->finisher = std::thread(finisher_loop, raw);
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
    if (S->finisher.joinable()) S->finisher.join();
    Ctx* c = S->ctx;
    {
        std::lock_guard<std::mutex> lock(c->mutex);
        hipSetDevice(S->device);
        hipStreamSynchronize(S->stream);
        // release the scene references of every request that was accepted
        for (auto& r : S->reqs)
            if (r->job.scene >= 0 && r->job.scene < (int)c->scene_refs.size()) c->scene_refs[r->job.scene]--;
        for (auto& r : S->reqs)
            if (r->copied) hipEventDestroy(r->copied);
        for (auto& kv : S->staging_sizes) hipHostFree(kv.first);
        for (int k = 0; k < 2; k++) {
            hipHostFree(S->done_host[k]); hipHostFree(S->ctr_host[k]); hipHostFree(S->total_host[k]);
            hipEventDestroy(S->round_ev[k]);
        }
        S->arena_mem.release(); S->jobs_mem.release(); S->done_mem.release(); S->order_mem.release();
        S->table_mem.release(); S->overflow.release();
        for (auto& b : S->wf_mem) b.release();
        hipStreamDestroy(S->stream);
    }
    delete S;
}

int64_t server_submit(Server* S, const dfr_job& job) {
    if (job.width < 1 || job.height < 1 || job.spp < 1) return DFR_E_ARGUMENT;
    {
        // the scene must stay resident while a request renders it
        Ctx* c = S->ctx;
        std::lock_guard<std::mutex> cl(c->mutex);
        if (job.scene < 0 || job.scene >= (int)c->scenes.size() || !c->scenes[job.scene]) return DFR_E_ARGUMENT;
        if ((int)c->scene_refs.size() < (int)c->scenes.size()) c->scene_refs.resize(c->scenes.size(), 0);
        c->scene_refs[job.scene]++;
    }
    auto r = std::make_unique<Request>();
    r->job = job;
    r->pixels = film_pixels(job);
    r->work = r->pixels * job.spp;
    r->t_submit = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> l(S->m);
    int64_t id = (int64_t)S->reqs.size();
    S->reqs.push_back(std::move(r));
    if (S->stopping) { S->reqs[id]->state = -1; return DFR_E_STATE; }
    S->pending.push_back(id);
    S->submits++;
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
    if (film) {
        // the film is handed over once: its host copy is released
        if (r.film.empty() && r.pixels > 0) return DFR_E_STATE;
        std::memcpy(film, r.film.data(), r.film.size() * sizeof(float));
        std::vector<float>().swap(r.film);
    }
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
    rc = compile_scene(host, {}, *cs, c->light_pdf_mode, c->stream, c->error);
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
    if (scene < (int)c->scene_refs.size() && c->scene_refs[scene] > 0)
        return c->fail(DFR_E_STATE, "a serving loop still has requests on this scene");
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
