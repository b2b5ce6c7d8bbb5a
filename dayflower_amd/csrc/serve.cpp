// serve.cpp — resident scenes, batched jobs and the serving loop of the C-ABI.
#include <algorithm>
#include <cstring>

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
        if (c->counters.alloc(64) != hipSuccess) return c->fail(DFR_E_MEMORY, "counters");
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
                                  (int64_t*)c->counters.ptr, c->stream);
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
    (void)jobs; (void)n;
    if (out) { out[0] = c->wf_params.block; out[1] = c->wf_params.capacity; out[2] = c->wf_params.sort_keys; }
    return c->fail(DFR_E_UNSUPPORTED, "autotune not built yet");
}

dfr_server* dfr_server_create(dfr_ctx* h, const dfr_server_opts* opts) {
    (void)h; (void)opts;
    return nullptr;
}
void dfr_server_destroy(dfr_server* s) { (void)s; }
int64_t dfr_server_submit(dfr_server* s, const dfr_job* job) { (void)s; (void)job; return DFR_E_UNSUPPORTED; }
int32_t dfr_server_poll(dfr_server* s, int64_t r) { (void)s; (void)r; return DFR_E_UNSUPPORTED; }
int32_t dfr_server_wait(dfr_server* s, int64_t r, float* film) { (void)s; (void)r; (void)film; return DFR_E_UNSUPPORTED; }
int32_t dfr_server_latency(dfr_server* s, int64_t r, double* sec) { (void)s; (void)r; (void)sec; return DFR_E_UNSUPPORTED; }
int32_t dfr_server_get_stats(dfr_server* s, dfr_server_stats* st) { (void)s; (void)st; return DFR_E_UNSUPPORTED; }

}  // extern "C"
