// kernels.hip — HIP kernels of the path-tracing hot path for gfx950.
//
//  * megakernel: one work-item per pixel, all bounces in registers (the
//    plain per-job kernel; GPURenderer.run(), M/renderer/gpu/GPURenderer.java
//    :65-77, restated over the CPURenderer semantics of M/scene/Scene.java
//    :272-326, 451-718).
//  * wavefront: generate -> extend (closest hit) -> material sort -> shade ->
//    shadow (NEE visibility) -> commit, over a path pool in HBM with
//    block-aggregated queue appends (wave ballot + popcount prefix, one
//    atomic per block and queue).
//  * ray queries: closest/any hit for a ray array (hit-parity checks and the
//    traversal benchmark).
#include <mutex>

#include "device.h"
#include "kernels.h"

namespace dfr {

// ---------------------------------------------------------------------------
// wave-level helpers (wave64)
// ---------------------------------------------------------------------------
DEV int lane_id() { return __lane_id(); }

// Block-wide queue reservation.  Each thread asks for `a` continuing-path
// slots (0/1), `b` any-hit shadow slots (0..2) and `c` closest-hit shadow
// slots (0..2); the block makes one atomic on n_next and one 64-bit atomic on
// n_shadow2 (both shadow partitions at once).  Returns the thread's first slot
// of each kind.  Every thread of the block must call it.
template <int BLOCK>
DEV void block_reserve(Counters* ctr, bool a, int b, int c, int& sa, int& sb, int& sc) {
    constexpr int NW = BLOCK / 64;
    __shared__ int wsum[3][NW];
    __shared__ int wbase[3][NW];
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint64_t ma = __ballot(a), mb0 = __ballot(b & 1), mb1 = __ballot(b >> 1);
    const uint64_t mc0 = __ballot(c & 1), mc1 = __ballot(c >> 1);
    const int pa = __popcll(ma & lt);
    const int pb = __popcll(mb0 & lt) + 2 * __popcll(mb1 & lt);
    const int pc = __popcll(mc0 & lt) + 2 * __popcll(mc1 & lt);
    if (lane == 0) {
        wsum[0][w] = __popcll(ma);
        wsum[1][w] = __popcll(mb0) + 2 * __popcll(mb1);
        wsum[2][w] = __popcll(mc0) + 2 * __popcll(mc1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int ta = 0, tb = 0, tc = 0;
        for (int k = 0; k < NW; k++) {
            wbase[0][k] = ta; ta += wsum[0][k];
            wbase[1][k] = tb; tb += wsum[1][k];
            wbase[2][k] = tc; tc += wsum[2][k];
        }
        int ga = ta ? atomicAdd(&ctr->n_next, ta) : 0;
        unsigned long long g = 0;
        if (tb | tc)
            g = atomicAdd((unsigned long long*)&ctr->n_shadow2, (unsigned long long)tb | ((unsigned long long)tc << 32));
        const int gb = (int)(g & 0xffffffffull), gc = (int)(g >> 32);
        for (int k = 0; k < NW; k++) { wbase[0][k] += ga; wbase[1][k] += gb; wbase[2][k] += gc; }
    }
    __syncthreads();
    sa = wbase[0][w] + pa;
    sb = wbase[1][w] + pb;
    sc = wbase[2][w] + pc;
}

// Adds `n` to job_done[job] for every lane with pred, one atomic per distinct
// job in the wave.
DEV void wave_count_jobs(int* job_done, bool pred, uint32_t job) {
    uint64_t left = __ballot(pred);
    while (left) {
        const int leader = __ffsll((unsigned long long)left) - 1;
        const uint32_t j = __shfl(job, leader);
        const uint64_t same = __ballot(pred && job == j) & left;
        if (lane_id() == leader) atomicAdd(&job_done[j], __popcll(same));
        left &= ~same;
    }
}

// Adds the block's traversal visit counts to counter[0] (nodes) and
// counter[1] (triangles): wave sums, then one atomic pair per block.  Every
// thread of the block must call it.
template <int BLOCK>
DEV void add_visit_counts(unsigned long long* counter, const VisitCount& vc) {
    __shared__ unsigned long long part[2];
    unsigned long long a = vc.nodes, b = vc.tris;
    for (int off = 32; off > 0; off >>= 1) { a += __shfl_down(a, off); b += __shfl_down(b, off); }
    if (threadIdx.x == 0) { part[0] = 0; part[1] = 0; }
    __syncthreads();
    if (lane_id() == 0 && (a | b)) { atomicAdd(&part[0], a); atomicAdd(&part[1], b); }
    __syncthreads();
    if (threadIdx.x == 0 && (part[0] | part[1])) { atomicAdd(&counter[0], part[0]); atomicAdd(&counter[1], part[1]); }
}

DEV void film_add(float4* film, int64_t idx, V c) {
    if (!(isfinite(c.x) && isfinite(c.y) && isfinite(c.z))) return;
    if (c.x != 0.0f) unsafeAtomicAdd(&film[idx].x, c.x);
    if (c.y != 0.0f) unsafeAtomicAdd(&film[idx].y, c.y);
    if (c.z != 0.0f) unsafeAtomicAdd(&film[idx].z, c.z);
}

// ---------------------------------------------------------------------------
// NEE as deferred visibility requests (Scene.doEstimateDirectLight,
// M/scene/Scene.java:1530-1633, isSpecular = false).
//   mode 0: contribute c if nothing is hit in (T_MIN, tmax)          (delta / sky lights)
//   mode 1: contribute c if the closest hit in (T_MIN, tmax) is the light    (area lights)
//   mode 2: BSDF-sampled ray toward an area light: closest hit in
//           (T_MIN, inf); the light's emitted radiance there times c.  (A
//           BSDF-sampled ray toward an infinite light is a mode-0 request
//           whose c already holds the light's radiance along the ray.)
// ---------------------------------------------------------------------------
struct ShadowReq {
    V o, d;
    float tmax;
    int mode;   // 0, 1, 2 as above; -1: no request
    int light;
    V c;
};

// Fills `ls_req` with the light-sample request (mode 0 or 1) and `bs_req`
// with the BSDF-sample request (mode 2); a slot without a request keeps
// mode -1.  Two named slots (not an array) keep both in registers.
template <bool EXT>
DEV void nee_requests(const DevScene& sc, const Bsdf& b, const Surf& si, int li, const Rng& rng, uint32_t dim,
                      V scale, ShadowReq& ls_req, ShadowReq& bs_req) {
    ls_req.mode = -1;
    bs_req.mode = -1;
    const DevLight& L = sc.lights[li];
    const uint32_t T = BX_ALL_EXCEPT_SPECULAR;
    LightSmp ls;
    bool have = sample_light<EXT>(sc, L, si, rng.get(dim + 1), rng.get(dim + 2), rng.get(dim + 5), ls);
    V ns = si.s.w;
    if (delta_light<EXT>(L.type)) {
        if (have && !black(ls.L) && ls.pdf > 0.0f) {
            V f = bsdf_f<EXT>(b, T, ls.wi) * fabsf(dot(ls.wi, ns));
            if (!black(f)) {
                ShadowReq& r = ls_req;
                r.o = si.p; r.d = nrm(ls.p - si.p); r.tmax = fabsf(len(si.p - ls.p)) + 0.001f;
                r.mode = 0; r.light = li; r.c = scale * f * ls.L / ls.pdf;
            }
        }
        return;
    }
    if (have && !black(ls.L) && ls.pdf > 0.0f) {
        V f = bsdf_f<EXT>(b, T, ls.wi) * fabsf(dot(ls.wi, ns));
        float sp = bsdf_pdf<EXT>(b, T, ls.wi);
        if (!black(f)) {
            float w = power_heuristic(ls.pdf, sp);
            ShadowReq& r = ls_req;
            r.o = si.p; r.d = nrm(ls.p - si.p); r.tmax = fabsf(len(si.p - ls.p)) + 0.001f;
            r.mode = L.type == DFR_LIGHT_DIFFUSE_AREA ? 1 : 0; r.light = li;
            r.c = scale * f * ls.L * w / ls.pdf;
        }
    }
    LSample bs;
    if (bsdf_sample<EXT>(b, T, rng.get(dim + 3), rng.get(dim + 4), bs)) {
        V f = bs.f * fabsf(dot(bs.wi, ns));
        bool spec = (bs.type & BX_SPEC) != 0;
        if (!black(f) && bs.pdf > 0.0f) {
            float w = 1.0f;
            bool ok = true;
            // the sky's pdf and radiance along the sampled direction share
            // its object-space form
            const bool sky = L.type == DFR_LIGHT_PEREZ;
            const V wd = nrm(bs.wi);
            const V io = sky ? perez_to_object(sc.perez[L.perez], wd) : wd;
            if (!spec) {
                float lp = sky ? perez_pdf_object(sc.perez[L.perez], io) : light_pdf<EXT>(sc, L, si, bs.wi);
                if (lp == 0.0f) ok = false;
                else w = power_heuristic(bs.pdf, lp);
            }
            if (ok) {
                ShadowReq& r = bs_req;
                r.o = si.p; r.d = wd; r.tmax = FLT_MAX;
                r.light = li; r.c = scale * f * w / bs.pdf;
                if (L.type == DFR_LIGHT_DIFFUSE_AREA) {
                    r.mode = 2;
                } else {
                    // an infinite light is seen exactly when nothing is hit:
                    // its radiance folds into c and the query is any-hit
                    r.c = r.c * (sky ? perez_sky_object(sc.perez[L.perez], io) : emitted_ray<EXT>(sc, L, r.d));
                    r.mode = black(r.c) ? -1 : 0;
                }
            }
        }
    }
}

// Contribution of a traced request given its hit (`hit` false on a miss).
DEV V resolve_hit(const DevScene& sc, const ShadowReq& r, bool hit, const HitRec& h) {
    if (r.mode == 0) return hit ? mk(0, 0, 0) : r.c;
    if (r.mode == 1) return (hit && sc.prims[h.prim].area_light == r.light) ? r.c : mk(0, 0, 0);
    if (hit) {
        if (sc.prims[h.prim].area_light != r.light) return mk(0, 0, 0);
        Surf s2 = make_surface(sc, r.o, r.d, h);
        V Li = emitted(sc, s2, -r.d);
        return r.c * Li;
    }
    // closest-hit requests aim at area lights, which a miss does not see
    return r.c * emitted_ray<false>(sc, sc.lights[r.light], r.d);
}

template <int SHORT>
DEV V resolve_shadow(const DevScene& sc, const ShadowReq& r, TraversalStack<SHORT>& st, VisitCount& vc) {
    HitRec h;
    bool hit = r.mode == 0 ? trace<true, SHORT>(sc, r.o, r.d, T_MIN, r.tmax, h, st, vc)
                           : trace<false, SHORT>(sc, r.o, r.d, T_MIN, r.tmax, h, st, vc);
    return resolve_hit(sc, r, hit, h);
}

// ---------------------------------------------------------------------------
// Plain megakernel (one work-item per pixel)
// ---------------------------------------------------------------------------
template <bool EXT, int SHORT>
DEV V radiance_pt(const DevScene& sc, V o, V d, const Rng& rng, int max_bounce, int min_rr, bool preview,
                  TraversalStack<SHORT>& st, int64_t* cnt) {
    V L = mk(0, 0, 0), beta = mk(1, 1, 1);
    bool spec = false;
    float eta_scale = 1.0f;
    int pass = 0;
    int64_t n_closest = 0, n_any = 0;
    for (int bounce = 0;; bounce++) {
        HitRec h;
        n_closest++;
        if (trace<false, SHORT>(sc, o, d, T_MIN, FLT_MAX, h, st)) {
            Surf si = make_surface(sc, o, d, h);
            V wo = -d;
            if (bounce == 0 || spec) L = L + beta * emitted<EXT>(sc, si, wo);
            if (bounce >= max_bounce) break;
            const DevMaterial& m = sc.mats[sc.prims[h.prim].material];
            TexHit th;
            if (!shade_material<EXT>(sc, m, si, d, th)) {
                o = si.p;
                bounce--;
                if (++pass > 256) break;
                continue;
            }
            const uint32_t dim = DIM_BOUNCE + DIMS_PER_BOUNCE * (uint32_t)bounce;
            Bsdf b = make_bsdf<EXT>(sc, m, si, d, rng, dim, beta, th);
            if (sc.n_lights > 0 && count_nonspecular(b) > 0) {
                int n = sc.n_lights;
                int li = min((int)(rng.get(dim) * n), n - 1);
                ShadowReq r0, r1;
                nee_requests<EXT>(sc, b, si, li, rng, dim, beta * (float)n, r0, r1);
                VisitCount vc;
                if (r0.mode >= 0) {
                    if (r0.mode == 0) n_any++; else n_closest++;
                    L = L + resolve_shadow<SHORT>(sc, r0, st, vc);
                }
                if (r1.mode >= 0) {
                    if (r1.mode == 0) n_any++; else n_closest++;
                    L = L + resolve_shadow<SHORT>(sc, r1, st, vc);
                }
            }
            LSample bs;
            if (!bsdf_sample<EXT>(b, BX_ALL, rng.get(dim + 6), rng.get(dim + 7), bs)) break;
            if (black(bs.f) || bs.pdf == 0.0f) break;
            beta = beta * (bs.f * fabsf(dot(bs.wi, si.s.w)) / bs.pdf);
            spec = (bs.type & BX_SPEC) != 0;
            if ((bs.type & BX_TRANS) && spec) eta_scale *= dot(wo, si.g.w) > 0.0f ? b.eta * b.eta : 1.0f / (b.eta * b.eta);
            o = si.p;
            d = nrm(bs.wi);
            V rr = beta * eta_scale;
            if (maxc(rr) < 1.0f && bounce >= min_rr) {
                float q = fmaxf(0.05f, 1.0f - maxc(rr));
                if (rng.get(dim + 8) < q) break;
                beta = beta / (1.0f - q);
            }
        } else {
            // Scene.java:603-606: in preview mode a primary ray that leaves
            // the scene is white
            if (bounce == 0 && preview) L = mk(1.0f, 1.0f, 1.0f);
            else if (bounce == 0 || spec)
                for (int i = 0; i < sc.n_lights; i++) L = L + beta * emitted_ray<EXT>(sc, sc.lights[i], d);
            break;
        }
    }
    cnt[0] += n_closest;
    cnt[1] += n_any;
    return L;
}

template <int SHORT>
DEV V radiance_ao(const DevScene& sc, V o, V d, const Rng& rng, float max_dist, int samples, bool preview,
                  TraversalStack<SHORT>& st, int64_t* cnt) {
    HitRec h;
    cnt[0]++;
    if (!trace<false, SHORT>(sc, o, d, T_MIN, FLT_MAX, h, st)) {
        const float w = preview ? 1.0f : 0.0f;  // Scene.java:308-309
        return mk(w, w, w);
    }
    Surf si = make_surface(sc, o, d, h);
    float acc = 0.0f;
    for (int i = 0; i < samples; i++) {
        float u = rng.get(2 + 2 * i), v = rng.get(3 + 2 * i);
        float st_ = sqrtf(fmaxf(0.0f, 1.0f - u * u)), phi = 2.0f * PI_F * v;
        V hs = mk(st_ * cosf(phi), st_ * sinf(phi), u);
        V dir = nrm(mk(hs.x * si.g.u.x + hs.y * si.g.v.x + hs.z * si.g.w.x,
                       hs.x * si.g.u.y + hs.y * si.g.v.y + hs.z * si.g.w.y,
                       hs.x * si.g.u.z + hs.y * si.g.v.z + hs.z * si.g.w.z));
        HitRec h2;
        if (max_dist > 0.0f) {
            // hits beyond max_dist weigh the same as a miss, so the query may stop there
            cnt[0]++;
            if (trace<false, SHORT>(sc, si.p, dir, T_MIN, max_dist, h2, st)) {
                float t = fminf(fmaxf(h2.t, 0.0f), max_dist);
                acc += t / max_dist;
            } else {
                acc += 1.0f;
            }
        } else {
            cnt[1]++;
            if (!trace<true, SHORT>(sc, si.p, dir, T_MIN, FLT_MAX, h2, st)) acc += 1.0f;
        }
    }
    float v = acc * (PI_F / samples);
    v = v / PI_F;
    return mk(v, v, v);
}

template <int SHORT>
DEV V radiance_rc(const DevScene& sc, V o, V d, bool preview, TraversalStack<SHORT>& st, int64_t* cnt) {
    HitRec h;
    cnt[0]++;
    if (trace<false, SHORT>(sc, o, d, T_MIN, FLT_MAX, h, st)) {
        Surf si = make_surface(sc, o, d, h);
        float g = 0.5f * fabsf(dot(si.s.w, d));
        return mk(g, g, g);
    }
    if (preview) return mk(1.0f, 1.0f, 1.0f);  // Scene.java:692-693
    V L = mk(0, 0, 0);
    for (int i = 0; i < sc.n_lights; i++) L = L + emitted_ray<true>(sc, sc.lights[i], d);
    return L;
}

template <int BLOCK, int SHORT, bool EXT>
__global__ __launch_bounds__(BLOCK) void k_megakernel(const DevScene* __restrict__ scenes, const DevJob* __restrict__ jobp,
                                                      float4* __restrict__ film, int* __restrict__ overflow, int n_pixels,
                                                      unsigned long long* __restrict__ counters) {
    __shared__ int lds[SHORT * BLOCK];
    const DevJob& job = *jobp;
    const DevScene& sc = scenes[job.scene];
    int pix = blockIdx.x * BLOCK + threadIdx.x;
    int64_t cnt[2] = {0, 0};
    if (pix < n_pixels) {
        TraversalStack<SHORT> st;
        st.init(lds, overflow, BLOCK);
        int x = pix % job.width, y = pix / job.width;
        V sum = mk(0, 0, 0);
        int count = 0;
        for (int k = 0; k < job.spp; k++) {
            uint32_t sample = (uint32_t)(job.sample_base + k);
            Rng rng = make_rng(job.seed, (uint32_t)pix, sample);
            V o, d;
            if (!primary_ray(job.cam, (float)x, (float)y, rng.get(0), rng.get(1), rng, o, d)) continue;
            V L;
            const bool pv = job.preview != 0;
            if (job.algorithm == DFR_ALGO_PATH_TRACING) L = radiance_pt<EXT, SHORT>(sc, o, d, rng, job.max_bounce, job.min_rr, pv, st, cnt);
            else if (job.algorithm == DFR_ALGO_AMBIENT_OCCLUSION) L = radiance_ao<SHORT>(sc, o, d, rng, job.max_distance, job.ao_samples, pv, st, cnt);
            else if (job.algorithm == DFR_ALGO_RAY_CASTING) L = radiance_rc<SHORT>(sc, o, d, pv, st, cnt);
            else L = mk(0, 0, 0);
            // AbstractCPURenderer.java:229: drop NaN/Inf and negative-luminance samples
            float lum = 0.212671f * L.x + 0.715160f * L.y + 0.072169f * L.z;
            if (!(isfinite(L.x) && isfinite(L.y) && isfinite(L.z)) || lum < -1.0e-5f) continue;
            if (job.sample_film)
                film[job.film_offset + (int64_t)(sample - (uint32_t)job.film_sample_base) * n_pixels + pix] =
                    make_float4(L.x, L.y, L.z, 1.0f);
            sum = sum + L;
            count++;
        }
        if (!job.sample_film) {
            int64_t fi = job.film_offset + pix;
            float4 f = film[fi];
            f.x += sum.x; f.y += sum.y; f.z += sum.z; f.w += (float)count;
            film[fi] = f;
        }
    }
    // per-wave reduction of the ray counters, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        cnt[0] += __shfl_down(cnt[0], off);
        cnt[1] += __shfl_down(cnt[1], off);
    }
    if (lane_id() == 0 && counters) {
        atomicAdd(&counters[0], (unsigned long long)cnt[0]);
        atomicAdd(&counters[1], (unsigned long long)cnt[1]);
    }
}

int overflow_entries_per_thread(int max_depth) {
    int need = max_depth + 2 - SHORT_STACK;
    return need > 1 ? need : 1;
}

template <bool EXT>
static void launch_megakernel_t(const DevScene* scenes, const DevJob* job, float4* film, int* overflow, int n_pixels,
                                int block, unsigned long long* c, hipStream_t s) {
    int grid = (n_pixels + block - 1) / block;
    switch (block) {
        case 64: hipLaunchKernelGGL((k_megakernel<64, SHORT_STACK, EXT>), dim3(grid), dim3(64), 0, s, scenes, job, film, overflow, n_pixels, c); break;
        case 128: hipLaunchKernelGGL((k_megakernel<128, SHORT_STACK, EXT>), dim3(grid), dim3(128), 0, s, scenes, job, film, overflow, n_pixels, c); break;
        default: hipLaunchKernelGGL((k_megakernel<256, SHORT_STACK, EXT>), dim3((n_pixels + 255) / 256), dim3(256), 0, s, scenes, job, film, overflow, n_pixels, c); break;
    }
}
void launch_megakernel(const DevScene* scenes, const DevJob* job, float4* film, int* overflow, int n_pixels, int block,
                       bool ext_features, int64_t* counters, hipStream_t s) {
    auto c = reinterpret_cast<unsigned long long*>(counters);
    if (ext_features) launch_megakernel_t<true>(scenes, job, film, overflow, n_pixels, block, c, s);
    else launch_megakernel_t<false>(scenes, job, film, overflow, n_pixels, block, c, s);
}

// ---------------------------------------------------------------------------
// Ray queries
// ---------------------------------------------------------------------------
template <int BLOCK, int SHORT>
__global__ __launch_bounds__(BLOCK) void k_trace_rays(const DevScene* __restrict__ scene, const float4* __restrict__ rays,
                                                      int64_t n, int any_hit, int32_t* __restrict__ prim,
                                                      int64_t* __restrict__ tri, float* __restrict__ tout,
                                                      int* __restrict__ overflow) {
    __shared__ int lds[SHORT * BLOCK];
    TraversalStack<SHORT> st;
    st.init(lds, overflow, BLOCK);
    const DevScene& sc = *scene;
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
        float4 a = rays[2 * i], b = rays[2 * i + 1];
        HitRec h;
        bool hit = any_hit ? trace<true, SHORT>(sc, f4v(a), f4v(b), a.w, b.w, h, st)
                           : trace<false, SHORT, true>(sc, f4v(a), f4v(b), a.w, b.w, h, st);
        prim[i] = hit ? h.prim : -1;
        tri[i] = hit ? h.tri : -1;
        tout[i] = hit ? h.t : INFINITY;
    }
}

void launch_trace_rays(const DevScene* scene, const float4* rays, int64_t n, int any_hit, int32_t* prim, int64_t* tri,
                       float* t, int* overflow, int block, int grid, hipStream_t s) {
    (void)block;
    hipLaunchKernelGGL((k_trace_rays<256, SHORT_STACK>), dim3(grid), dim3(256), 0, s, scene, rays, n, any_hit, prim, tri, t,
                       overflow);
}

// ---------------------------------------------------------------------------
// Ray queries without allocation: the traversal stacks' HBM overflow lives in
// a file-scope device array (`scratch`) instead of a buffer the library
// allocates.  The launch has a fixed grid of NOALLOC_BLOCKS blocks of 256
// threads that stride over the rays, and every thread owns NOALLOC_STACK
// overflow entries (entry k of thread g at scratch[k * NOALLOC_THREADS + g]),
// so the array holds NOALLOC_THREADS * NOALLOC_STACK ints (24 MiB).  Every
// call on one device shares that device's array: calls that overlap on a
// device race.  Nothing in it needs clearing (entries are written before
// they are read).
// ---------------------------------------------------------------------------
constexpr int NOALLOC_BLOCKS = 1024;
constexpr int NOALLOC_THREADS = NOALLOC_BLOCKS * 256;
constexpr int NOALLOC_STACK = 24;
static __device__ __align__(256) int scratch[(size_t)NOALLOC_THREADS * NOALLOC_STACK];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
int* g_scratch_ptr[MAX_DEVICES] = {};
}  // namespace

int noalloc_scratch_init(int device) {
    if (device < 0 || device >= MAX_DEVICES) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return (int)hipSuccess;
    // the device is current: its own copy of the array
    int* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    g_scratch_ptr[device] = ptr;
    return (int)hipSuccess;
}

int* noalloc_scratch(int device) {
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_ptr[device];
}

int noalloc_stack_entries() { return NOALLOC_STACK; }

void launch_trace_rays_noalloc(const DevScene* scene, const float4* rays, int64_t n, int any_hit, int32_t* prim,
                               int64_t* tri, float* t, int* scratch_ptr, hipStream_t s) {
    hipLaunchKernelGGL((k_trace_rays<256, SHORT_STACK>), dim3(NOALLOC_BLOCKS), dim3(256), 0, s, scene, rays, n, any_hit,
                       prim, tri, t, scratch_ptr);
}

template <int BLOCK, int SHORT>
__global__ __launch_bounds__(BLOCK) void k_cast_primary(const DevScene* __restrict__ scene, const DevJob* __restrict__ jobp,
                                                        float* __restrict__ rays, int32_t* __restrict__ prim,
                                                        int64_t* __restrict__ tri, float* __restrict__ tout,
                                                        int* __restrict__ overflow, int n_pixels) {
    __shared__ int lds[SHORT * BLOCK];
    const DevJob& job = *jobp;
    int pix = blockIdx.x * BLOCK + threadIdx.x;
    if (pix >= n_pixels) return;
    TraversalStack<SHORT> st;
    st.init(lds, overflow, BLOCK);
    int x = pix % job.width, y = pix / job.width;
    Rng rng = make_rng(job.seed, (uint32_t)pix, (uint32_t)job.sample_base);
    V o, d;
    float* r = rays + 8 * (size_t)pix;
    if (!primary_ray(job.cam, (float)x, (float)y, rng.get(0), rng.get(1), rng, o, d)) {
        for (int k = 0; k < 8; k++) r[k] = 0.0f;
        prim[pix] = -1; tri[pix] = -1; tout[pix] = INFINITY;
        return;
    }
    r[0] = o.x; r[1] = o.y; r[2] = o.z; r[3] = T_MIN; r[4] = d.x; r[5] = d.y; r[6] = d.z; r[7] = FLT_MAX;
    HitRec h;
    bool hit = trace<false, SHORT, true>(*scene, o, d, T_MIN, FLT_MAX, h, st);
    prim[pix] = hit ? h.prim : -1;
    tri[pix] = hit ? h.tri : -1;
    tout[pix] = hit ? h.t : INFINITY;
}

void launch_cast_primary(const DevScene* scene, const DevJob* job, float* rays, int32_t* prim, int64_t* tri, float* t,
                         int* overflow, int n_pixels, hipStream_t s) {
    hipLaunchKernelGGL((k_cast_primary<256, SHORT_STACK>), dim3((n_pixels + 255) / 256), dim3(256), 0, s, scene, job, rays,
                       prim, tri, t, overflow, n_pixels);
}

// Tone map operators of AbstractImageKernel.java:1189-1324 at exposure 1,
// the only exposure GPURenderer.java:312-335 passes.  Ordinals as in
// AbstractGPURenderer.java:50-79; any other value leaves the colour as is.
__device__ __forceinline__ float tone_map(int op, float a) {
    switch (op) {
        case 0: {  // filmic curve, ACES modified: subtract 0, minimum -FLT_MAX
            const float b = fmaxf(a, -3.40282347e38f);
            return fminf(fmaxf((b * (2.51f * b + 0.03f)) / (b * (2.43f * b + 0.59f) + 0.14f), 0.0f), 1.0f);
        }
        case 2: return a / (1.0f + a);                          // Reinhard
        case 3: return a * (1.0f + a / 16.0f) / (1.0f + a);     // Reinhard, white point 4
        case 4: return 1.0f - expf(-a);                         // 1 - exp(-exposure^2 a)
        case 5: return a / (a + 0.155f) * 1.019f;               // Unreal 3
        default: return a;
    }
}

// imageBegin, tone map, imageRedoGammaCorrection, imageEnd
// (AbstractImageKernel.java:1104-1168, sRGB constants :39-47) over the
// (sum, count) film: one pixel per thread, 16 B in, 4 B out.
__global__ void k_rgba8(const float4* __restrict__ film, uint8_t* __restrict__ out, int n, int tone) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 f = film[i];
    float inv = f.w > 0.0f ? 1.0f / f.w : 0.0f;
    float c[3] = {f.x * inv, f.y * inv, f.z * inv};
    uint32_t px = 0xff000000u;
    for (int k = 0; k < 3; k++) {
        float v = tone_map(tone, c[k]);
        v = v <= 0.00304f ? v * 12.9230566f : 1.05500054f * powf(v, 0.41666666f) - 0.05500052f;
        px |= (uint32_t)(uint8_t)(int)(fminf(fmaxf(v, 0.0f), 1.0f) * 255.0f + 0.5f) << (8 * k);
    }
    reinterpret_cast<uint32_t*>(out)[i] = px;
}

void launch_rgba8(const float4* film, uint8_t* out, int n, int tone_mapper, hipStream_t s) {
    hipLaunchKernelGGL(k_rgba8, dim3((n + 255) / 256), dim3(256), 0, s, film, out, n, tone_mapper);
}

// ---------------------------------------------------------------------------
// CPURenderer's film: PixelImageF.filmAddColorXYZ with GaussianFilter2F(2, 2,
// 2) (M/image/PixelImageF.java:569-617) and filmRender (:691-705) in
// ColorSpaceF.S_R_G_B.  The reference splats each sample into every pixel
// whose centre lies within two pixels of it; here each output pixel gathers
// the samples of its 5 x 5 neighbourhood, in the order the reference adds
// them to that pixel (pass by pass, samples in row order), from per-sample
// records (rgb, 1 when the sample exists).  A sample's position is its
// pixel plus the camera offset drawn from its stream's dimensions 0 and 1.
// Products and sums are kept unfused so each term rounds as the
// reference's does.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_film_gauss(const float4* __restrict__ smp, int W, int H, int spp, uint64_t seed,
                                                    int sample_base, const FilmConsts* __restrict__ fc,
                                                    float4* __restrict__ xyz, float4* __restrict__ film) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    const int64_t npix = (int64_t)W * H;
    const int p = y * W + x;
    const float* m = fc->rgb_to_xyz;
    float4 acc = xyz[p];
    float4 own = film[p];
    for (int k = 0; k < spp; k++) {
        const uint32_t sample = (uint32_t)(sample_base + k);
        const float4* rec = smp + (int64_t)k * npix;
        for (int qy = max(y - 2, 0); qy <= min(y + 2, H - 1); qy++) {
            for (int qx = max(x - 2, 0); qx <= min(x + 2, W - 1); qx++) {
                const int q = qy * W + qx;
                const float4 r = rec[q];
                if (r.w == 0.0f) continue;
                const Rng rng = make_rng(seed, (uint32_t)q, sample);
                const float dx = __fadd_rn((float)qx, rng.get(0)) - 0.5f;
                const float dy = __fadd_rn((float)qy, rng.get(1)) - 0.5f;
                const int x0 = (int)fmaxf(ceilf(dx - 2.0f), 0.0f), x1 = (int)fminf(floorf(dx + 2.0f), (float)(W - 1));
                const int y0 = (int)fmaxf(ceilf(dy - 2.0f), 0.0f), y1 = (int)fminf(floorf(dy + 2.0f), (float)(H - 1));
                if (x < x0 || x > x1 || y < y0 || y > y1) continue;
                const int ox = min((int)floorf(fabsf(__fmul_rn(__fmul_rn((float)x - dx, 0.5f), 16.0f))), 15);
                const int oy = min((int)floorf(fabsf(__fmul_rn(__fmul_rn((float)y - dy, 0.5f), 16.0f))), 15);
                const float w = fc->table[oy * 16 + ox];
                const float X = __fadd_rn(__fadd_rn(__fmul_rn(m[0], r.x), __fmul_rn(m[3], r.y)), __fmul_rn(m[6], r.z));
                const float Y = __fadd_rn(__fadd_rn(__fmul_rn(m[1], r.x), __fmul_rn(m[4], r.y)), __fmul_rn(m[7], r.z));
                const float Z = __fadd_rn(__fadd_rn(__fmul_rn(m[2], r.x), __fmul_rn(m[5], r.y)), __fmul_rn(m[8], r.z));
                acc.x = __fadd_rn(acc.x, __fmul_rn(X, w));
                acc.y = __fadd_rn(acc.y, __fmul_rn(Y, w));
                acc.z = __fadd_rn(acc.z, __fmul_rn(Z, w));
                acc.w = __fadd_rn(acc.w, w);
                if (q == p) { own.x += r.x; own.y += r.y; own.z += r.z; own.w += 1.0f; }
            }
        }
    }
    xyz[p] = acc;
    film[p] = own;
}

void launch_film_gauss(const float4* samples, int width, int height, int spp, uint64_t seed, int sample_base,
                       const FilmConsts* consts, float4* xyz, float4* film, hipStream_t s) {
    if (spp <= 0 || width <= 0 || height <= 0) return;
    hipLaunchKernelGGL(k_film_gauss, dim3((width + 15) / 16, (height + 15) / 16), dim3(256), 0, s, samples, width, height,
                       spp, seed, sample_base, consts, xyz, film);
}

__global__ void k_film_rgb(const float4* __restrict__ xyz, float* __restrict__ rgb, int n, const FilmConsts* __restrict__ fc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 f = xyz[i];
    const float* m = fc->xyz_to_rgb;
    float c[3];
    for (int k = 0; k < 3; k++)
        c[k] = __fadd_rn(__fadd_rn(__fmul_rn(m[3 * k], f.x), __fmul_rn(m[3 * k + 1], f.y)), __fmul_rn(m[3 * k + 2], f.z));
    if (f.w != 0.0f) {
        const float r = 1.0f / f.w;
        for (int k = 0; k < 3; k++) c[k] = fmaxf(__fmul_rn(c[k], r), 0.0f) + 0.0f;
    }
    for (int k = 0; k < 3; k++) {
        const float v = c[k];
        rgb[3 * i + k] = v <= fc->break_point
                             ? __fmul_rn(v, fc->slope)
                             : __fmul_rn(fc->slope_match, (float)pow((double)v, (double)fc->gamma_recip)) - fc->segment_offset;
    }
}

void launch_film_rgb(const float4* xyz, float* rgb, int n, const FilmConsts* consts, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_film_rgb, dim3((n + 255) / 256), dim3(256), 0, s, xyz, rgb, n, consts);
}

// ---------------------------------------------------------------------------
// Wavefront path tracer
// ---------------------------------------------------------------------------
DEV int find_job(const DevJob* jobs, int n_jobs, int64_t item) {
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].work_begin <= item) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// New camera paths appended after the continuing ones (up to capacity).
// Type (0..2) of the light a path's NEE at `bounce` will sample, 3 without
// lights: the light choice is the stateless random number of that bounce.
// Paths carry it in bits 9-10 of info.w, so the sort key of a hit needs only
// the hit's material.
constexpr uint32_t INFO_LT_SHIFT = 9;
// bits of info.w that path records also keep in the origin's w
constexpr uint32_t ORIGIN_W_BITS = 0x80000000u | (3u << INFO_LT_SHIFT);
DEV uint32_t nee_light_type(const DevScene& sc, const Rng& rng, uint32_t bounce) {
    if (sc.n_lights == 0) return 3u;
    const int li = min((int)(rng.get(DIM_BOUNCE + DIMS_PER_BOUNCE * bounce) * sc.n_lights), sc.n_lights - 1);
    const int t = sc.lights[li].type;
    // infinite lights share class 2, delta lights class 1
    return t == DFR_LIGHT_PEREZ || t == DFR_LIGHT_IMAGE ? 2u : (t == DFR_LIGHT_DIFFUSE_AREA ? 0u : 1u);
}

// The job of work item `item` through a generation-order table: entries in
// increasing work_begin, each naming the job slot that owns the range from its
// work_begin to the next entry's.
DEV int find_job_ordered(const GenEntry* order, int n_order, int64_t item) {
    int lo = 0, hi = n_order - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (order[mid].work_begin <= item) lo = mid; else hi = mid - 1;
    }
    return order[lo].slot;
}

__global__ void k_generate(const DevScene* __restrict__ scenes, const DevJob* __restrict__ jobs, int n_jobs,
                           const GenEntry* __restrict__ order, int n_order,
                           PathSoA out, Counters* __restrict__ ctr, float4* __restrict__ film, int capacity) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int base = ctr->n_next;
    int64_t next = ctr->work_next, total = ctr->work_total;
    int64_t room = (int64_t)capacity - base;
    int64_t avail = total - next;
    int64_t take = room < avail ? room : avail;
    if (i >= take) return;
    int64_t item = next + i;
    int j = order ? find_job_ordered(order, n_order, item) : find_job(jobs, n_jobs, item);
    const DevJob& job = jobs[j];
    int64_t local = item - job.work_begin;
    int npix = job.width * job.height;
    int pix = (int)(local % npix);
    uint32_t sample = (uint32_t)(job.sample_base + local / npix);
    int x = pix % job.width, y = pix / job.width;
    Rng rng = make_rng(job.seed, (uint32_t)pix, sample);
    V o, d;
    int slot = base + i;
    uint32_t flags = 0;
    if (!primary_ray(job.cam, (float)x, (float)y, rng.get(0), rng.get(1), rng, o, d)) {
        flags = 0x80000000u;  // rejected camera sample: the path is dropped at shade
        o = mk(0, 0, 0); d = mk(0, 0, 1);
    } else {
        const int64_t fi = job.sample_film ? (int64_t)(sample - (uint32_t)job.film_sample_base) * npix + pix : pix;
        unsafeAtomicAdd(&film[job.film_offset + fi].w, 1.0f);
    }
    flags |= nee_light_type(scenes[job.scene], rng, 0u) << INFO_LT_SHIFT;
    out.o[slot] = make_float4(o.x, o.y, o.z, __uint_as_float(flags & ORIGIN_W_BITS));
    out.d[slot] = make_float4(d.x, d.y, d.z, 1.0f);
    out.beta[slot] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    out.info[slot] = make_uint4((uint32_t)pix, sample, (uint32_t)j, flags);
}

__global__ void k_commit(Counters* ctr, int capacity) {
    int base = ctr->n_next;
    int64_t room = (int64_t)capacity - base;
    int64_t avail = ctr->work_total - ctr->work_next;
    int64_t take = room < avail ? room : avail;
    if (take < 0) take = 0;
    ctr->work_next += take;
    ctr->n_active = base + (int)take;
    ctr->bounces += ctr->n_active;
    ctr->n_next = 0;
    ctr->n_shadow2 = 0;
}

// Sort key of a hit: material class x the type of the light its NEE will
// sample (0..2; 3 with no lights), so that a wave of the shade kernel runs
// one material and one light-sampling routine.  The class is the material
// type for matte .. Disney (0..5); glossy, substrate, translucent, clear
// coat and uber share class 6.  The light is chosen by
// the same stateless random number the shade kernel draws.  30 = miss,
// 31 = dropped camera sample.
constexpr int NSORT_KEYS = 32;
constexpr int SORT_TILE = 16384;

DEV int sort_key_in(const DevScene& sc, int prim, uint32_t info_w) {
    if (prim == -2) return 31;
    if (prim < 0) return 30;
    return sc.prim_class[prim] + (int)((info_w >> INFO_LT_SHIFT) & 3u);
}
DEV int sort_key(const DevScene* __restrict__ scenes, const DevJob* __restrict__ jobs, uint4 info, float4 h) {
    const int prim = __float_as_int(h.y);
    return prim < 0 ? sort_key_in(scenes[0], prim, info.w) : sort_key_in(scenes[jobs[info.z].scene], prim, info.w);
}

// Traversal of a ray queue with ray replacement: block b owns RPL * BLOCK
// consecutive queue entries; a lane whose ray is finished takes the next
// unclaimed entry of the block's range (one LDS atomic per wave), so the
// lanes of a wave keep working while its slowest ray finishes.  Traversal is
// the speculative while-while walk of trace().
//   SHADOW false: closest hit of every queued path, written to hits[i].
//   SHADOW true:  the any-hit part of the shadow queue (entries [0, n_any)):
//                 a request that hits nothing adds its contribution to the film.
// queue entries per thread of a traversal block: 8 for the extension rays,
// 4 for the shorter-lived shadow requests (whose queue is shorter too)
constexpr int EXTEND_RPL = 8;
constexpr int SHADOW_RPL = 4;

// At most 72 VGPRs (7 waves per SIMD): measured faster than 6 waves with no
// register limit and than 8 waves with spills.
// UNI: every job of the launch uses the scene `uscene`, whose node and
// triangle arrays' bases then stay in scalar registers; otherwise each ray
// reaches its arrays through its job's scene.
template <int BLOCK, int SHORT, int RPL, bool COUNT, bool SHADOW, bool UNI>
__global__ __launch_bounds__(BLOCK, 7) void k_trace_queue(const DevScene* __restrict__ scenes,
                                                         const DevJob* __restrict__ jobs, PathSoA q, ShadowSoA sh,
                                                         float4* __restrict__ hits, float4* __restrict__ film,
                                                         const Counters* __restrict__ ctr, int* __restrict__ overflow,
                                                         unsigned long long* __restrict__ trav,
                                                         const DevScene* __restrict__ uscene, int32_t* __restrict__ keys,
                                                         int32_t* __restrict__ bhist, int tiles, RaySample smp) {
    __shared__ int lds[SHORT * BLOCK];
    __shared__ int next;
    __shared__ int s_hist[NSORT_KEYS];
    // scene-uniform extension with sorted shading: the finished ray writes its
    // sort key (the material of the hit, the NEE light type from info.w)
    uint32_t info_w = 0;
    constexpr uint32_t DONE = TRAVERSAL_DONE;
    const int n = SHADOW ? (int)(ctr->n_shadow2 & 0xffffffffull) : ctr->n_active;
    const int base = blockIdx.x * RPL * BLOCK;
    const int limit = min(n - base, RPL * BLOCK);
    if (threadIdx.x == 0) next = BLOCK;
    if (UNI && !SHADOW && threadIdx.x < NSORT_KEYS) s_hist[threadIdx.x] = 0;
    __syncthreads();
    if (limit <= 0) return;
    const int lane = lane_id();
    const uint64_t lt = (1ull << lane) - 1ull;
    TraversalStack<SHORT> st;
    st.init(lds, overflow, BLOCK);
    VisitCount vc;
    int e = threadIdx.x;  // entry of the block's range this lane traces
    bool active = e < limit;
    PRay r;
    float tbest = FLT_MAX;
    int prim = -1, tri = -1;
    uint32_t cur = DONE, leaf = 0;
    // with COUNT: this ray's visits and the far limit it entered the BVH with
    int ray_nodes = 0, ray_tris = 0;
    float t_entry = 0.0f;
    bool entered = false;
    const float4* nodes = UNI ? uscene->nodes : nullptr;
    const float4* tris = UNI ? uscene->tris : nullptr;
    PlaneSel ps{0u, 0u, 0u};
    bool fresh = active;  // a ray to set up at the top of the loop
    for (;;) {
        if (fresh) {
            const int i = base + e;
            fresh = false;
            prim = -1; tri = -1; leaf = 0; st.sp = 0; cur = DONE;
            if (COUNT) { ray_nodes = 0; ray_tris = 0; entered = false; }
            // the ray, the item walk's result, the infinite planes, then the
            // triangles' BVH root (the scene-uniform kernel skips the job ->
            // scene hops)
            auto setup = [&](const float4 o4, const float4 d4, uint32_t job) {
                const DevScene& sc = UNI ? *uscene : scenes[jobs[job].scene];
                const V o = f4v(o4), d = f4v(d4);
                r = prep_ray(o, d);
                if (sc.prepass) {
                    // k_items_queue ran ahead: the nearest sphere, rectangle
                    // or instance triangle, or an occluded shadow ray's tmax
                    // of -inf
                    if (SHADOW) {
                        if (tbest == -INFINITY) prim = 0;
                    } else {
                        const float4 h = hits[i];
                        tbest = h.x;
                        prim = __float_as_int(h.y);
                        tri = __float_as_int(h.z);
                    }
                }
                for (int k = 0; k < sc.n_analytic; k++) {
                    const int pi = sc.analytic[k];
                    const float t = analytic_t(sc.prims[pi], o, d, T_MIN, tbest);
                    if (t != 0.0f && t < tbest) { tbest = t; prim = pi; tri = -1; }
                }
                if (!UNI) { nodes = sc.nodes; tris = sc.tris; }
                ps = plane_sel(r);
                cur = sc.has_bvh && !(SHADOW && prim >= 0) ? 0u : DONE;
                if (COUNT) { t_entry = tbest; entered = cur == 0u; }
            };
            if (SHADOW) {
                const float4 o4 = sh.o[i];
                tbest = o4.w;
                setup(o4, sh.d[i], UNI ? 0u : (uint32_t)sh.aux[i].y);
            } else if (UNI) {
                // the origin's w carries the dropped flag and the NEE light
                // type: no info load
                const float4 o4 = q.o[i], d4 = q.d[i];
                info_w = __float_as_uint(o4.w);
                tbest = FLT_MAX;
                if (info_w & 0x80000000u) prim = -2;  // dropped camera sample
                else setup(o4, d4, 0u);
            } else {
                // the three loads issued together (the ray's are wasted on a
                // dropped camera sample)
                const uint4 info = q.info[i];
                const float4 o4 = q.o[i], d4 = q.d[i];
                info_w = info.w;
                tbest = FLT_MAX;
                if (info.w & 0x80000000u) prim = -2;  // dropped camera sample
                else setup(o4, d4, info.z);
            }
        }
        if (active && (cur != DONE || leaf != 0)) {
            while (cur < DONE) {
                if (COUNT) { vc.nodes++; ray_nodes++; }
                cur = visit_node<SHORT, !SHADOW>(nodes, cur, r, ps, T_MIN, tbest, st);
                if ((cur & LEAF_FLAG_DEV) && leaf == 0) {
                    leaf = cur;
                    cur = st.sp == 0 ? DONE : (uint32_t)st.pop();
                }
                if (__popcll(__ballot(leaf == 0)) <= TRAVERSAL_STRAGGLERS) break;
            }
            while (leaf != 0) {
                // one triangle per round: the leaf reference steps to its next
                // triangle ((first + 1) << 3 | (count - 2) = reference + 7)
                const uint32_t first = (leaf & 0x7fffffffu) >> 3;
                if (COUNT) { vc.tris++; ray_tris++; }
                float4 a, b, c;
                load_tri(tris, first, a, b, c);
                const float t = tri_t(f4v(a), f4v(b), f4v(c), r.o, r.d, r.kx, r.ky, r.kz, r.sx, r.sy, r.sz, T_MIN, tbest);
                bool occluded = false;
                if (t != 0.0f && t < tbest) {
                    tbest = t; tri = __float_as_int(a.w); prim = __float_as_int(b.w);
                    occluded = SHADOW;
                }
                leaf = (leaf & 7u) ? leaf + 7u : 0u;
                if (SHADOW && occluded) { leaf = 0; cur = DONE; break; }
                if (leaf == 0 && (cur & LEAF_FLAG_DEV)) {
                    leaf = cur;
                    cur = st.sp == 0 ? DONE : (uint32_t)st.pop();
                }
                if (__popcll(__ballot(leaf != 0)) <= TRAVERSAL_STRAGGLERS) break;
            }
        }
        // finished rays: record the result, then claim the next entry of the range
        const bool finished = active && cur == DONE && leaf == 0;
        if (finished) {
            const int i = base + e;
            if (COUNT && smp.rec && entered && i % smp.stride == 0) {
                // every stride-th ray that entered the BVH, for the count check
                const unsigned long long k = atomicAdd(&trav[4], 1ull);
                if (k < (unsigned long long)smp.cap) {
                    smp.rec[3 * k] = make_float4(r.o.x, r.o.y, r.o.z, t_entry);
                    smp.rec[3 * k + 1] = make_float4(r.d.x, r.d.y, r.d.z, __int_as_float(SHADOW ? 1 : 0));
                    smp.rec[3 * k + 2] = make_float4(__int_as_float(ray_nodes), __int_as_float(ray_tris),
                                                     __int_as_float(prim), __int_as_float(tri));
                }
            }
            if (SHADOW) {
                if (prim < 0) {
                    const float4 c = sh.c[i];
                    film_add(film, ((int64_t)sh.aux[i].z << 32) | __float_as_uint(c.w), f4v(c));
                }
            } else {
                hits[i] = make_float4(tbest, __int_as_float(prim), __int_as_float(tri), 0.0f);
                if (UNI && keys) {
                    const int key = sort_key_in(*uscene, prim, info_w);
                    keys[i] = key;
                    atomicAdd(&s_hist[key], 1);
                }
            }
        }
        const uint64_t m = __ballot(finished);
        if (m) {
            const int leader = __ffsll((unsigned long long)m) - 1;
            int got = 0;
            if (lane == leader) got = atomicAdd(&next, __popcll(m));
            got = __shfl(got, leader);
            if (finished) {
                e = got + __popcll(m & lt);
                active = e < limit;
                fresh = active;
            }
        }
        if (__ballot(active) == 0) break;
    }
    if (UNI && !SHADOW && keys) {
        // the block's range lies in one sort tile: add its key counts
        __syncthreads();
        if (threadIdx.x < NSORT_KEYS && s_hist[threadIdx.x])
            atomicAdd(&bhist[threadIdx.x * tiles + base / SORT_TILE], s_hist[threadIdx.x]);
    }
    if (COUNT) add_visit_counts<BLOCK>(trav + (SHADOW ? 2 : 0), vc);
}

// Counting sort of the path queue by sort key, over tiles of SORT_TILE
// entries: per-tile key histograms (here with wave ballots, or added by the
// scene-uniform extension kernel), per-key scans of the key-major histogram
// array, then a scatter in which every wave ranks its entries per key with
// ballots and takes its base with one LDS atomic per key.

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sort_hist(const DevScene* __restrict__ scenes, const DevJob* __restrict__ jobs,
                                                     PathSoA q, const float4* __restrict__ hits,
                                                     int32_t* __restrict__ keys, int32_t* __restrict__ bhist,
                                                     const Counters* __restrict__ ctr) {
    __shared__ int hist[NSORT_KEYS];
    const int n = ctr->n_active;
    const int t0 = blockIdx.x * SORT_TILE;
    if (threadIdx.x < NSORT_KEYS) hist[threadIdx.x] = 0;
    __syncthreads();
    // per wave, one pass per distinct key present: the leader lane's key,
    // its lanes counted by one ballot; lane kk of the wave accumulates key kk
    int mine = 0;
    for (int k = threadIdx.x; k < SORT_TILE; k += BLOCK) {
        const int i = t0 + k;
        int key = -1;
        if (i < n) {
            key = sort_key(scenes, jobs, q.info[i], hits[i]);
            keys[i] = key;
        }
        uint64_t left = __ballot(key >= 0);
        while (left) {
            const int kk = __shfl(key, __ffsll((unsigned long long)left) - 1);
            const uint64_t m = __ballot(key == kk);
            if (lane_id() == kk) mine += __popcll(m);
            left &= ~m;
        }
    }
    if (lane_id() < NSORT_KEYS && mine) atomicAdd(&hist[lane_id()], mine);
    __syncthreads();
    if (threadIdx.x < NSORT_KEYS) bhist[threadIdx.x * gridDim.x + blockIdx.x] = hist[threadIdx.x];
}

// Per-key exclusive scan of the tile histograms, one block per key (tiles <=
// 4096, so at most 16 entries per thread); the key totals go to
// bhist[NSORT_KEYS * tiles + key] and k_sort_scatter adds their prefix.
__global__ __launch_bounds__(256) void k_sort_scan_rows(int32_t* __restrict__ bhist, int tiles) {
    __shared__ int wsum[4];
    const int key = blockIdx.x, t = threadIdx.x, lane = lane_id(), w = t >> 6;
    int32_t* row = bhist + key * tiles;
    const int per = (tiles + 255) / 256;
    const int b0 = min(tiles, t * per), b1 = min(tiles, b0 + per);
    int v[16];
    int sum = 0;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        v[e] = b0 + e < b1 ? row[b0 + e] : 0;
        sum += v[e];
    }
    int incl = sum;
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(incl, off);
        if (lane >= off) incl += u;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int pre = 0;
    for (int k = 0; k < w; k++) pre += wsum[k];
    int run = pre + incl - sum;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        if (b0 + e < b1) row[b0 + e] = run;
        run += v[e];
    }
    if (t == 255) bhist[NSORT_KEYS * tiles + key] = pre + incl;
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sort_scatter(const int32_t* __restrict__ keys, const int32_t* __restrict__ boff,
                                                        int32_t* __restrict__ perm, const Counters* __restrict__ ctr) {
    __shared__ int base[NSORT_KEYS];
    const int n = ctr->n_active;
    const int t0 = blockIdx.x * SORT_TILE;
    const int lane = lane_id();
    const uint64_t lt = (1ull << lane) - 1ull;
    if (threadIdx.x < NSORT_KEYS) {
        // exclusive prefix of the key totals (lanes 0..31 of wave 0)
        const int tot = boff[NSORT_KEYS * gridDim.x + threadIdx.x];
        int incl = tot;
        for (int off = 1; off < NSORT_KEYS; off <<= 1) {
            const int u = __shfl_up(incl, off);
            if ((int)threadIdx.x >= off) incl += u;
        }
        base[threadIdx.x] = boff[threadIdx.x * gridDim.x + blockIdx.x] + incl - tot;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < SORT_TILE; k += BLOCK) {
        const int i = t0 + k;
        const int key = i < n ? keys[i] : -1;
        int slot = 0;
        uint64_t left = __ballot(key >= 0);
        while (left) {
            const int leader = __ffsll((unsigned long long)left) - 1;
            const int kk = __shfl(key, leader);
            const uint64_t m = __ballot(key == kk);
            int b = 0;
            if (lane == leader) b = atomicAdd(&base[kk], __popcll(m));
            b = __shfl(b, leader);
            if (key == kk) slot = b + __popcll(m & lt);
            left &= ~m;
        }
        if (key >= 0) perm[slot] = i;
    }
}

DEV void put_shadow(ShadowSoA sh, int s, const ShadowReq& R, int64_t film_idx, int job) {
    sh.o[s] = make_float4(R.o.x, R.o.y, R.o.z, R.tmax);
    sh.d[s] = make_float4(R.d.x, R.d.y, R.d.z, __int_as_float(R.mode));
    sh.c[s] = make_float4(R.c.x, R.c.y, R.c.z, __uint_as_float((uint32_t)film_idx));
    sh.aux[s] = make_int4(R.light, job, (int)(film_idx >> 32), 0);
}

// At most 128 VGPRs (4 waves per SIMD): the few spills this costs are
// cheaper than running at 3 waves per SIMD.  Launched in SHADE_BLOCK-thread
// blocks: the block-wide queue reservation's two counter atomics are
// device-wide serial points, so fewer, larger blocks shade faster (512:
// -6 % against 256; 1024 loses more at the block barriers than it saves).
constexpr int SHADE_BLOCK = 512;
// UNI 1: every job of the launch uses `uscene`, read through the kernel
// argument so that the scene record's pointers stay in scalar registers
// (−3 % shade time against the per-job lookup, which puts them in VGPRs);
// UNI 2: moreover the launch has one job, whose record is then uniform too
// (−1.7 %).  UNI 0: per-job lookups.
template <int BLOCK, bool EXT, int UNI>
__global__ __launch_bounds__(BLOCK, 4) void k_shade(const DevScene* __restrict__ scenes, const DevScene* __restrict__ uscene,
                                                 const DevJob* __restrict__ jobs,
                                                 PathSoA q, const float4* __restrict__ hits,
                                                 const int32_t* __restrict__ perm, PathSoA out, ShadowSoA sh,
                                                 Counters* __restrict__ ctr, float4* __restrict__ film, int use_perm,
                                                 int* __restrict__ job_done, int capacity) {
    int k = blockIdx.x * BLOCK + threadIdx.x;
    int n = ctr->n_active;
    bool active = k < n;
    int i = active ? (use_perm ? perm[k] : k) : 0;
    bool cont = false;
    V no = mk(0, 0, 0), nd = mk(0, 0, 1), nbeta = mk(0, 0, 0);
    float neta = 1.0f;
    uint4 ninfo = make_uint4(0, 0, 0, 0);
    ShadowReq r0, r1;
    r0.mode = -1;
    r1.mode = -1;
    int64_t film_idx = 0;
    if (active) {
        uint4 info = q.info[i];
        float4 hv = hits[i];
        int prim = __float_as_int(hv.y);
        if (prim != -2) {
            const DevJob& job = UNI == 2 ? jobs[0] : jobs[info.z];
            const DevScene& sc = UNI ? *uscene : scenes[job.scene];
            film_idx = job.film_offset + info.x;
            if (job.sample_film)
                film_idx += (int64_t)(info.y - (uint32_t)job.film_sample_base) * (job.width * job.height);
            float4 o4 = q.o[i], d4 = q.d[i], b4 = q.beta[i];
            V o = f4v(o4), d = f4v(d4), beta = f4v(b4);
            float eta_scale = d4.w;
            int bounce = (int)(info.w & 0xffu);
            bool spec = (info.w >> 8) & 1u;
            int pass = (int)((info.w >> 16) & 0x1ffu);
            Rng rng = make_rng(job.seed, info.x, info.y);
            V L = mk(0, 0, 0);
            if (prim >= 0) {
                HitRec h; h.t = hv.x; h.prim = prim; h.tri = __float_as_int(hv.z);
                Surf si = make_surface(sc, o, d, h);
                V wo = -d;
                if (bounce == 0 || spec) L = L + beta * emitted<EXT>(sc, si, wo);
                if (bounce < job.max_bounce) {
                    const DevMaterial& m = sc.mats[sc.prims[prim].material];
                    TexHit th;
                    if (!shade_material<EXT>(sc, m, si, d, th)) {
                        if (pass + 1 <= 256) {
                            cont = true; no = si.p; nd = d; nbeta = beta; neta = eta_scale;
                            ninfo = make_uint4(info.x, info.y, info.z, (uint32_t)bounce | ((uint32_t)spec << 8) | ((uint32_t)(pass + 1) << 16) |
                                                                        (info.w & (3u << INFO_LT_SHIFT)));
                        }
                    } else {
                        const uint32_t dim = DIM_BOUNCE + DIMS_PER_BOUNCE * (uint32_t)bounce;
                        Bsdf b = make_bsdf<EXT>(sc, m, si, d, rng, dim, beta, th);
                        if (sc.n_lights > 0 && count_nonspecular(b) > 0) {
                            int nl = sc.n_lights;
                            int li = min((int)(rng.get(dim) * nl), nl - 1);
                            nee_requests<EXT>(sc, b, si, li, rng, dim, beta * (float)nl, r0, r1);
                        }
                        LSample bs;
                        if (bsdf_sample<EXT>(b, BX_ALL, rng.get(dim + 6), rng.get(dim + 7), bs) && !black(bs.f) && bs.pdf != 0.0f) {
                            V nb = beta * (bs.f * fabsf(dot(bs.wi, si.s.w)) / bs.pdf);
                            bool nspec = (bs.type & BX_SPEC) != 0;
                            float ne = eta_scale;
                            if ((bs.type & BX_TRANS) && nspec) ne *= dot(wo, si.g.w) > 0.0f ? b.eta * b.eta : 1.0f / (b.eta * b.eta);
                            V rr = nb * ne;
                            bool alive = true;
                            if (maxc(rr) < 1.0f && bounce >= job.min_rr) {
                                float qv = fmaxf(0.05f, 1.0f - maxc(rr));
                                if (rng.get(dim + 8) < qv) alive = false;
                                else nb = nb / (1.0f - qv);
                            }
                            if (alive) {
                                cont = true; no = si.p; nd = nrm(bs.wi); nbeta = nb; neta = ne;
                                ninfo = make_uint4(info.x, info.y, info.z, (uint32_t)(bounce + 1) | ((uint32_t)nspec << 8) | ((uint32_t)pass << 16) |
                                                                            (nee_light_type(sc, rng, (uint32_t)bounce + 1u) << INFO_LT_SHIFT));
                            }
                        }
                    }
                }
            } else if (bounce == 0 && job.preview) {
                L = mk(1.0f, 1.0f, 1.0f);  // Scene.java:603-606
            } else if (bounce == 0 || spec) {
                for (int l = 0; l < sc.n_lights; l++) L = L + beta * emitted_ray<EXT>(sc, sc.lights[l], d);
            }
            if (!black(L)) film_add(film, film_idx, L);
            // each NEE request carries the film index of its path
        }
    }
    if (job_done) wave_count_jobs(job_done, active && !cont, active ? q.info[i].z : 0u);
    // queue appends: one reservation per block
    const bool has0 = r0.mode >= 0, has1 = r1.mode >= 0;
    const bool any0 = has0 && r0.mode == 0, any1 = has1 && r1.mode == 0;
    const int nany = (any0 ? 1 : 0) + (any1 ? 1 : 0);
    const int ncl = (has0 && !any0 ? 1 : 0) + (has1 && !any1 ? 1 : 0);
    int s_path, s_any, s_cl;
    block_reserve<BLOCK>(ctr, cont, nany, ncl, s_path, s_any, s_cl);
    if (cont) {
        out.o[s_path] = make_float4(no.x, no.y, no.z, __uint_as_float(ninfo.w & ORIGIN_W_BITS));
        out.d[s_path] = make_float4(nd.x, nd.y, nd.z, neta);
        out.beta[s_path] = make_float4(nbeta.x, nbeta.y, nbeta.z, 0.0f);
        out.info[s_path] = ninfo;
    }
    const int last = 2 * capacity - 1;  // closest-hit requests fill the shadow queue from here down
    const int job = active ? (int)q.info[i].z : 0;
    if (has0) put_shadow(sh, any0 ? s_any : last - s_cl, r0, film_idx, job);
    if (has1) put_shadow(sh, any1 ? s_any + (any0 ? 1 : 0) : last - (s_cl + (has0 && !any0 ? 1 : 0)), r1, film_idx, job);
}

// Shadow queue, closest-hit partition (n_cl entries from the queue's end):
// area-light visibility (mode 1) and BSDF-sampled rays toward area lights
// (mode 2).
template <int BLOCK, int SHORT, bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_shadow_closest(const DevScene* __restrict__ scenes,
                                                          const DevJob* __restrict__ jobs, ShadowSoA sh,
                                                          const Counters* __restrict__ ctr, float4* __restrict__ film,
                                                          int* __restrict__ overflow,
                                                          unsigned long long* __restrict__ trav, int capacity) {
    __shared__ int lds[SHORT * BLOCK];
    const int j = blockIdx.x * BLOCK + threadIdx.x;
    const int n_cl = (int)(ctr->n_shadow2 >> 32);
    VisitCount vc;
    if (j < n_cl) {
        const int i = 2 * capacity - 1 - j;
        TraversalStack<SHORT> st;
        st.init(lds, overflow, BLOCK);
        const float4 o = sh.o[i], d = sh.d[i], c = sh.c[i];
        const int4 aux = sh.aux[i];
        const DevScene& sc = scenes[jobs[aux.y].scene];
        ShadowReq r;
        r.o = f4v(o); r.tmax = o.w; r.d = f4v(d); r.mode = __float_as_int(d.w); r.light = aux.x; r.c = f4v(c);
        HitRec h;
        const bool hit = trace<false, SHORT>(sc, r.o, r.d, T_MIN, r.tmax, h, st, vc);
        const V v = resolve_hit(sc, r, hit, h);
        if (!black(v)) film_add(film, ((int64_t)aux.z << 32) | __float_as_uint(c.w), v);
    }
    if (COUNT) add_visit_counts<BLOCK>(trav + 2, vc);
}

__global__ void k_count_rays(Counters* ctr) {
    const unsigned long long n2 = ctr->n_shadow2;
    ctr->rays_closest += ctr->n_active + (int64_t)(n2 >> 32);
    ctr->rays_any += (int64_t)(n2 & 0xffffffffull);
}

void wavefront_init(WavefrontBuffers& w, const DevJob* jobs, int n_jobs, int64_t work_total, hipStream_t s) {
    (void)jobs; (void)n_jobs;
    Counters c{};
    c.work_total = work_total;
    hipMemcpyAsync(w.ctr, &c, sizeof(Counters), hipMemcpyHostToDevice, s);
}

// Spheres, rectangles and instances of shared meshes ahead of a traversal
// launch over the path queue (SHADOW false) or the any-hit shadow queue
// (SHADOW true), one ray per thread: each ray of a scene with an item BVH or
// instances walks them, and the closest hit goes to hits[i] as (t,
// primitive, triangle) for k_trace_queue to start from; an occluded shadow
// ray's tmax becomes -inf.  Kept out of k_trace_queue, whose
// register budget (72 at 7 waves per SIMD) the walk's code would spill, and
// whose rays set up one at a time as lanes free up; here a whole wave walks
// together.  At most 72 VGPRs (7 waves per SIMD): 5.4 % faster on 5000
// spheres and rectangles, 2 % on 64 mesh instances, than 90 VGPRs at 5.
template <int BLOCK, int SHORT, bool SHADOW>
__global__ __launch_bounds__(BLOCK, 7) void k_items_queue(const DevScene* __restrict__ scenes,
                                                       const DevJob* __restrict__ jobs, PathSoA q, ShadowSoA sh,
                                                       float4* __restrict__ hits, const Counters* __restrict__ ctr,
                                                       int* __restrict__ overflow, const DevScene* __restrict__ uscene) {
    __shared__ int lds[SHORT * BLOCK];
    const int n = SHADOW ? (int)(ctr->n_shadow2 & 0xffffffffull) : ctr->n_active;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    float4 o4, d4;
    uint32_t job = 0;
    if (SHADOW) {
        o4 = sh.o[i];
        d4 = sh.d[i];
        if (!uscene) job = (uint32_t)sh.aux[i].y;
    } else {
        o4 = q.o[i];
        d4 = q.d[i];
        uint32_t flags;
        if (uscene) {
            flags = __float_as_uint(o4.w);
        } else {
            const uint4 info = q.info[i];
            flags = info.w;
            job = info.z;
        }
        if (flags & 0x80000000u) return;  // dropped camera sample
    }
    const DevScene& sc = uscene ? *uscene : scenes[jobs[job].scene];
    if (!sc.prepass) return;
    TraversalStack<SHORT> st;
    st.init(lds, overflow, BLOCK);
    const V o = f4v(o4), d = f4v(d4);
    const PRay r = prep_ray(o, d);
    const PlaneSel ps = plane_sel(r);
    float tbest = SHADOW ? o4.w : FLT_MAX;
    int prim = -1, tri = -1;
    bool hit = false;
    if (sc.n_items > 0) hit = trace_items<SHORT, SHADOW>(sc, r, ps, T_MIN, tbest, prim, st);
    if (sc.n_instances > 0 && !(SHADOW && hit)) hit = trace_instances<SHORT, SHADOW>(sc, o, d, T_MIN, tbest, prim, tri, st) || hit;
    if (SHADOW) {
        if (hit) sh.o[i] = make_float4(o4.x, o4.y, o4.z, -INFINITY);
    } else {
        hits[i] = make_float4(tbest, __int_as_float(prim), __int_as_float(tri), 0.0f);
    }
}

// One traversal launch over the path queue (SHADOW false) or the any-hit
// shadow queue (SHADOW true); the scene-uniform variant when p.uscene is set.
// With p.items, the item walk over the same queue runs first, in `igrid`
// blocks of B threads (one per ray).
template <int B, bool COUNT, bool SHADOW>
static void launch_trace_queue(int grid, int igrid, int tiles, const DevScene* scenes, const DevJob* jobs,
                               float4* film, WavefrontBuffers& w, const WavefrontParams& p, unsigned long long* trav,
                               hipStream_t s) {
    if (p.items)
        hipLaunchKernelGGL((k_items_queue<B, SHORT_STACK, SHADOW>), dim3(igrid), dim3(B), 0, s, scenes, jobs, w.a, w.sh,
                           w.hit, w.ctr, w.overflow, p.uscene);
    const RaySample smp = COUNT ? w.sample : RaySample{};
    if (BVH_WIDTH == 4 && p.uscene)
        hipLaunchKernelGGL((k_trace_queue<B, SHORT_STACK, SHADOW ? SHADOW_RPL : EXTEND_RPL, COUNT, SHADOW, true>), dim3(grid), dim3(B), 0, s,
                           scenes, jobs, w.a, w.sh, w.hit, film, w.ctr, w.overflow, trav, p.uscene,
                           SHADOW || !p.sort_keys ? nullptr : w.keys, w.bhist, tiles, smp);
    else
        hipLaunchKernelGGL((k_trace_queue<B, SHORT_STACK, SHADOW ? SHADOW_RPL : EXTEND_RPL, COUNT, SHADOW, false>), dim3(grid), dim3(B), 0, s,
                           scenes, jobs, w.a, w.sh, w.hit, film, w.ctr, w.overflow, trav, p.uscene,
                           SHADOW || !p.sort_keys ? nullptr : w.keys, w.bhist, tiles, smp);
}

__global__ void k_zero_i32(int32_t* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

template <int B>
static void iteration_b(const DevScene* scenes, const DevJob* jobs, int n_jobs, float4* film, WavefrontBuffers& w,
                        const WavefrontParams& p, const IterProfile* prof, hipStream_t s) {
    const int cap = w.capacity;
    // the queues' layout follows the capacity; the grids cover `qb` paths
    const int qb = p.queue_bound > 0 && p.queue_bound < cap ? p.queue_bound : cap;
    const int grid256 = p.generate ? (cap + 255) / 256 : 1;
    // top up the queue with new camera paths, then commit the counters
    hipLaunchKernelGGL(k_generate, dim3(grid256), dim3(256), 0, s, scenes, jobs, n_jobs, p.order, p.n_order, w.a, w.ctr,
                       film, cap);
    hipLaunchKernelGGL(k_commit, dim3(1), dim3(1), 0, s, w.ctr, cap);
    const bool timing = prof && prof->timing;
    const int use_perm = p.sort_keys > 0;
    const int tiles = (qb + SORT_TILE - 1) / SORT_TILE;  // <= 4096 (k_sort_scan_rows)
    // the scene-uniform extension kernel writes the sort keys and adds their
    // per-tile counts: clear the counts first
    const bool fused_hist = use_perm && BVH_WIDTH == 4 && p.uscene;
    if (fused_hist)
        hipLaunchKernelGGL(k_zero_i32, dim3((NSORT_KEYS * tiles + 255) / 256), dim3(256), 0, s, w.bhist, NSORT_KEYS * tiles);
    if (timing) hipEventRecord(prof->ev[0], s);
    unsigned long long* trav = prof ? prof->trav : nullptr;
    const int egrid = (qb + EXTEND_RPL * B - 1) / (EXTEND_RPL * B);
    const int eigrid = (qb + B - 1) / B;
    if (trav) launch_trace_queue<B, true, false>(egrid, eigrid, tiles, scenes, jobs, film, w, p, trav, s);
    else launch_trace_queue<B, false, false>(egrid, eigrid, tiles, scenes, jobs, film, w, p, trav, s);
    if (timing) hipEventRecord(prof->ev[1], s);
    if (use_perm) {
        if (!fused_hist)
            hipLaunchKernelGGL((k_sort_hist<256>), dim3(tiles), dim3(256), 0, s, scenes, jobs, w.a, w.hit, w.keys, w.bhist,
                               w.ctr);
        hipLaunchKernelGGL(k_sort_scan_rows, dim3(NSORT_KEYS), dim3(256), 0, s, w.bhist, tiles);
        hipLaunchKernelGGL((k_sort_scatter<256>), dim3(tiles), dim3(256), 0, s, w.keys, w.bhist, w.perm, w.ctr);
    }
    if (timing) hipEventRecord(prof->ev[2], s);
    // shading in 512-thread blocks whatever the traversal block: one queue
    // reservation (two same-address atomics) per 512 paths
    {
        const dim3 grid((qb + SHADE_BLOCK - 1) / SHADE_BLOCK), block(SHADE_BLOCK);
        const DevScene* us = p.uscene;
        const int uni = us ? (n_jobs == 1 ? 2 : 1) : 0;
#define DFR_SHADE(E, U)                                                                                                  \
    hipLaunchKernelGGL((k_shade<SHADE_BLOCK, E, U>), grid, block, 0, s, scenes, us, jobs, w.a, w.hit, w.perm, w.b, w.sh, \
                       w.ctr, film, use_perm, w.job_done, cap)
        if (p.ext_features) {
            if (uni == 2) DFR_SHADE(true, 2); else if (uni == 1) DFR_SHADE(true, 1); else DFR_SHADE(true, 0);
        } else {
            if (uni == 2) DFR_SHADE(false, 2); else if (uni == 1) DFR_SHADE(false, 1); else DFR_SHADE(false, 0);
        }
#undef DFR_SHADE
    }
    if (timing) hipEventRecord(prof->ev[3], s);
    hipLaunchKernelGGL(k_count_rays, dim3(1), dim3(1), 0, s, w.ctr);
    if (timing) hipEventRecord(prof->ev[4], s);
    const int sgrid = (int)((2 * (int64_t)qb + B - 1) / B);  // shadow requests <= 2 per path
    // the any-hit queue runs in blocks of at most 128 threads (−0.8 % against
    // 256; the extension rays prefer 256); the thread count, and so the
    // overflow stack, is the same
    constexpr int SB = B < 128 ? B : 128;
    const int agrid = (int)((2 * (int64_t)qb + SHADOW_RPL * SB - 1) / (SHADOW_RPL * SB));
    const int aigrid = (int)((2 * (int64_t)qb + SB - 1) / SB);
    if (trav) {
        launch_trace_queue<SB, true, true>(agrid, aigrid, tiles, scenes, jobs, film, w, p, trav, s);
        if (p.closest_shadow)
            hipLaunchKernelGGL((k_shadow_closest<B, SHORT_STACK, true>), dim3(sgrid), dim3(B), 0, s, scenes, jobs, w.sh,
                               w.ctr, film, w.overflow, trav, cap);
    } else {
        launch_trace_queue<SB, false, true>(agrid, aigrid, tiles, scenes, jobs, film, w, p, trav, s);
        if (p.closest_shadow)
            hipLaunchKernelGGL((k_shadow_closest<B, SHORT_STACK, false>), dim3(sgrid), dim3(B), 0, s, scenes, jobs, w.sh,
                               w.ctr, film, w.overflow, trav, cap);
    }
    if (timing) hipEventRecord(prof->ev[5], s);
    std::swap(w.a, w.b);
}

// One wavefront iteration; the queue state lives on the device, so a
// sequence of iterations needs no host round trip.  Block sizes 128 and 256
// are built (the autotuner picks one).
void wavefront_iteration(const DevScene* scenes, const DevJob* jobs, int n_jobs, float4* film, WavefrontBuffers& w,
                         const WavefrontParams& p, const IterProfile* prof, hipStream_t s) {
    if (p.block == 128) iteration_b<128>(scenes, jobs, n_jobs, film, w, p, prof, s);
    else iteration_b<256>(scenes, jobs, n_jobs, film, w, p, prof, s);
}

}  // namespace dfr
