// kernels.h — host-side launch interface of the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "device_types.h"

namespace dfr {

struct DevScene;

// One render job as the kernels see it.
struct DevJob {
    uint64_t seed;
    int32_t width, height;
    int32_t spp;              // samples per pixel of this launch sequence
    int32_t sample_base;      // first sample index (progressive passes)
    int32_t max_bounce, min_rr;
    int32_t algorithm, ao_samples;
    float max_distance;
    int32_t preview;          // setPreviewMode: a primary ray that misses is white
    int64_t film_offset;      // first pixel of this job in the film array
    int64_t work_begin;       // first work item of this job (prefix sum)
    int64_t work_count;       // width * height * spp
    float cam[19];
    int32_t scene;            // index into the scene table
    // 1: the film holds one record per sample, sample k of pixel p at
    // film_offset + (k - film_sample_base) * width * height + p (the
    // filtered-film mode); 0: one running sum per pixel
    int32_t sample_film;
    int32_t film_sample_base;
};

// Wavefront path state (structure of arrays).
struct PathSoA {
    float4* o;      // origin.xyz
    float4* d;      // direction.xyz, w = etaScale
    float4* beta;   // throughput rgb
    uint4* info;    // pixel, sample, job, bounce | specular << 8 | passthrough << 16
};

struct ShadowSoA {
    float4* o;      // origin.xyz, tmax
    float4* d;      // direction.xyz, mode (int bits)
    float4* c;      // contribution rgb, film index low 32 bits
    int4* aux;      // light index, job, film index high 32 bits, -
};

struct Counters {
    int32_t n_active;      // paths in the current queue
    int32_t n_next;        // continuing paths appended by shade
    int32_t taken;         // new paths appended by generate
    int32_t pad;
    // shadow requests appended by shade: any-hit requests fill the shadow
    // queue from the front (count in the low 32 bits), closest-hit requests
    // from the back (count in the high 32 bits); one 64-bit atomic reserves
    // both.
    uint64_t n_shadow2;
    int64_t work_next;     // next work item
    int64_t work_total;
    int64_t rays_closest;
    int64_t rays_any;
    int64_t bounces;
};
static_assert(sizeof(Counters) == 64, "Counters is one 64-byte line");

// Threads of the widest traversal launch (the shadow stage over a pool of
// `capacity` paths: 2 * capacity requests plus one wave of padding between
// the any-hit and closest-hit partitions, rounded up to a 256-thread block);
// the HBM overflow stacks are sized for it.
inline int64_t wavefront_stack_threads(int capacity) { return 2 * (int64_t)capacity + 64 + 256; }

// Largest path pool: the material sort's one-block scan holds 128 histogram
// entries per thread (32 keys x capacity / 16384-entry tiles <= 1024 x 128).
constexpr int MAX_POOL_PATHS = 1 << 26;

// With traversal counting: every `stride`-th traced ray that enters the BVH
// (extension and any-hit shadow queues) is recorded as three float4 --
// (origin, far limit at BVH entry), (direction, 1 for any hit else 0 as int
// bits), (node visits, triangle tests, hit primitive, hit triangle as int
// bits) -- in up to `cap` records; trav[4] counts the records claimed.
struct RaySample {
    float4* rec = nullptr;
    int cap = 0;
    int stride = 1;
};

struct WavefrontBuffers {
    PathSoA a, b;
    float4* hit;       // t, prim, tri (int bits)
    int32_t* perm;     // material-sorted order
    int32_t* keys;
    int32_t* bhist;    // per-tile key histograms / offsets (32 keys x tiles of 16384 entries), then 32 key totals
    ShadowSoA sh;
    int* overflow;     // traversal-stack overflow
    int capacity;      // path pool size
    Counters* ctr;
    int* job_done;     // optional: paths finished per job (serving)
    RaySample sample;  // optional: traced-ray sample of counting runs
};

// Launch helpers (kernels.hip).  `scenes` points to a device array of
// DevScene records (one per resident scene); jobs to a device array.
void launch_megakernel(const DevScene* scenes, const DevJob* job, float4* film, int* overflow, int n_pixels,
                       int block, bool ext_features, int64_t* counters, hipStream_t s);
void launch_trace_rays(const DevScene* scene, const float4* rays, int64_t n, int any_hit, int32_t* prim,
                       int64_t* tri, float* t, int* overflow, int block, int grid, hipStream_t s);
void launch_cast_primary(const DevScene* scene, const DevJob* job, float* rays, int32_t* prim, int64_t* tri,
                         float* t, int* overflow, int n_pixels, hipStream_t s);
void launch_rgba8(const float4* film, uint8_t* out, int n, int tone_mapper, hipStream_t s);

// CPURenderer's film (PixelImageF with GaussianFilter2F(2, 2, 2), sRGB).
// FilmConsts carries the 16 x 16 filter table, the colour-space arrays
// (ColorSpaceF storage order, 12 floats each) and the gamma constants.
struct FilmConsts {
    float table[256];
    float rgb_to_xyz[12], xyz_to_rgb[12];
    float break_point, slope, slope_match, segment_offset, gamma_recip;
};
// Splats the per-sample records of `spp` samples (sample_base ..) of a
// width x height job into the XYZ film (X, Y, Z, weight sum per pixel), in
// the reference's order per pixel, and adds each pixel's own samples to the
// (sum, count) film.
void launch_film_gauss(const float4* samples, int width, int height, int spp, uint64_t seed, int sample_base,
                       const FilmConsts* consts, float4* xyz, float4* film, hipStream_t s);
// filmRender(): XYZ film -> gamma-encoded RGB (3 floats per pixel).
void launch_film_rgb(const float4* xyz, float* rgb, int n, const FilmConsts* consts, hipStream_t s);

// Ray queries with the traversal overflow in the library's static device
// array (no allocation).  noalloc_scratch_init takes the array's address on
// `device`, which must be current, and returns a hipError_t code;
// noalloc_scratch gives that address (null before the init succeeded);
// noalloc_stack_entries is the number of overflow entries each thread owns.
int noalloc_scratch_init(int device);
int* noalloc_scratch(int device);
int noalloc_stack_entries();
void launch_trace_rays_noalloc(const DevScene* scene, const float4* rays, int64_t n, int any_hit, int32_t* prim,
                               int64_t* tri, float* t, int* scratch_ptr, hipStream_t s);

// One entry of a generation-order table (the serving loop, whose job slots
// are reused and so not ordered by their work ranges): the job in `slot` owns
// the work items from work_begin up to the next entry's work_begin.
struct GenEntry {
    int64_t work_begin;
    int32_t slot;
    int32_t pad;
};

struct WavefrontParams {
    int block;          // threads per block
    int sort_keys;      // 0: unsorted; 1: shade in (material, light type) order
    int grid_cap;       // blocks of the queue kernels (capacity / block)
    bool closest_shadow;  // some scene has area lights: launch the closest-hit shadow kernel
    // the scene record every job of the run uses (node and triangle arrays
    // below 4 GiB), or null: selects the scene-uniform traversal kernels
    const DevScene* uscene = nullptr;
    // some scene uses a glossy, substrate, translucent, clear coat or uber
    // material or a spot or directional light: launch the shade kernel
    // built with their code
    bool ext_features = false;
    // some scene has a BVH over its spheres and rectangles or instances of
    // shared meshes: each traversal launch is preceded by their walk over
    // its queue (k_items_queue)
    bool items = false;
    // generation-order table, or null when the job table itself is ordered
    // by work_begin (batches)
    const GenEntry* order = nullptr;
    int n_order = 0;
    // Bound on the paths of every queue of the next iterations, from the
    // host's last counter poll (0: the capacity).  Once every work item has
    // been generated the queues only shrink, so the grids of those
    // iterations are sized to it and `generate` is off.
    int queue_bound = 0;
    bool generate = true;
};

// Optional instrumentation of one wavefront iteration: HIP events around the
// extend, shade and shadow launches, and BVH visit counters for extend.
struct IterProfile {
    hipEvent_t ev[6];
    bool timing;
    unsigned long long* trav;  // [extend nodes, extend triangles, shadow nodes, shadow triangles, ray samples] or null
};

void wavefront_init(WavefrontBuffers& w, const DevJob* jobs, int n_jobs, int64_t work_total, hipStream_t s);
void wavefront_iteration(const DevScene* scenes, const DevJob* jobs, int n_jobs, float4* film, WavefrontBuffers& w,
                         const WavefrontParams& p, const IterProfile* prof, hipStream_t s);
int overflow_entries_per_thread(int max_depth);
constexpr int SHORT_STACK = 16;

}  // namespace dfr
