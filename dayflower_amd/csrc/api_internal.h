// api_internal.h — host-side state behind the C-ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dayflower_hip.h"
#include "bvh.h"
#include "device.h"
#include "kernels.h"

namespace dfr {

// Owning device allocation.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void release() {
        if (ptr) hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    // (Re)allocates when the current block is smaller than `n`; returns
    // hipSuccess or the allocation error.
    hipError_t alloc(size_t n) {
        if (n <= bytes && ptr) return hipSuccess;
        release();
        hipError_t e = hipMalloc(&ptr, n);
        if (e != hipSuccess) { ptr = nullptr; return e; }
        bytes = n;
        return hipSuccess;
    }
};

// Host copy of a dfr_scene_desc (setScene copies; setup compiles).
struct SceneHost {
    std::vector<float> positions, normals;
    std::vector<int32_t> indices;
    std::vector<dfr_primitive> prims;
    std::vector<dfr_material> mats;
    std::vector<dfr_light> lights;
    float camera[DFR_CAMERA_LENGTH];
    struct Image { int32_t width = 0, height = 0; std::vector<float> rgb; };
    std::vector<Image> images;
    std::vector<float> uvs;            // 2 per vertex, or empty
    std::vector<dfr_texture> textures;
};

// A validated object-to-world map of one primitive: the linear part `a`,
// the translation `t`, the normal matrix `n` (inverse transpose of `a`), its
// determinant `det`, the mean column length `scale`; `linear_identity` when
// a is the identity (a translation at most), `similarity` when a is a
// scaled rotation or reflection.
struct Xform {
    double a[3][3], n[3][3], t[3];
    double det = 1.0, scale = 1.0;
    bool identity = true, linear_identity = true, similarity = true;
};

// Mesh primitives that name the same triangle range are instances of one
// mesh: its triangles stay in object space, in a four-wide BVH of their own
// (one per shared range), and a top-level BVH over the instances' world
// boxes leads rays to them.  Built by instances.cpp.
struct InstanceSet {
    struct Group {
        int64_t first = 0, count = 0;  // the shared triangle range
        uint32_t root = 0;             // its tree's root within `bottom`
        float box[6];                  // object-space bounds: lo.xyz, hi.xyz
    };
    std::vector<int32_t> prim_group;   // per primitive: its group, or -1
    std::vector<int32_t> tri_group;    // per triangle: its group, or -1
    std::vector<Group> groups;
    std::vector<Node4> top;            // top-level tree over instances (root 0)
    std::vector<Node4> bottom;         // the groups' trees, references relative to `bottom`
    std::vector<float4> tris;          // the groups' triangles, object space, tree order
    std::vector<int32_t> order;        // top-level position -> primitive
    int max_stack = 0;                 // deepest stack of a top-level plus a group walk
    bool any() const { return !groups.empty(); }
};

// Finds the shared triangle ranges of `in` (identical ranges of two or more
// mesh primitives; ranges that overlap otherwise are an error, as is an area
// light on a shared range) into is.prim_group, tri_group and groups.
int find_instances(const SceneHost& in, InstanceSet& is, std::string& err);
// Builds every group's tree from `opos` (object-space vertices, nine floats
// per triangle) and the top-level tree over the instances' world boxes.
void build_instances(const std::vector<float>& opos, const std::vector<Xform>& xf, InstanceSet& is);
// Refits the top-level tree to new transforms (the groups' trees are object
// space and stay as they are).
void refit_instances(const std::vector<Xform>& xf, InstanceSet& is);
// The device records of the instances in top-level order.
std::vector<DevInstance> instance_records(const std::vector<Xform>& xf, const InstanceSet& is);
// The top-level tree followed by the groups' trees, with the latter's
// references moved past the former: the node array the kernels walk.
std::vector<Node4> instance_nodes(const InstanceSet& is);

// A scene compiled into SoA device buffers, resident in HBM.
struct CompiledScene {
    DeviceBuffer nodes, tris, tshade, prims, prim_class, analytic, mats, lobes, lights, perez, xforms, images, texels,
        textures, item_nodes, item_prims, inst_nodes, inst_tris, instances;
    DevScene dev{};
    int max_depth = 1;
    bool area_lights = false;  // any DiffuseAreaLight (the only lights whose NEE needs closest hits)
    bool ext_features = false;  // any glossy, substrate, translucent, clear coat or uber material, or spot or directional light
    int64_t n_nodes = 0, n_tris = 0;
    int64_t n_bvh_tris = 0;            // triangles in the scene's BVH (those of unshared meshes)
    double sah_cost = 0.0, build_seconds = 0.0;
    size_t bytes = 0;
    float camera[DFR_CAMERA_LENGTH];
    std::vector<Node4> nodes4_host;    // the BVH, kept for refits after transform updates
    std::vector<uint32_t> tri_order;   // BVH position -> triangle
    // the spheres' and rectangles' BVH, kept for refits
    std::vector<Node4> item_nodes_host;
    std::vector<int32_t> item_prim_order;  // item BVH position -> primitive
    int64_t n_items = 0;
    InstanceSet inst;                  // shared meshes, kept for refits
};

int copy_scene_desc(const dfr_scene_desc* d, SceneHost& h, std::string& err);
// The reference's CompiledScene arrays as a scene description and one
// object-to-world matrix per primitive (compiled_scene.cpp); *holder owns
// both until free_transcoded.
int transcode_compiled_scene(const dfr_compiled_scene* s, const dfr_scene_desc** desc, const float** xforms,
                             void** holder, std::string& err);
void free_transcoded(void* holder);
struct WavefrontBuffers;
// Allocates (or grows) the 17 wavefront buffers of a path pool of `capacity`
// paths into `mem` and points `wf` at them; false when an allocation fails.
bool alloc_wavefront_buffers(WavefrontBuffers& wf, DeviceBuffer* mem, int capacity);
// Bytes of per-sample records one chunk of a filtered render may use.
constexpr int64_t SAMPLE_FILM_BYTES = int64_t(2) << 30;

int make_xforms(const SceneHost& in, const std::vector<float>& m, std::vector<Xform>& out, std::string& err);
int compile_scene(const SceneHost& in, const std::vector<Xform>& xf, CompiledScene& out, int light_pdf_mode,
                  hipStream_t s, std::string& err);
int refit_scene(const SceneHost& in, const std::vector<Xform>& xf, CompiledScene& out, hipStream_t s, std::string& err);
size_t wavefront_bytes(int capacity);

struct WavefrontConfig {
    int block = 256;
    int capacity = 1 << 26;  // path pool size (64M paths, about 19 GB of HBM)
    int sort_keys = 1;       // 1: shade in (material, light type) order; 0: queue order
};

struct Ctx {
    std::mutex mutex;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string error;
    std::atomic<bool> stop{false};
    int scratch_status = 0;    // hipError_t of noalloc_scratch_init on this device

    // GPURenderer settings (AbstractGPURenderer.java:143-159)
    int width = 800, height = 800;
    int max_bounce = 20, min_rr = 5;
    int algorithm = DFR_ALGO_PATH_TRACING;
    int ao_samples = 10;
    float max_distance = 20.0f;
    int tone_mapper = DFR_TONE_MAPPER_NONE, preview = 0;
    uint64_t seed = 0;
    int kernel = 0;            // 0 wavefront, 1 plain megakernel
    int light_pdf_mode = 0;
    int spp_per_pass = 1;
    int profile = 0;           // time kernel launches with HIP events
    int count_traversal = 0;   // count BVH visits of the extend kernel
    int ray_sample_stride = 0; // with count_traversal: record every n-th traced ray (0: none)
    DeviceBuffer sample_buf;   // the recorded rays (RaySample records)
    int64_t sample_count = 0;  // records of the last counting run
    float camera[DFR_CAMERA_LENGTH] = {0};

    // scene state
    SceneHost scene_host;
    std::vector<float> xform_host;   // 16 floats per primitive of scene_host (object to world), or empty
    bool has_scene = false, scene_dirty = false, film_dirty = true, is_setup = false, clearing = false;
    int render_pass = 0;
    std::vector<std::unique_ptr<CompiledScene>> scenes;
    std::vector<int> scene_refs;  // per scene: requests of serving loops that render it
    uint64_t scene_gen = 0;       // bumped whenever the device scene table changes
    int main_scene = -1;
    DeviceBuffer scene_table, job_buf, film, overflow, counters, image;

    // wavefront engine
    WavefrontConfig wf_params;
    WavefrontBuffers wf{};
    DeviceBuffer wf_mem[17];
    int wf_capacity = 0;
    std::vector<hipEvent_t> prof_events;   // 6 per wavefront iteration, reused
    // Captured runs of WAVEFRONT_POLL iterations, replayed between counter
    // polls; each key holds every pointer, size and grid bound its launches
    // bake in.
    bool use_graphs = true;
    struct GraphEntry {
        std::vector<uintptr_t> key;
        hipGraphExec_t exec;
        uint64_t used;
    };
    static constexpr size_t MAX_GRAPHS = 8;  // least recently used evicted
    std::vector<GraphEntry> graphs;
    uint64_t graph_clock = 0;
    int replay_iterations(const std::vector<uintptr_t>& key, const DevScene* scenes, const DevJob* jobs, int n_jobs,
                          float4* film, const WavefrontParams& p, int iterations);

    // CPURenderer's Gaussian XYZ film (DFR_KEY_FILM_FILTER)
    int film_filter = 0;
    DeviceBuffer sample_records;   // per-sample records of one chunk of a filtered render
    DeviceBuffer xyz_film;         // X, Y, Z, filter weight sum per pixel
    DeviceBuffer film_consts;      // FilmConsts
    DeviceBuffer film_rgb;         // filmRender output staging
    int64_t xyz_pixels = -1;       // pixels xyz_film holds, -1 when stale
    int ensure_filtered_film();

    dfr_stats stats{};
    dfr_stats prof_stats{};   // instrumentation of the last wavefront run

    int fail(int code, const std::string& msg);
    int hip_fail(hipError_t e, const char* what);
    int sync_scene_table();
    int max_depth() const;
    bool any_area_lights() const;
    bool any_ext_features() const;
    bool any_items() const;
    int ensure_overflow(int64_t threads);
    int ensure_film(int64_t pixels);
    int ensure_wavefront(int capacity);
    int upload_jobs(const std::vector<DevJob>& jobs);
    int jobs_scene = -1;  // the scene every job of the last upload_jobs uses, -1 if they differ
    const DevScene* uniform_scene() const;
    int run_wavefront(int n_jobs, int64_t work_total, float4* film_ptr, Counters* out);
    DevJob main_job(int spp, int sample_base) const;
};

}  // namespace dfr
