"""ctypes mirror of include/dayflower_hip.h (records and entry points).

The record layouts are checked against the C header's sizes by the CPU test
suite; any change to the header must be mirrored here.
"""
import ctypes
import os

c_int32, c_int64, c_float, c_double, c_uint64 = (
    ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_double, ctypes.c_uint64)

CAMERA_LENGTH = 19

ALGO_AMBIENT_OCCLUSION, ALGO_DEPTH_CAMERA, ALGO_PATH_TRACING, ALGO_RAY_CASTING, ALGO_RAY_TRACING = range(5)
(TONE_MAPPER_FILMIC_CURVE_ACES_MODIFIED_VERSION_1, TONE_MAPPER_NONE, TONE_MAPPER_REINHARD,
 TONE_MAPPER_REINHARD_MODIFIED_VERSION_1, TONE_MAPPER_REINHARD_MODIFIED_VERSION_2, TONE_MAPPER_UNREAL_3) = range(6)
SHAPE_TRIANGLE_MESH, SHAPE_SPHERE, SHAPE_RECTANGLE, SHAPE_PLANE = range(4)
(MAT_MATTE, MAT_MIRROR, MAT_GLASS, MAT_METAL, MAT_PLASTIC, MAT_DISNEY, MAT_GLOSSY, MAT_SUBSTRATE, MAT_TRANSLUCENT,
 MAT_CLEAR_COAT, MAT_UBER) = range(11)
LIGHT_DIFFUSE_AREA, LIGHT_POINT, LIGHT_PEREZ, LIGHT_SPOT, LIGHT_DIRECTIONAL, LIGHT_IMAGE = range(6)
LENS_FISHEYE, LENS_THIN = 0, 1

KEY_MAXIMUM_BOUNCE = 1
KEY_MINIMUM_BOUNCE_RR = 2
KEY_RENDERING_ALGORITHM = 3
KEY_SAMPLES = 4
KEY_TONE_MAPPER = 5
KEY_PREVIEW_MODE = 6
KEY_SEED = 7
KEY_SEED_HI = 8
KEY_KERNEL = 9
KEY_LIGHT_PDF_MODE = 10
KEY_SPP_PER_PASS = 11
KEY_PROFILE = 12
KEY_COUNT_TRAVERSAL = 13
KEY_WF_BLOCK = 14
KEY_WF_POOL_PATHS = 15
KEY_WF_SORT_KEYS = 16
KEY_BVH_WIDTH = 17
KEY_USE_GRAPHS = 18
KEY_RAY_SAMPLE = 19
KEY_FILM_FILTER = 20
KEY_MAXIMUM_DISTANCE = 101

E_OK, E_NULL, E_ARGUMENT, E_STATE, E_DEVICE, E_MEMORY, E_STOPPED, E_UNSUPPORTED = 0, -1, -2, -3, -4, -5, -6, -7

BATCH_PLAIN, BATCH_PACKED = 0, 1


class Material(ctypes.Structure):
    _fields_ = [
        ("type", c_int32), ("remap_roughness", c_int32), ("thin", c_int32), ("emission_texture", c_int32),
        ("emission", c_float * 3), ("kd", c_float * 3), ("ks", c_float * 3), ("kr", c_float * 3),
        ("kt", c_float * 3), ("eta", c_float * 3), ("k", c_float * 3),
        ("angle", c_float), ("roughness_u", c_float), ("roughness_v", c_float),
        ("disney_metallic", c_float), ("disney_roughness", c_float), ("disney_specular_tint", c_float),
        ("disney_anisotropic", c_float), ("disney_sheen", c_float), ("disney_sheen_tint", c_float),
        ("disney_clearcoat", c_float), ("disney_clearcoat_gloss", c_float), ("disney_spec_trans", c_float),
        ("disney_diff_trans", c_float), ("disney_flatness", c_float),
        ("textures", c_int32 * 8),
    ]


SLOT_KD, SLOT_KS, SLOT_KR, SLOT_KT, SLOT_ROUGHNESS_U, SLOT_ROUGHNESS_V, SLOT_ETA, SLOT_K = range(8)
(TEX_CONSTANT, TEX_CHECKERBOARD, TEX_UV, TEX_SURFACE_NORMAL, TEX_DOT_PRODUCT, TEX_BULLSEYE, TEX_POLKA_DOT, TEX_BLEND,
 TEX_IMAGE, TEX_MARBLE, TEX_SIMPLEX_FBM) = range(11)


class Texture(ctypes.Structure):
    _fields_ = [("type", c_int32), ("a", c_int32), ("b", c_int32), ("image", c_int32),
                ("color", c_float * 3), ("angle", c_float), ("scale", c_float * 2), ("origin", c_float * 3),
                ("cell", c_float), ("radius", c_float), ("t", c_float * 3)]


class Primitive(ctypes.Structure):
    _fields_ = [
        ("shape", c_int32), ("material", c_int32), ("area_light", c_int32), ("reserved0", c_int32),
        ("first_triangle", c_int64), ("triangle_count", c_int64), ("data", c_float * 12),
    ]


class Light(ctypes.Structure):
    _fields_ = [
        ("type", c_int32), ("primitive", c_int32), ("triangle", c_int64),
        ("two_sided", c_int32), ("reserved0", c_int32),
        ("radiance", c_float * 3), ("position", c_float * 3), ("direction", c_float * 3),
        ("turbidity", c_float), ("cone_angle", c_float), ("cone_delta", c_float), ("radius", c_float),
        ("reserved1", c_float),
    ]


class Image(ctypes.Structure):
    _fields_ = [("width", c_int32), ("height", c_int32), ("rgb", ctypes.POINTER(c_float))]


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("n_vertices", c_int64), ("positions", ctypes.POINTER(c_float)), ("normals", ctypes.POINTER(c_float)),
        ("uvs", ctypes.POINTER(c_float)), ("n_triangles", c_int64), ("indices", ctypes.POINTER(c_int32)),
        ("n_primitives", c_int32), ("primitives", ctypes.POINTER(Primitive)),
        ("n_materials", c_int32), ("materials", ctypes.POINTER(Material)),
        ("n_lights", c_int32), ("lights", ctypes.POINTER(Light)),
        ("camera", c_float * CAMERA_LENGTH),
        ("n_images", c_int32), ("images", ctypes.POINTER(Image)),
        ("n_textures", c_int32), ("textures", ctypes.POINTER(Texture)),
    ]


class FArray(ctypes.Structure):
    _fields_ = [("data", ctypes.POINTER(c_float)), ("length", c_int64)]


class IArray(ctypes.Structure):
    _fields_ = [("data", ctypes.POINTER(c_int32)), ("length", c_int64)]


# dfr_compiled_scene: (field, kind) in header order; kind "f" float[], "i" int[]
COMPILED_SCENE_ARRAYS = [
    ("axis_aligned_bounding_boxes", "f"), ("bounding_spheres", "f"),
    ("cones", "f"), ("cylinders", "f"), ("disks", "f"), ("hyperboloids", "f"), ("paraboloids", "f"),
    ("polygons", "f"), ("rectangles", "f"), ("rectangular_cuboids", "f"), ("tori", "f"), ("triangles", "f"),
    ("triangle_meshes", "i"),
    ("blend_textures", "f"), ("bullseye_textures", "f"), ("checkerboard_textures", "f"), ("constant_textures", "f"),
    ("ldr_image_textures", "f"), ("ldr_image_texture_offsets", "i"),
    ("marble_textures", "f"), ("polka_dot_textures", "f"), ("simplex_fbm_textures", "f"),
    ("bullseye_materials", "f"), ("checkerboard_materials", "f"),
    ("clear_coat_materials", "i"), ("disney_materials", "i"), ("glass_materials", "i"), ("glossy_materials", "i"),
    ("matte_materials", "i"), ("metal_materials", "i"), ("mirror_materials", "i"), ("plastic_materials", "i"),
    ("polka_dot_materials", "f"),
    ("substrate_materials", "i"), ("translucent_materials", "i"), ("uber_materials", "i"),
    ("normal_map_ldr_image_modifiers", "f"), ("normal_map_ldr_image_modifier_offsets", "i"),
    ("simplex_noise_normal_map_modifiers", "f"),
    ("light_ids_and_offsets", "i"),
    ("diffuse_area_lights", "f"), ("directional_lights", "f"), ("image_lights", "f"), ("image_light_offsets", "i"),
    ("perez_lights", "f"), ("perez_light_offsets", "i"), ("point_lights", "f"), ("spot_lights", "f"),
    ("camera", "f"),
    ("primitives", "i"), ("primitive_matrices", "f"),
]


class CompiledScene(ctypes.Structure):
    _fields_ = [(n, FArray if k == "f" else IArray) for n, k in COMPILED_SCENE_ARRAYS] + [
        ("primitive_count", c_int32), ("light_count", c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("seconds", c_double), ("path_samples", c_int64), ("rays_closest", c_int64),
                ("rays_any", c_int64), ("bounces", c_int64), ("extend_seconds", c_double),
                ("shade_seconds", c_double), ("shadow_seconds", c_double), ("extend_launches", c_int64),
                ("extend_rays", c_int64), ("extend_nodes", c_int64), ("extend_tris", c_int64),
                ("shadow_rays", c_int64), ("shadow_nodes", c_int64), ("shadow_tris", c_int64)]


class Job(ctypes.Structure):
    _fields_ = [("scene", c_int32), ("width", c_int32), ("height", c_int32), ("spp", c_int32),
                ("maximum_bounce", c_int32), ("minimum_bounce_rr", c_int32), ("seed", c_uint64),
                ("camera", c_float * CAMERA_LENGTH)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", c_int64), ("max_active", c_int32), ("pool_paths", c_int32),
                ("spp_quantum", c_int32), ("slot_capacity", c_int32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", c_int64), ("rejected", c_int64), ("path_samples", c_int64),
                ("busy_seconds", c_double), ("peak_bytes", c_int64), ("peak_active", c_int32),
                ("reserved", c_int32)]


# Every entry point declared in include/dayflower_hip.h, with its ctypes
# signature: (name, restype, argtypes).
P = ctypes.POINTER
_c_void_p = ctypes.c_void_p
ENTRY_POINTS = [
    ("dfr_create", _c_void_p, [c_int32]),
    ("dfr_destroy", None, [_c_void_p]),
    ("dfr_last_error", ctypes.c_char_p, [_c_void_p]),
    ("dfr_set_scene", c_int32, [_c_void_p, P(SceneDesc)]),
    ("dfr_update_transforms", c_int32, [_c_void_p, _c_void_p, c_int32, c_int32]),
    ("dfr_update_matrix44fs", c_int32, [_c_void_p, _c_void_p, c_int64]),
    ("dfr_set_compiled_scene", c_int32, [_c_void_p, P(CompiledScene)]),
    ("dfr_transcode_compiled_scene", c_int32, [P(CompiledScene), P(_c_void_p), ctypes.c_char_p, c_int32]),
    ("dfr_transcoded_desc", P(SceneDesc), [_c_void_p]),
    ("dfr_transcoded_transforms", P(c_float), [_c_void_p]),
    ("dfr_transcoded_free", None, [_c_void_p]),
    ("dfr_set_resolution", c_int32, [_c_void_p, c_int32, c_int32]),
    ("dfr_set_i32", c_int32, [_c_void_p, c_int32, c_int32]),
    ("dfr_get_i32", c_int32, [_c_void_p, c_int32, P(c_int32)]),
    ("dfr_set_f32", c_int32, [_c_void_p, c_int32, c_float]),
    ("dfr_get_f32", c_int32, [_c_void_p, c_int32, P(c_float)]),
    ("dfr_setup", c_int32, [_c_void_p, c_int32]),
    ("dfr_render", c_int32, [_c_void_p, c_int32]),
    ("dfr_clear", c_int32, [_c_void_p]),
    ("dfr_get_render_pass", c_int32, [_c_void_p]),
    ("dfr_request_stop", c_int32, [_c_void_p]),
    ("dfr_update_camera", c_int32, [_c_void_p, P(c_float)]),
    ("dfr_read_film", c_int32, [_c_void_p, P(c_float), P(c_int32)]),
    ("dfr_read_rgba8", c_int32, [_c_void_p, P(ctypes.c_uint8)]),
    ("dfr_read_film_xyz", c_int32, [_c_void_p, _c_void_p]),
    ("dfr_read_film_filtered", c_int32, [_c_void_p, _c_void_p]),
    ("dfr_get_stats", c_int32, [_c_void_p, P(Stats)]),
    ("dfr_trace_rays", c_int32, [_c_void_p, _c_void_p, c_int64, c_int32, _c_void_p, _c_void_p, _c_void_p, c_int32]),
    ("dfr_trace_rays_noalloc", c_int32, [_c_void_p, _c_void_p, c_int64, c_int32, _c_void_p, _c_void_p, _c_void_p]),
    ("dfr_read_ray_sample", c_int32, [_c_void_p, _c_void_p, c_int64, _c_void_p]),
    ("dfr_read_bvh", c_int32, [_c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_void_p]),
    ("dfr_cast_primary", c_int32, [_c_void_p, c_int32, P(c_float), P(c_int32), P(c_int64), P(c_float)]),
    ("dfr_scene_add", c_int32, [_c_void_p, P(SceneDesc)]),
    ("dfr_scene_remove", c_int32, [_c_void_p, c_int32]),
    ("dfr_batch_render", c_int32, [_c_void_p, P(Job), c_int32, c_int32, P(P(c_float))]),
    ("dfr_autotune", c_int32, [_c_void_p, P(Job), c_int32, P(c_int32)]),
    ("dfr_server_create", _c_void_p, [_c_void_p, P(ServerOpts)]),
    ("dfr_wavefront_bytes", c_int64, [c_int32]),
    ("dfr_server_destroy", None, [_c_void_p]),
    ("dfr_server_submit", c_int64, [_c_void_p, P(Job)]),
    ("dfr_server_poll", c_int32, [_c_void_p, c_int64]),
    ("dfr_server_wait", c_int32, [_c_void_p, c_int64, P(c_float)]),
    ("dfr_server_latency", c_int32, [_c_void_p, c_int64, P(c_double)]),
    ("dfr_server_get_stats", c_int32, [_c_void_p, P(ServerStats)]),
]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdayflower_hip.so")

_lib = None


def load(path=LIB_PATH):
    """Load the HIP C-ABI library; raises if it has not been built."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the HIP extension first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    for name, res, args in ENTRY_POINTS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib
