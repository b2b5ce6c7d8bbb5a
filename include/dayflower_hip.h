/*
 * dayflower_hip.h — C-ABI of the MI355X-native path-tracing hot path.
 *
 * The entry points below replace, one for one, the renderer surface that the
 * reference's GPURenderer exposes to its callers (JavaFX GUI, README usage,
 * ImageGenerator).  Citations use M/ = src/main/java/org/dayflower/ of the
 * reference tree.
 *
 * Conventions
 *   - Every function returns 0 (DFR_OK) or a negative DFR_E_* code; the text of
 *     the last error on a context is available from dfr_last_error().
 *   - The caller owns every input array; the library copies what it needs
 *     before returning.  The context owns all device memory.
 *   - One context serialises its calls on an internal mutex (the reference
 *     methods are `synchronized`); dfr_request_stop() is lock-free.
 *   - No torch or HIP types appear here: plain pointers, sizes and PODs.
 */
#ifndef DAYFLOWER_HIP_H
#define DAYFLOWER_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Error codes                                                              */
/* ------------------------------------------------------------------------ */
#define DFR_OK                 0
#define DFR_E_NULL            -1  /* NullPointerException in the reference   */
#define DFR_E_ARGUMENT        -2  /* IllegalArgumentException                */
#define DFR_E_STATE           -3  /* call order (e.g. render before setup)   */
#define DFR_E_DEVICE          -4  /* HIP runtime error                       */
#define DFR_E_MEMORY          -5  /* allocation failed / over budget         */
#define DFR_E_STOPPED         -6  /* render stopped by dfr_request_stop      */
#define DFR_E_UNSUPPORTED     -7

/* ------------------------------------------------------------------------ */
/* Enumerations (ordinals follow the reference)                              */
/* ------------------------------------------------------------------------ */
/* M/renderer/RenderingAlgorithm.java:33-53; GPU ordinals
 * M/renderer/gpu/AbstractGPURenderer.java:83-87 */
#define DFR_ALGO_AMBIENT_OCCLUSION 0
#define DFR_ALGO_DEPTH_CAMERA      1
#define DFR_ALGO_PATH_TRACING      2
#define DFR_ALGO_RAY_CASTING       3
#define DFR_ALGO_RAY_TRACING       4

/* Tone map operators (AbstractGPURenderer.java:50-79; ordinals as there).
 * Only path tracing applies them, with exposure 1 (GPURenderer.java:312-335);
 * any other value, and every other algorithm, sRGB-encodes the film as is. */
#define DFR_TONE_MAPPER_FILMIC_CURVE_ACES_MODIFIED_VERSION_1 0
#define DFR_TONE_MAPPER_NONE                                 1
#define DFR_TONE_MAPPER_REINHARD                             2
#define DFR_TONE_MAPPER_REINHARD_MODIFIED_VERSION_1          3
#define DFR_TONE_MAPPER_REINHARD_MODIFIED_VERSION_2          4
#define DFR_TONE_MAPPER_UNREAL_3                             5

/* Shapes: M/geometry/shape/{TriangleMesh3F,Sphere3F,Rectangle3F,Plane3F}.java */
#define DFR_SHAPE_TRIANGLE_MESH 0
#define DFR_SHAPE_SPHERE        1
#define DFR_SHAPE_RECTANGLE     2
#define DFR_SHAPE_PLANE         3

/* Materials: M/scene/material/{Matte,Mirror,Glass,Metal,Plastic,Disney,
 * Glossy,Substrate,Translucent,ClearCoat,Uber}Material.java */
#define DFR_MAT_MATTE       0
#define DFR_MAT_MIRROR      1
#define DFR_MAT_GLASS       2
#define DFR_MAT_METAL       3
#define DFR_MAT_PLASTIC     4
#define DFR_MAT_DISNEY      5
#define DFR_MAT_GLOSSY      6
#define DFR_MAT_SUBSTRATE   7
#define DFR_MAT_TRANSLUCENT 8
#define DFR_MAT_CLEAR_COAT  9
#define DFR_MAT_UBER       10

/* Lights: M/scene/light/{DiffuseAreaLight,PointLight,PerezLight,SpotLight,
 * DirectionalLight,ImageLight}.java */
#define DFR_LIGHT_DIFFUSE_AREA 0
#define DFR_LIGHT_POINT        1
#define DFR_LIGHT_PEREZ        2
#define DFR_LIGHT_SPOT         3
#define DFR_LIGHT_DIRECTIONAL  4
#define DFR_LIGHT_IMAGE        5

/* Camera lens ordinals: M/scene/Lens.java:31-36 */
#define DFR_LENS_FISHEYE 0
#define DFR_LENS_THIN    1

/* Camera record: the 19-float layout of
 * M/scene/compiler/CompiledCameraCache.java:40-102,158-197 */
#define DFR_CAMERA_FOV_X        0  /* tan(+fovX/2)                        */
#define DFR_CAMERA_FOV_Y        1  /* tan(-fovY/2)                        */
#define DFR_CAMERA_LENS         2
#define DFR_CAMERA_U            3
#define DFR_CAMERA_V            6
#define DFR_CAMERA_W            9
#define DFR_CAMERA_EYE         12
#define DFR_CAMERA_APERTURE    15
#define DFR_CAMERA_FOCAL       16
#define DFR_CAMERA_RES_X       17
#define DFR_CAMERA_RES_Y       18
#define DFR_CAMERA_LENGTH      19

/* ------------------------------------------------------------------------ */
/* Scene description (input of setScene; compiled into SoA at setup)        */
/* ------------------------------------------------------------------------ */

/* One material.  Which fields are read depends on `type`:
 *   MATTE   kd, angle (degrees; 0 => Lambertian, else Oren-Nayar), emission
 *   MIRROR  kr, emission
 *   GLASS   kr, kt, eta[0], roughness_u/v, remap_roughness, emission
 *   METAL   eta[0..2], k[0..2], roughness_u/v, remap_roughness, emission
 *   PLASTIC kd, ks, roughness_u, remap_roughness, emission
 *   DISNEY  kd (base color), the disney_* scalars, eta[0], emission (thin
 *           != 0 is refused with DFR_E_UNSUPPORTED)
 *   GLOSSY      kr, roughness_u (used as is), emission
 *   SUBSTRATE   kd, ks, roughness_u/v, remap_roughness, emission
 *   TRANSLUCENT kd, ks, kr (reflectance), kt (transmittance), roughness_u,
 *               remap_roughness, emission
 *   CLEAR_COAT  kd, ks, emission
 *   UBER        kd, ks, kr, kt, k (opacity), eta[0], roughness_u/v,
 *               remap_roughness, emission
 * textures[k] (k = DFR_SLOT_*) replaces that parameter by a texture of the
 * scene (index + 1; 0 keeps the constant): MATTE kd; PLASTIC kd, ks,
 * roughness_u; MIRROR kr; GLASS kr, kt, eta (the colour average);
 * METAL eta, k, roughness_u,
 * roughness_v; SUBSTRATE kd, ks, roughness_u, roughness_v; GLOSSY kr,
 * roughness_u (as is); CLEAR_COAT kd, ks.  A roughness slot reads the texture's colour average
 * (Texture.getFloat), remapped when remap_roughness is set.  A texture on
 * another material or slot is refused with DFR_E_UNSUPPORTED.  emission_texture textures the emission
 * of every material type (Material.emittance reads textureEmission). */
#define DFR_SLOT_KD 0
#define DFR_SLOT_KS 1
#define DFR_SLOT_KR 2
#define DFR_SLOT_KT 3
#define DFR_SLOT_ROUGHNESS_U 4
#define DFR_SLOT_ROUGHNESS_V 5
#define DFR_SLOT_ETA 6
#define DFR_SLOT_K 7
typedef struct dfr_material {
    int32_t type;
    int32_t remap_roughness;
    int32_t thin;
    int32_t emission_texture;   /* any material: a texture replacing emission (index + 1; 0: the constant) */
    float emission[3];
    float kd[3];
    float ks[3];
    float kr[3];
    float kt[3];
    float eta[3];
    float k[3];
    float angle;
    float roughness_u;
    float roughness_v;
    float disney_metallic;
    float disney_roughness;
    float disney_specular_tint;
    float disney_anisotropic;
    float disney_sheen;
    float disney_sheen_tint;
    float disney_clearcoat;
    float disney_clearcoat_gloss;
    float disney_spec_trans;
    float disney_diff_trans;
    float disney_flatness;
    int32_t textures[8];
} dfr_material;

/* Textures: M/scene/texture/{Constant,Checkerboard,UV,SurfaceNormal,
 * DotProduct,Bullseye,PolkaDot,Blend,LDRImage,Marble,
 * SimplexFractionalBrownianMotion}Texture.java.
 *   CONSTANT        color
 *   CHECKERBOARD    a, b (texture indices), angle (degrees), scale (u, v)
 *   UV              (u, v, 0)
 *   SURFACE_NORMAL  (n + 1) / 2 of the shading normal facing the ray
 *   DOT_PRODUCT     |dot(shading normal, ray direction)|
 *   BULLSEYE        a, b, origin, scale[0]
 *   POLKA_DOT       a, b, angle (degrees), cell (resolution), radius
 *   BLEND           a, b (not blends themselves), t (per component)
 *   IMAGE           image (index into the scene's images), angle (degrees),
 *                   scale (u, v); bilinear with wrap-around
 *   MARBLE          colours A = color, B = origin, C = t; scale[0] the
 *                   frequency, scale[1] the scale, cell the octaves (an
 *                   integer, 0..64) of the Perlin turbulence at the
 *                   world-space hit point
 *   SIMPLEX_FBM     color times simplex-noise fBm in [0, 1] at the
 *                   world-space hit point; scale[0] the frequency, scale[1]
 *                   the gain, cell the octaves (an integer, 0..64) */
#define DFR_TEX_CONSTANT       0
#define DFR_TEX_CHECKERBOARD   1
#define DFR_TEX_UV             2
#define DFR_TEX_SURFACE_NORMAL 3
#define DFR_TEX_DOT_PRODUCT    4
#define DFR_TEX_BULLSEYE       5
#define DFR_TEX_POLKA_DOT      6
#define DFR_TEX_BLEND          7
#define DFR_TEX_IMAGE          8
#define DFR_TEX_MARBLE         9
#define DFR_TEX_SIMPLEX_FBM   10
typedef struct dfr_texture {
    int32_t type;
    int32_t a, b;
    int32_t image;
    float color[3];
    float angle;
    float scale[2];
    float origin[3];
    float cell;
    float radius;
    float t[3];
} dfr_texture;

/* One primitive (M/scene/Primitive.java), given in its object space; its
 * object-to-world matrix is the identity unless dfr_update_transforms sets
 * one (a caller may also fold transforms in and leave the identity).
 *   TRIANGLE_MESH  triangles [first_triangle, first_triangle+triangle_count);
 *                  mesh primitives that name the same range are instances
 *                  of one mesh (its triangles are kept once, in object
 *                  space, under a BVH of their own, with a top-level BVH
 *                  over the instances' world boxes; the reference keeps one
 *                  record per distinct TriangleMesh3F, Shape3FCache.java
 *                  :176-177, 298-300); ranges that overlap without being
 *                  equal are DFR_E_ARGUMENT, an area light on a shared range
 *                  DFR_E_UNSUPPORTED
 *   SPHERE         data = center.xyz, radius
 *   RECTANGLE      data = corners a.xyz, b.xyz, c.xyz  (Rectangle3F(a,b,c))
 *   PLANE          data = origin.xyz, u.xyz, v.xyz, w.xyz (w = normal)
 * area_light is an index into the light array or -1. */
typedef struct dfr_primitive {
    int32_t shape;
    int32_t material;
    int32_t area_light;
    int32_t reserved0;
    int64_t first_triangle;
    int64_t triangle_count;
    float data[12];
} dfr_primitive;

/* One light.
 *   DIFFUSE_AREA  radiance, two_sided, primitive (its shape), triangle
 *                 (global triangle index when the shape is one triangle of a
 *                 mesh primitive; -1 when the shape is a sphere primitive,
 *                 whose area_light must name this light)
 *   POINT         radiance = intensity, position
 *   PEREZ         turbidity, direction = sun direction (world space)
 *   SPOT          radiance = intensity, position, direction = cone axis
 *                 (world space), cone_angle and cone_delta in degrees
 *   DIRECTIONAL   radiance, direction = towards the light (world space,
 *                 normalised by the compiler), radius = scene radius (the
 *                 shadow ray's length is twice it)
 *   IMAGE         an environment image (ImageLight.java:134-150, 544-584):
 *                 primitive = index into the scene's images, cone_angle =
 *                 the image's rotation in degrees (AngleF), direction[0..1]
 *                 = its scale (u, v; (1, 1) when both are 0), radius (10 when
 *                 0; sampled directions end at twice it); looked up at the
 *                 spherical coordinates of the direction in the light's
 *                 object frame with bilinear filtering and wrap-around */
typedef struct dfr_light {
    int32_t type;
    int32_t primitive;
    int64_t triangle;
    int32_t two_sided;
    int32_t reserved0;
    float radiance[3];
    float position[3];
    float direction[3];
    float turbidity;
    float cone_angle;
    float cone_delta;
    float radius;
    float reserved1;
} dfr_light;

/* One image: width * height linear RGB texels (3 floats each), row 0 first
 * (an ImageF's colours; an 8-bit image's bytes divided by 255, as
 * Color3F.unpack gives an LDRImageTexture's texels). */
typedef struct dfr_image {
    int32_t width, height;
    const float* rgb;
} dfr_image;

typedef struct dfr_scene_desc {
    int64_t n_vertices;
    const float* positions;   /* 3 * n_vertices                         */
    const float* normals;     /* 3 * n_vertices, or NULL (flat)         */
    const float* uvs;         /* 2 * n_vertices, or NULL                */
    int64_t n_triangles;
    const int32_t* indices;   /* 3 * n_triangles                        */
    int32_t n_primitives;
    const dfr_primitive* primitives;
    int32_t n_materials;
    const dfr_material* materials;
    int32_t n_lights;
    const dfr_light* lights;
    float camera[DFR_CAMERA_LENGTH];
    int32_t n_images;
    const struct dfr_image* images;  /* images of IMAGE lights and textures */
    int32_t n_textures;
    const dfr_texture* textures;
} dfr_scene_desc;

/* ------------------------------------------------------------------------ */
/* Renderer context: the GPURenderer surface                                */
/* ------------------------------------------------------------------------ */
typedef struct dfr_ctx dfr_ctx;

/* Integer / float settings (setters of AbstractGPURenderer.java:447-526). */
#define DFR_KEY_MAXIMUM_BOUNCE          1  /* default 20                       */
#define DFR_KEY_MINIMUM_BOUNCE_RR       2  /* default 5                        */
#define DFR_KEY_RENDERING_ALGORITHM     3  /* default DFR_ALGO_PATH_TRACING    */
#define DFR_KEY_SAMPLES                 4  /* AO samples, default 10           */
#define DFR_KEY_TONE_MAPPER             5  /* default DFR_TONE_MAPPER_NONE      */
#define DFR_KEY_PREVIEW_MODE            6  /* default 0                        */
#define DFR_KEY_SEED                    7  /* 64-bit seed, low 32 bits         */
#define DFR_KEY_SEED_HI                 8  /* 64-bit seed, high 32 bits        */
#define DFR_KEY_KERNEL                  9  /* 0 wavefront (default), 1 plain   */
#define DFR_KEY_LIGHT_PDF_MODE         10  /* 0 reference (default), 1 area    */
#define DFR_KEY_SPP_PER_PASS           11  /* samples per pixel per pass, 1    */
#define DFR_KEY_PROFILE                12  /* 1: time each kernel launch       */
#define DFR_KEY_COUNT_TRAVERSAL        13  /* 1: count BVH node / triangle visits */
/* wavefront engine configuration (what dfr_autotune sets)                   */
#define DFR_KEY_WF_BLOCK               14  /* threads per block: 128 or 256     */
#define DFR_KEY_WF_POOL_PATHS          15  /* path pool capacity, 1024 .. 2^26  */
#define DFR_KEY_WF_SORT_KEYS           16  /* 1: shade in (material, light type) order, 0: queue order */
#define DFR_KEY_BVH_WIDTH              17  /* read-only: children per BVH node (2 or 4) */
#define DFR_KEY_USE_GRAPHS             18  /* 1 (default): replay captured HIP graphs of 4 wavefront iterations */
#define DFR_KEY_FILM_FILTER            20  /* 0 (default): GPURenderer's running-mean film only;
                                              1: also CPURenderer's Gaussian XYZ film (see
                                              dfr_read_film_filtered); a change clears both films */
#define DFR_KEY_RAY_SAMPLE             19  /* n > 0 with COUNT_TRAVERSAL: record every n-th traced ray */
#define DFR_KEY_MAXIMUM_DISTANCE      101  /* float, default 20                */

/* new GPURenderer(observer): M/renderer/gpu/GPURenderer.java:55-57 with the
 * defaults of M/renderer/gpu/AbstractGPURenderer.java:143-159.
 * device < 0 selects the current HIP device.  Returns NULL on failure. */
dfr_ctx* dfr_create(int32_t device);

/* dispose(): M/renderer/gpu/AbstractKernel.java:166-168 */
void dfr_destroy(dfr_ctx* ctx);

/* Text of the last error on this context (empty string if none). */
const char* dfr_last_error(const dfr_ctx* ctx);

/* setScene(Scene): M/renderer/gpu/AbstractSceneKernel.java:247-250.
 * Copies the description; nothing is compiled until dfr_setup(). */
int32_t dfr_set_scene(dfr_ctx* ctx, const dfr_scene_desc* scene);

/* updateMatrix44Fs(): M/renderer/gpu/AbstractSceneKernel.java:292-300, the
 * primitives' object-to-world matrices of CompiledPrimitiveCache.toMatrix44Fs
 * (M/scene/compiler/CompiledPrimitiveCache.java:323-350).  `object_to_world`
 * holds 16 floats per primitive of the scene set by dfr_set_scene, row major
 * (element (r, c) at 4r + c), in primitive order; `n_primitives` must equal
 * the scene's primitive count (the reference ignores an array of another
 * length; this returns DFR_E_ARGUMENT).  A primitive's record (vertices,
 * sphere centre and radius, rectangle corners, plane frame) is in its object
 * space; surface frames are built there and carried to world space through
 * the normal matrix, as the reference does.  Matrices must be affine (last
 * row 0 0 0 1) and invertible.  Spheres and rectangles take rotations,
 * reflections, uniform scales and translations; meshes and planes any affine
 * map.  A mesh carrying an area light takes rotations and translations only
 * and a sphere light translations only; anything else is DFR_E_UNSUPPORTED.
 * Before the scene is compiled the matrices are kept for dfr_setup; after,
 * the world-space geometry is recomputed and uploaded at once, and the BVH
 * keeps its topology with its boxes refitted (rebuild = 0) or is built
 * again (rebuild = 1).  Refused (DFR_E_STATE) while a serving loop renders
 * the scene.  dfr_set_scene resets every primitive to the identity. */
int32_t dfr_update_transforms(dfr_ctx* ctx, const float* object_to_world, int32_t n_primitives, int32_t rebuild);

/* updateMatrix44Fs() with the array the reference uploads verbatim:
 * CompiledPrimitiveCache.getMatrix44Fs() (M/scene/compiler/
 * CompiledPrimitiveCache.java:323-350), 32 floats per primitive (object to
 * world, then world to object, each row major).  The reference only takes an
 * array of the length it already has (AbstractSceneKernel.java:292-300);
 * here `length` must be 32 x the scene's primitive count (DFR_E_ARGUMENT
 * otherwise).  The object-to-world halves go to dfr_update_transforms (BVH
 * refit); the world-to-object halves are not read (the inverse is computed in
 * double precision). */
int32_t dfr_update_matrix44fs(dfr_ctx* ctx, const float* matrices, int64_t length);

/* ------------------------------------------------------------------------ */
/* The reference's compiled scene, verbatim                                 */
/* ------------------------------------------------------------------------ */

/* One array of a CompiledScene as the reference uploads it
 * (doGetCompatibleArray: an empty array travels as one zero). */
typedef struct dfr_farray { const float* data; int64_t length; } dfr_farray;
typedef struct dfr_iarray { const int32_t* data; int64_t length; } dfr_iarray;

/* Every array doSetupScene() puts on the device (M/renderer/gpu/
 * AbstractSceneKernel.java:1777-1855), by the getter that produced it, plus
 * the two counts it keeps (primitiveCount, lightCount).  A Java caller fills
 * it from SceneCompiler.compile(scene) without touching the arrays. */
typedef struct dfr_compiled_scene {
    /* CompiledBoundingVolume3FCache (read for bounds only: the library
     * builds its own BVH) */
    dfr_farray axis_aligned_bounding_boxes, bounding_spheres;
    /* CompiledShape3FCache */
    dfr_farray cones, cylinders, disks, hyperboloids, paraboloids, polygons, rectangles, rectangular_cuboids, tori,
        triangles;
    dfr_iarray triangle_meshes;
    /* CompiledTextureCache */
    dfr_farray blend_textures, bullseye_textures, checkerboard_textures, constant_textures, ldr_image_textures;
    dfr_iarray ldr_image_texture_offsets;
    dfr_farray marble_textures, polka_dot_textures, simplex_fbm_textures;
    /* CompiledMaterialCache */
    dfr_farray bullseye_materials, checkerboard_materials;
    dfr_iarray clear_coat_materials, disney_materials, glass_materials, glossy_materials, matte_materials,
        metal_materials, mirror_materials, plastic_materials;
    dfr_farray polka_dot_materials;
    dfr_iarray substrate_materials, translucent_materials, uber_materials;
    /* CompiledModifierCache */
    dfr_farray normal_map_ldr_image_modifiers;
    dfr_iarray normal_map_ldr_image_modifier_offsets;
    dfr_farray simplex_noise_normal_map_modifiers;
    /* CompiledLightCache */
    dfr_iarray light_ids_and_offsets;
    dfr_farray diffuse_area_lights, directional_lights, image_lights;
    dfr_iarray image_light_offsets;
    dfr_farray perez_lights;
    dfr_iarray perez_light_offsets;
    dfr_farray point_lights, spot_lights;
    /* CompiledCameraCache (19 floats) */
    dfr_farray camera;
    /* CompiledPrimitiveCache: 8 ints and 32 floats per primitive */
    dfr_iarray primitives;
    dfr_farray primitive_matrices;
    int32_t primitive_count, light_count;
} dfr_compiled_scene;

/* setScene(Scene) for a Java caller that hands over SceneCompiler.compile's
 * arrays (the drop-in path of M/renderer/gpu/AbstractSceneKernel.java
 * :1777-1855): the arrays are transcoded into a dfr_scene_desc plus one
 * object-to-world matrix per primitive, then set as dfr_set_scene and
 * dfr_update_transforms would.  What the kernels cannot follow is refused
 * with DFR_E_UNSUPPORTED and a message naming it: Cone3F, Cylinder3F, Disk3F,
 * Hyperboloid3F, Paraboloid3F, Polygon3F, RectangularCuboid3F and Torus3F
 * primitives; the Bullseye, Checkerboard and PolkaDot materials;
 * FunctionTexture; any modifier other than NoOpModifier; a texture on a
 * parameter other than the colour slots and the emission of dfr_material
 * (a non-constant texture elsewhere); an area light on a shape other than a Triangle3F or a
 * Sphere3F, or one no primitive carries.  Malformed arrays (an offset outside
 * its array, an unknown type ID) give DFR_E_ARGUMENT. */
int32_t dfr_set_compiled_scene(dfr_ctx* ctx, const dfr_compiled_scene* scene);

/* The same transcoding without a context or a device: on success *out holds
 * the scene (dfr_transcoded_desc, dfr_transcoded_transforms) until
 * dfr_transcoded_free; on failure *out is NULL and `error` (capacity bytes,
 * may be NULL) holds the message. */
typedef struct dfr_transcoded dfr_transcoded;
int32_t dfr_transcode_compiled_scene(const dfr_compiled_scene* scene, dfr_transcoded** out, char* error,
                                     int32_t error_capacity);
const dfr_scene_desc* dfr_transcoded_desc(const dfr_transcoded* t);
const float* dfr_transcoded_transforms(const dfr_transcoded* t);  /* 16 per primitive, row major */
void dfr_transcoded_free(dfr_transcoded* t);

/* setImage(): M/renderer/gpu/AbstractGPURenderer.java:417-440.  Sets the
 * film resolution; takes effect at the next setup/render. */
int32_t dfr_set_resolution(dfr_ctx* ctx, int32_t width, int32_t height);

/* setMaximumBounce, setMinimumBounceRussianRoulette, setRenderingAlgorithm,
 * setSamples, setToneMapper, setPreviewMode, setMaximumDistance:
 * M/renderer/gpu/AbstractGPURenderer.java:447-526 */
int32_t dfr_set_i32(dfr_ctx* ctx, int32_t key, int32_t value);
int32_t dfr_get_i32(const dfr_ctx* ctx, int32_t key, int32_t* value);
int32_t dfr_set_f32(dfr_ctx* ctx, int32_t key, float value);
int32_t dfr_get_f32(const dfr_ctx* ctx, int32_t key, float* value);

/* setup() / setup(boolean): M/renderer/gpu/AbstractGPURenderer.java:532-556.
 * Compiles the scene (BVH + SoA) when compile_scene != 0, uploads it and
 * allocates the film. */
int32_t dfr_setup(dfr_ctx* ctx, int32_t compile_scene);

/* render(n): M/renderer/gpu/AbstractGPURenderer.java:231-332.  n < 1 is
 * DFR_E_ARGUMENT.  Each pass adds SPP_PER_PASS samples to every pixel. */
int32_t dfr_render(dfr_ctx* ctx, int32_t passes);

/* clear(): lazy reset of the film at the next render
 * (M/renderer/gpu/AbstractGPURenderer.java:209-211, 382-411). */
int32_t dfr_clear(dfr_ctx* ctx);
int32_t dfr_get_render_pass(const dfr_ctx* ctx);

/* renderShutdown(): M/renderer/gpu/AbstractGPURenderer.java:342-344. */
int32_t dfr_request_stop(dfr_ctx* ctx);

/* updateCamera(): M/renderer/gpu/AbstractSceneKernel.java:278-300 */
int32_t dfr_update_camera(dfr_ctx* ctx, const float camera[DFR_CAMERA_LENGTH]);

/* Film readback (the running mean of M/renderer/gpu/AbstractImageKernel.java
 * :960-995).  rgb: 3 floats per pixel (row-major, y down), count: samples
 * per pixel.  Either pointer may be NULL. */
int32_t dfr_read_film(dfr_ctx* ctx, float* rgb, int32_t* count);

/* CPURenderer's film: PixelImageF's XYZ film with GaussianFilter2F(2, 2, 2)
 * (M/image/PixelImageF.java:569-617, 662-705; M/filter/GaussianFilter2F.java)
 * in ColorSpaceF.S_R_G_B, kept beside the running-mean film while
 * DFR_KEY_FILM_FILTER is 1.  Every sample splats into the pixels whose
 * centres lie within two pixels of its position (pixel + camera offset),
 * weighted by the 16 x 16 Gaussian table; samples the reference would drop
 * (no camera ray; non-finite or negative-luminance radiance on the plain
 * kernel) add nothing.  dfr_read_film_xyz copies width * height * 4 floats
 * (X, Y, Z, filter weight sum); dfr_read_film_filtered copies width *
 * height * 3 floats of filmRender()'s output: linear sRGB divided by the
 * weight sum, clamped at zero and gamma encoded.  DFR_E_STATE when the
 * filter is off. */
int32_t dfr_read_film_xyz(dfr_ctx* ctx, float* xyzw);
int32_t dfr_read_film_filtered(dfr_ctx* ctx, float* rgb);

/* RGBA8 image (imageBegin..imageEnd, AbstractImageKernel.java:1104-1324;
 * GPURenderer.java:312-335): the tone map selected by DFR_KEY_TONE_MAPPER
 * when the algorithm is path tracing, sRGB encode (break point 0.00304,
 * slope 12.923, gamma 2.4), then (int)(clamp(v, 0, 1) * 255 + 0.5); alpha
 * 255.  rgba: 4 bytes per pixel, row-major, y down. */
int32_t dfr_read_rgba8(dfr_ctx* ctx, uint8_t* rgba);

/* Counters of the last render call. */
typedef struct dfr_stats {
    double   seconds;            /* device time of the last render call     */
    int64_t  path_samples;       /* camera samples traced                   */
    int64_t  rays_closest;       /* closest-hit queries                     */
    int64_t  rays_any;           /* any-hit (shadow / AO) queries           */
    int64_t  bounces;            /* path vertices shaded                    */
    /* with DFR_KEY_PROFILE: summed HIP-event durations per kernel kind     */
    double   extend_seconds;     /* closest-hit traversal (wavefront extend) */
    double   shade_seconds;
    double   shadow_seconds;
    int64_t  extend_launches;
    int64_t  extend_rays;
    /* with DFR_KEY_COUNT_TRAVERSAL: visits made by the extend kernel       */
    int64_t  extend_nodes;       /* two-wide nodes fetched                  */
    int64_t  extend_tris;        /* triangle tests                          */
    int64_t  shadow_rays;        /* shadow-stage queries (any and closest)  */
    int64_t  shadow_nodes;       /* with DFR_KEY_COUNT_TRAVERSAL            */
    int64_t  shadow_tris;
} dfr_stats;
int32_t dfr_get_stats(const dfr_ctx* ctx, dfr_stats* stats);

/* Measurement hooks (no counterpart in the reference): the rays a counting
 * render recorded and the compiled BVH, so that an independent traversal can
 * recount the node visits and triangle tests the roofline prices.
 *
 * dfr_read_ray_sample: records of 12 floats -- origin.xyz, far limit at BVH
 * entry, direction.xyz, kind (int bits: 0 closest hit, 1 any hit), node
 * visits, triangle tests, hit primitive, hit triangle (int bits) -- of the
 * last render with DFR_KEY_COUNT_TRAVERSAL and DFR_KEY_RAY_SAMPLE set (at
 * most 2^20).  With records NULL only *n_records is set.
 *
 * dfr_read_bvh: the main scene's four-wide BVH (over the triangles of the
 * meshes no other primitive shares), 32 floats per node (lo.x[4],
 * hi.x[4], lo.y[4], hi.y[4], lo.z[4], hi.z[4], child references[4] as uint32
 * bits, 4 unused), and its triangles in BVH order, 12 floats each (a.xyz,
 * original triangle id as int bits, b.xyz, primitive as int bits, c.xyz, 0).
 * A child reference is a node index, 0xffffffff (empty slot) or a leaf:
 * bit 31 set, bits 3..30 the first triangle, bits 0..2 count - 1.  On entry
 * *n_nodes / *n_tris give the buffers' capacities; on return the counts.
 * NULL buffers query the counts only. */
int32_t dfr_read_ray_sample(dfr_ctx* ctx, float* records, int64_t max_records, int64_t* n_records);
int32_t dfr_read_bvh(dfr_ctx* ctx, float* nodes, int64_t* n_nodes, float* tris, int64_t* n_tris);

/* ------------------------------------------------------------------------ */
/* Ray queries (ray casting / AO hit parity, BVH bench)                      */
/* ------------------------------------------------------------------------ */
/* A ray is 8 floats: origin.xyz, tmin, direction.xyz, tmax.
 * Hit records: prim (primitive index or -1), tri (global triangle index or
 * -1), t (world-space distance; undefined on a miss).
 * When `device_ptrs` != 0 every pointer is a device pointer already resident
 * in HBM; otherwise they are host pointers copied in and out.
 * any_hit != 0: stop at the first hit (M/scene/Scene.java intersects). */
int32_t dfr_trace_rays(dfr_ctx* ctx, const float* rays, int64_t n_rays,
                       int32_t any_hit, int32_t* prim, int64_t* tri, float* t,
                       int32_t device_ptrs);

/* The closest/any-hit query of dfr_trace_rays on device pointers, allocating
 * nothing: the traversal stacks' overflow lives in a static device array of
 * 24 MiB that the library holds on each device it runs on (1024 blocks of
 * 256 threads stride over the rays; 24 overflow entries per thread).  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  rays, prim, tri and t are device pointers.  Returns
 * DFR_E_UNSUPPORTED when the scene's BVH needs more overflow entries per
 * thread than the array holds.  Calls that overlap on one device share the
 * array and race. */
int32_t dfr_trace_rays_noalloc(dfr_ctx* ctx, const float* rays, int64_t n_rays, int32_t any_hit, int32_t* prim,
                               int64_t* tri, float* t);

/* Primary rays of the current camera for pass `sample_index` with the
 * uniform sub-pixel offset of ray casting (GPURenderer.java:338-369), traced
 * to the closest hit.  Writes the rays used (8 floats each) and hit records,
 * one per pixel.  Host pointers. */
int32_t dfr_cast_primary(dfr_ctx* ctx, int32_t sample_index, float* rays,
                         int32_t* prim, int64_t* tri, float* t);

/* ------------------------------------------------------------------------ */
/* Resident scenes, batched jobs and serving                                 */
/* ------------------------------------------------------------------------ */
/* A render job: one camera over one resident scene at one resolution and
 * spp.  Jobs never share RNG streams: sample s of pixel p of job j uses the
 * counter (seed_j, p, s), so a job's film is the same alone or batched. */
typedef struct dfr_job {
    int32_t scene;          /* id from dfr_scene_add                        */
    int32_t width;
    int32_t height;
    int32_t spp;
    int32_t maximum_bounce;
    int32_t minimum_bounce_rr;
    uint64_t seed;
    float camera[DFR_CAMERA_LENGTH];
} dfr_job;

/* Compile and keep a scene resident on the device; returns its id (>= 0)
 * or a negative error code. */
int32_t dfr_scene_add(dfr_ctx* ctx, const dfr_scene_desc* scene);
int32_t dfr_scene_remove(dfr_ctx* ctx, int32_t scene);

#define DFR_BATCH_PLAIN   0   /* one launch sequence per job, in order      */
#define DFR_BATCH_PACKED  1   /* all jobs share persistent wavefront launches */

/* Render `n_jobs` jobs to completion.  films[j] receives width*height*3
 * floats (mean radiance) of job j (host pointers; NULL to skip readback).
 * Returns 0 or an error; per-batch timing in dfr_get_stats. */
int32_t dfr_batch_render(dfr_ctx* ctx, const dfr_job* jobs, int32_t n_jobs,
                         int32_t mode, float* const* films);

/* Autotune the packed wavefront launch (block size, path-pool capacity,
 * material-sort granularity) on a job list; the chosen values are kept on
 * the context and reported through `out` (3 ints) when it is not NULL. */
int32_t dfr_autotune(dfr_ctx* ctx, const dfr_job* jobs, int32_t n_jobs, int32_t* out);

/* Serving: a native request loop with continuous batching. */
typedef struct dfr_server dfr_server;

typedef struct dfr_server_opts {
    int64_t hbm_budget_bytes;   /* admission budget for films + queues      */
    int32_t max_active;         /* at most this many requests on the device */
    int32_t pool_paths;         /* path pool size of the wavefront engine   */
    int32_t spp_quantum;        /* reserved (0)                              */
    int32_t slot_capacity;      /* job slots (0: 65536); slots of finished
                                   requests are reused                       */
} dfr_server_opts;

dfr_server* dfr_server_create(dfr_ctx* ctx, const dfr_server_opts* opts);
/* Device bytes of the wavefront engine for a path pool of `pool_paths`
 * (counted against hbm_budget_bytes before any film). */
int64_t dfr_wavefront_bytes(int32_t pool_paths);
void dfr_server_destroy(dfr_server* server);
/* Submit a request (non-blocking); returns its id (>= 0) or an error. */
int64_t dfr_server_submit(dfr_server* server, const dfr_job* job);
/* 1 done, 0 pending, <0 error (including rejected by admission). */
int32_t dfr_server_poll(dfr_server* server, int64_t request);
/* Block until the request is done; then copy its film (3 floats/pixel).  A
 * film is handed over once: a second wait with film != NULL returns
 * DFR_E_STATE (NULL only waits). */
int32_t dfr_server_wait(dfr_server* server, int64_t request, float* film);
/* Latency (submit -> film ready) in seconds of request `request`. */
int32_t dfr_server_latency(dfr_server* server, int64_t request, double* seconds);
typedef struct dfr_server_stats {
    int64_t completed;
    int64_t rejected;
    int64_t path_samples;
    double  busy_seconds;
    int64_t peak_bytes;
    int32_t peak_active;
    int32_t reserved;
} dfr_server_stats;
int32_t dfr_server_get_stats(dfr_server* server, dfr_server_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* DAYFLOWER_HIP_H */
