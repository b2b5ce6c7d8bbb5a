// device_types.h — records of the compiled scene as they live in HBM.
// Shared by the host scene compiler (api.cpp) and the kernels.
#pragma once
#include <cstdint>

namespace dfr {

// Branching factor of the BVH the scene compiler emits and the kernels walk:
// 2 (64-byte child-pair nodes) or 4 (128-byte four-child nodes).
constexpr int BVH_WIDTH = 4;

// Lobe kinds of the compiled BSDF (M/scene/bxdf/*).
enum LobeKind : int32_t {
    LOBE_LAMBERT = 0,          // LambertianBRDF
    LOBE_OREN_NAYAR = 1,       // OrenNayarBRDF
    LOBE_SPECULAR_R = 2,       // SpecularBRDF (constant Fresnel)
    LOBE_FRESNEL_SPECULAR = 3, // FresnelSpecularBXDF
    LOBE_TS_R = 4,             // TorranceSparrowBRDF
    LOBE_TS_T = 5,             // TorranceSparrowBTDF
    LOBE_DISNEY_DIFFUSE = 6,
    LOBE_DISNEY_RETRO = 7,
    LOBE_DISNEY_SHEEN = 8,
    LOBE_DISNEY_CLEARCOAT = 9,
    LOBE_FRESNEL_BLEND = 10,   // FresnelBlendBRDF (R = diffuse, T = specular scale)
    LOBE_LAMBERT_T = 11,       // LambertianBTDF
    LOBE_SPECULAR_T = 12,      // SpecularBTDF (dielectric etaA -> etaB)
};

// BXDFType bit flags (M/scene/BXDFType.java).
enum : uint32_t { BX_REFL = 1, BX_TRANS = 2, BX_DIFF = 4, BX_GLOSSY = 8, BX_SPEC = 16 };
constexpr uint32_t BX_ALL = 31, BX_ALL_EXCEPT_SPECULAR = 15;

enum FresnelKind : int32_t { FR_CONST = 0, FR_DIELECTRIC = 1, FR_CONDUCTOR = 2, FR_DISNEY = 3 };

struct DevLobe {           // 128 bytes
    int32_t kind;
    uint32_t type;
    int32_t fresnel;
    int32_t separable;     // microfacet shadowing model
    float R[3], T[3];
    float A, B;            // Oren-Nayar
    float etaA, etaB;      // dielectric pair of transmission lobes
    float frEtaI, frEtaT;  // dielectric Fresnel
    float condEta[3], condK[3];
    float r0[3];           // Disney Fresnel
    float disneyEta, disneyMetallic;
    float alphaX, alphaY;
    float roughness, gloss, weight;
};
static_assert(sizeof(DevLobe) == 128, "lobe record size");

struct DevMaterial {       // 80 bytes
    int32_t first_lobe;
    int32_t n_lobes;       // 0 with has_bsdf = 0: pass-through surface
    int32_t has_bsdf;
    int32_t type;          // DFR_MAT_* (material-sort key)
    float eta;
    float emission[3];
    // per colour slot (DFR_SLOT_*): texture index + 1, or 0.  A textured
    // material keeps its full lobe list; which lobes exist at a hit, and
    // their colours, follow from the textures there.
    int32_t tex[8];
    int32_t tex_emission;  // texture index + 1 replacing `emission`, or 0
    int32_t remap;         // remap_roughness (for textured roughness)
    int32_t pad[2];
};

// A texture (M/scene/texture/*Texture.java); angles as cos/sin.
struct DevTexture {        // 80 bytes
    int32_t type, a, b;
    int32_t width, height; // image textures
    int32_t pad;
    int64_t texel;         // image textures: first texel in DevScene::texels
    float color[3];
    float cos_angle, sin_angle;
    float scale_u, scale_v;
    float origin[3];
    float cell, radius;
    float t[3];
};

struct DevPrim {           // 80 bytes
    int32_t shape, material, area_light;
    // bit 0: the surface frame is built in object space (DevScene::xforms);
    // bit 1 (PRIM_INSTANCE): a mesh instance whose triangle records are in
    // object space, `data` its object-to-world matrix (3 x 4, row-major)
    int32_t xformed;
    float data[12];
    float normal[4];       // rectangle: normalize(cross(normalize(b-a), normalize(c-a)))
};
constexpr int32_t PRIM_FRAME_XFORM = 1;
constexpr int32_t PRIM_INSTANCE = 2;

// One instance of a shared mesh, a leaf item of the top-level tree (80
// bytes): the world-to-object map that carries a ray into the mesh's object
// space (rows of a 3 x 4 matrix), the instance's world box (a top-level leaf
// bounds up to eight instances), the root of the mesh's tree in
// DevScene::inst_nodes and the instance's primitive.
struct DevInstance {
    float w2o[12];
    float lo[3];
    int32_t root;
    float hi[3];
    int32_t prim;
};

struct DevLight {          // 64 bytes
    int32_t type;
    int32_t prim;
    int32_t tri;           // original triangle index of a triangle area light
    int32_t two_sided;
    float radiance[3];
    float area;
    float position[3];
    int32_t perez;         // offset of the Perez record in the perez table
    // spot: cone axis, cos(cone angle) in `area`, cos(cone angle - delta) in
    // `param`; directional: unit direction towards the light, scene radius
    // in `param`
    float dir[3];
    float param;
};

// Perez sky record: parameters and the 32x32 importance distribution.
constexpr int PEREZ_RES = 32;
struct DevPerez {
    double zenith[3];
    double perez_y[5], perez_x[5], perez_lum[5];
    float sun_dir_world[3];
    float sun_dir_object[3];
    float sun_color[3];
    float theta;                 // sun zenith angle
    float radius;
    // per coefficient set (luminance, x, y) in single precision: the five
    // Perez coefficients and the denominator F(0, theta_s), which depends
    // on the sun only
    float coef_lum[5], coef_x[5], coef_y[5];
    float den_lum, den_x, den_y;
    float zd_lum, zd_x, zd_y;    // zenith value / denominator, per set
    float s0[3], s1[3], s2[3];   // daylight basis in XYZ
    float xyz_to_rgb[9];         // sRGB, row major
    float world_to_object[9];    // rotation, row major
    float object_to_world[9];
    float func_int;              // integral of the marginal
    float marginal_func[PEREZ_RES];
    float marginal_cdf[PEREZ_RES + 1];
    float cond_func[PEREZ_RES][PEREZ_RES];
    float cond_cdf[PEREZ_RES][PEREZ_RES + 1];
    float cond_int[PEREZ_RES];
};

// ImageLight record (M/scene/light/ImageLight.java:134-150, 544-584): the
// image (width x height RGB texels at `texel` in the scene's texel array),
// its rotation and scale in texture space, the light's object frame and the
// 32 x 32 importance distribution of luminance x sin(theta), laid out as the
// Perez record's.
struct DevImageLight {
    int32_t width, height;
    int64_t texel;               // first texel (3 floats each) in DevScene::texels
    float cos_angle, sin_angle;
    float scale_u, scale_v;
    float radius;
    float world_to_object[9];    // rotation, row major
    float object_to_world[9];
    float func_int;
    float marginal_func[PEREZ_RES];
    float marginal_cdf[PEREZ_RES + 1];
    float cond_func[PEREZ_RES][PEREZ_RES];
    float cond_cdf[PEREZ_RES][PEREZ_RES + 1];
    float cond_int[PEREZ_RES];
};

}  // namespace dfr
