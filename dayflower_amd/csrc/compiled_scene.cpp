// compiled_scene.cpp — the reference's CompiledScene arrays, transcoded into
// a dfr_scene_desc plus one object-to-world matrix per primitive.
//
// The arrays are the ones doSetupScene uploads (M/renderer/gpu/
// AbstractSceneKernel.java:1777-1855); their layouts are those of
// M/scene/compiler/Compiled{Primitive,Shape3F,Material,Texture,Light,
// Camera}Cache.java, cited where each record is read.  M/ =
// src/main/java/org/dayflower/ of the reference tree.
//
// Offsets in the arrays follow the compiler's two conventions: shapes and
// materials are addressed by element offset (SceneCompiler maps each distinct
// record to its first element), textures, lights and modifiers by index (the
// kernels multiply by the record length), and the variable-length records
// (LDR image textures, image and Perez lights) through their offset arrays.
// Texture and material references pack (type ID, offset) into one int: two
// 16-bit halves (Ints.pack(a, b), M/utility/Ints.java:237-242) in primitives,
// light tables and texture records, four bytes (Ints.pack(a, b, c, d),
// :258-265) in material records, first pair in the low bytes.
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dayflower_hip.h"

namespace dfr {
namespace {

// ---------------------------------------------------------------------------
// Type IDs (the ID constant of each class) and record layouts
// ---------------------------------------------------------------------------
// shapes (M/geometry/shape/*3F.java)
constexpr int SHAPE_CONE = 2, SHAPE_CYLINDER = 5, SHAPE_DISK = 6, SHAPE_HYPERBOLOID = 7, SHAPE_PARABOLOID = 9,
              SHAPE_PLANE = 10, SHAPE_POLYGON = 11, SHAPE_RECTANGLE = 13, SHAPE_RECTANGULAR_CUBOID = 14,
              SHAPE_SPHERE = 15, SHAPE_TORUS = 16, SHAPE_TRIANGLE = 17, SHAPE_TRIANGLE_MESH = 18;
// materials (M/scene/material/*Material.java)
constexpr int MAT_BULLSEYE = 1, MAT_CHECKERBOARD = 2, MAT_CLEAR_COAT = 3, MAT_DISNEY = 4, MAT_GLASS = 7,
              MAT_GLOSSY = 8, MAT_MATTE = 11, MAT_METAL = 12, MAT_MIRROR = 13, MAT_PLASTIC = 14,
              MAT_POLKA_DOT = 15, MAT_SUBSTRATE = 16, MAT_TRANSLUCENT = 18, MAT_UBER = 19;
// textures (M/scene/texture/*Texture.java)
constexpr int TEX_BLEND = 1, TEX_BULLSEYE = 2, TEX_CHECKERBOARD = 3, TEX_CONSTANT = 4, TEX_DOT_PRODUCT = 5,
              TEX_FUNCTION = 6, TEX_LDR_IMAGE = 7, TEX_MARBLE = 8, TEX_POLKA_DOT = 9, TEX_SIMPLEX_FBM = 10,
              TEX_SURFACE_NORMAL = 11, TEX_UV = 12;
// lights (M/scene/light/*Light.java)
constexpr int LIGHT_DIFFUSE_AREA = 1, LIGHT_DIRECTIONAL = 2, LIGHT_IMAGE = 3, LIGHT_PEREZ = 4, LIGHT_POINT = 5,
              LIGHT_SPOT = 7;
// modifiers (M/scene/modifier/*Modifier.java)
constexpr int MODIFIER_NO_OP = 2;
// bounding-volume hierarchy nodes of a triangle mesh (M/geometry/
// boundingvolume/hierarchy/{Leaf,Tree}BVHNode3F.java)
constexpr int BVH_LEAF = 1, BVH_TREE = 2;

// CompiledPrimitiveCache.java:48-93
constexpr int PRIMITIVE_LENGTH = 8, MATRIX_PAIR_LENGTH = 32;
constexpr int PRIMITIVE_OFFSET_AREA_LIGHT_ID_AND_OFFSET = 1, PRIMITIVE_OFFSET_MATERIAL_ID = 4,
              PRIMITIVE_OFFSET_MATERIAL_OFFSET = 5, PRIMITIVE_OFFSET_SHAPE_ID = 6, PRIMITIVE_OFFSET_SHAPE_OFFSET = 7;
// CompiledShape3FCache.java:260-404
constexpr int RECTANGLE_OFFSET_POSITION = 0, RECTANGLE_OFFSET_SIDE_A = 3, RECTANGLE_OFFSET_SIDE_B = 6,
              RECTANGLE_LENGTH = 16;
constexpr int TRIANGLE_LENGTH = 24, TRIANGLE_OFFSET_POSITION = 0, TRIANGLE_OFFSET_TEXTURE_COORDINATES = 9,
              TRIANGLE_OFFSET_NORMAL = 15;
constexpr int MESH_NODE_OFFSET_ID = 0, MESH_NODE_OFFSET_NEXT = 2, MESH_NODE_OFFSET_LEFT_OR_COUNT = 3;
// CompiledMaterialCache.java:136-391 (ints; word 0 of every record packs
// the emission texture and the modifier)
constexpr int CLEAR_COAT_LENGTH = 2, DISNEY_LENGTH = 16, GLASS_LENGTH = 8, GLOSSY_LENGTH = 2, MATTE_LENGTH = 2,
              METAL_LENGTH = 4, MIRROR_LENGTH = 2, PLASTIC_LENGTH = 4, SUBSTRATE_LENGTH = 4, TRANSLUCENT_LENGTH = 4,
              UBER_LENGTH = 8;
// CompiledTextureCache.java
constexpr int BLEND_TEXTURE_LENGTH = 8, BULLSEYE_TEXTURE_LENGTH = 8, CHECKERBOARD_TEXTURE_LENGTH = 8,
              CONSTANT_TEXTURE_LENGTH = 4, POLKA_DOT_TEXTURE_LENGTH = 8;
constexpr int MARBLE_TEXTURE_LENGTH = 8, MARBLE_TEXTURE_OFFSET_COLOR_A = 0, MARBLE_TEXTURE_OFFSET_COLOR_B = 1,
              MARBLE_TEXTURE_OFFSET_COLOR_C = 2, MARBLE_TEXTURE_OFFSET_FREQUENCY = 3, MARBLE_TEXTURE_OFFSET_SCALE = 4,
              MARBLE_TEXTURE_OFFSET_OCTAVES = 6;
constexpr int SIMPLEX_FBM_TEXTURE_LENGTH = 4, SIMPLEX_FBM_TEXTURE_OFFSET_COLOR = 0, SIMPLEX_FBM_TEXTURE_OFFSET_FREQUENCY = 1,
              SIMPLEX_FBM_TEXTURE_OFFSET_GAIN = 2, SIMPLEX_FBM_TEXTURE_OFFSET_OCTAVES = 3;
constexpr int LDR_IMAGE_OFFSET_ANGLE_RADIANS = 0, LDR_IMAGE_OFFSET_SCALE = 1, LDR_IMAGE_OFFSET_RESOLUTION_X = 3,
              LDR_IMAGE_OFFSET_RESOLUTION_Y = 4, LDR_IMAGE_OFFSET_IMAGE = 5;
// CompiledLightCache.java:68-263
constexpr int DIFFUSE_AREA_LIGHT_LENGTH = 40, DIFFUSE_AREA_LIGHT_OFFSET_RADIANCE = 32,
              DIFFUSE_AREA_LIGHT_OFFSET_SHAPE_ID = 35, DIFFUSE_AREA_LIGHT_OFFSET_SHAPE_OFFSET = 36,
              DIFFUSE_AREA_LIGHT_OFFSET_IS_TWO_SIDED = 38;
constexpr int DIRECTIONAL_LIGHT_LENGTH = 8, DIRECTIONAL_LIGHT_OFFSET_RADIANCE = 0,
              DIRECTIONAL_LIGHT_OFFSET_DIRECTION = 3, DIRECTIONAL_LIGHT_OFFSET_RADIUS = 6;
constexpr int IMAGE_LIGHT_OFFSET_ANGLE_RADIANS = 32, IMAGE_LIGHT_OFFSET_SCALE = 33, IMAGE_LIGHT_OFFSET_RADIUS = 35,
              IMAGE_LIGHT_OFFSET_RESOLUTION_X = 36, IMAGE_LIGHT_OFFSET_RESOLUTION_Y = 37, IMAGE_LIGHT_OFFSET_IMAGE = 2352;
constexpr int PEREZ_LIGHT_OFFSET_SUN_DIRECTION_WORLD_SPACE = 38, PEREZ_LIGHT_OFFSET_TURBIDITY = 41;
constexpr int POINT_LIGHT_LENGTH = 8, POINT_LIGHT_OFFSET_INTENSITY = 0, POINT_LIGHT_OFFSET_POSITION = 3;
constexpr int SPOT_LIGHT_LENGTH = 24, SPOT_LIGHT_OFFSET_WORLD_TO_OBJECT = 0, SPOT_LIGHT_OFFSET_INTENSITY = 16,
              SPOT_LIGHT_OFFSET_POSITION = 19, SPOT_LIGHT_OFFSET_COS_CONE_ANGLE = 22,
              SPOT_LIGHT_OFFSET_COS_CONE_ANGLE_MINUS_DELTA = 23;
constexpr int CAMERA_LENGTH = 19;

struct Failure {
    int code;
    std::string msg;
};
[[noreturn]] void fail(int code, const std::string& msg) { throw Failure{code, msg}; }

const char* shape_name(int id) {
    switch (id) {
        case SHAPE_CONE: return "Cone3F";
        case SHAPE_CYLINDER: return "Cylinder3F";
        case SHAPE_DISK: return "Disk3F";
        case SHAPE_HYPERBOLOID: return "Hyperboloid3F";
        case SHAPE_PARABOLOID: return "Paraboloid3F";
        case SHAPE_POLYGON: return "Polygon3F";
        case SHAPE_RECTANGULAR_CUBOID: return "RectangularCuboid3F";
        case SHAPE_TORUS: return "Torus3F";
        default: return nullptr;
    }
}

// Bounds-checked views of one array.
struct FA {
    const float* p;
    int64_t n;
    const char* name;
    float at(int64_t i) const {
        if (!p || i < 0 || i >= n) fail(DFR_E_ARGUMENT, std::string(name) + ": element " + std::to_string(i) + " outside the array");
        return p[i];
    }
    void need(int64_t first, int64_t count) const { if (count > 0) { at(first); at(first + count - 1); } }
};
struct IA {
    const int32_t* p;
    int64_t n;
    const char* name;
    int32_t at(int64_t i) const {
        if (!p || i < 0 || i >= n) fail(DFR_E_ARGUMENT, std::string(name) + ": element " + std::to_string(i) + " outside the array");
        return p[i];
    }
};
FA fa(const dfr_farray& a, const char* name) { return FA{a.data, a.data ? a.length : 0, name}; }
IA ia(const dfr_iarray& a, const char* name) { return IA{a.data, a.data ? a.length : 0, name}; }

// Ints.pack(a, b) -> (a, b);  Ints.pack(a, b, c, d) -> the pair k (0 or 1)
void unpack16(int32_t v, int& id, int& off) { id = (int)(((uint32_t)v >> 16) & 0xffffu); off = (int)((uint32_t)v & 0xffffu); }
void unpack8(int32_t v, int k, int& id, int& off) {
    const uint32_t u = (uint32_t)v;
    id = (int)((u >> (16 * k)) & 0xffu);
    off = (int)((u >> (16 * k + 8)) & 0xffu);
}

struct Out {
    std::vector<float> positions, normals, uvs;
    std::vector<int32_t> indices;
    std::vector<dfr_primitive> prims;
    std::vector<dfr_material> mats;
    std::vector<dfr_light> lights;
    std::vector<dfr_texture> textures;
    std::vector<std::vector<float>> image_rgb;
    std::vector<dfr_image> images;
    std::vector<float> xforms;  // 16 per primitive
    float camera[DFR_CAMERA_LENGTH];
};

class Transcoder {
public:
    Transcoder(const dfr_compiled_scene& s, Out& o) : s_(s), o_(o) {}

    void run() {
        const FA cam = fa(s_.camera, "camera");
        cam.need(0, CAMERA_LENGTH);
        for (int k = 0; k < CAMERA_LENGTH; k++) o_.camera[k] = cam.at(k);
        if (s_.primitive_count < 0 || s_.light_count < 0) fail(DFR_E_ARGUMENT, "negative primitive or light count");
        lights_first();
        for (int32_t i = 0; i < s_.primitive_count; i++) primitive(i);
        area_lights();
    }

private:
    const dfr_compiled_scene& s_;
    Out& o_;
    std::map<std::pair<int, int>, int> tex_memo_, mat_memo_;
    std::map<int64_t, std::pair<int64_t, int64_t>> mesh_memo_;  // mesh offset -> (first triangle, count)
    std::map<int, int> area_light_index_;   // DiffuseAreaLight offset -> IR light index
    std::vector<int> area_light_prim_;      // per IR light: the primitive that carries it (-1 none)
    std::vector<std::pair<int, int>> area_light_shape_;  // per IR light: (shape ID, shape offset)

    // -- textures ------------------------------------------------------------
    // A constant texture's colour, for a parameter the kernels read as a value.
    void constant_color(int id, int off, float* rgb, const char* what) {
        if (id != TEX_CONSTANT)
            fail(DFR_E_UNSUPPORTED, std::string(what) + ": a texture of type " + std::to_string(id) +
                                        " where the kernels read a constant (only the colour slots of dfr_material "
                                        "take textures)");
        const FA a = fa(s_.constant_textures, "constant_textures");
        const int64_t b = (int64_t)off * CONSTANT_TEXTURE_LENGTH;
        for (int k = 0; k < 3; k++) rgb[k] = a.at(b + k);
    }
    // Texture.getFloat: the average of the colour (textureEvaluateFloat,
    // M/renderer/gpu/AbstractTextureKernel.java:133-137)
    float constant_float(int id, int off, const char* what) {
        float c[3];
        constant_color(id, off, c, what);
        volatile float s = c[0] + c[1];
        s = s + c[2];
        return s / 3.0f;
    }
    // a colour packed as 0xAARRGGBB int bits stored in a float: byte / 255
    // (colorRGBIntToRFloat and kin)
    static void argb(float packed, float* rgb) {
        const uint32_t c = (uint32_t)(int32_t)packed;
        rgb[0] = (float)((c >> 16) & 0xffu) / 255.0f;
        rgb[1] = (float)((c >> 8) & 0xffu) / 255.0f;
        rgb[2] = (float)(c & 0xffu) / 255.0f;
    }
    static float degrees(float radians) { return (float)((double)radians * (180.0 / 3.14159265358979323846)); }

    int texture(int id, int off, int depth = 0) {
        if (depth > 64) fail(DFR_E_ARGUMENT, "texture references nested more than 64 deep (a cycle?)");
        auto key = std::make_pair(id, off);
        auto it = tex_memo_.find(key);
        if (it != tex_memo_.end()) return it->second;
        dfr_texture t;
        std::memset(&t, 0, sizeof(t));
        t.a = t.b = t.image = -1;
        t.scale[0] = t.scale[1] = 1.0f;
        auto sub = [&](float packed, int& dst) {
            const int32_t v = (int32_t)packed;
            int sid, soff;
            unpack16(v, sid, soff);
            dst = texture(sid, soff, depth + 1);
        };
        switch (id) {
            case TEX_CONSTANT: {  // CompiledTextureCache.java:1800-1812
                t.type = DFR_TEX_CONSTANT;
                constant_color(id, off, t.color, "texture");
                break;
            }
            case TEX_CHECKERBOARD: {  // :1731-1755 (angle degrees, radians, A, B, scale u, v)
                const FA a = fa(s_.checkerboard_textures, "checkerboard_textures");
                const int64_t b = (int64_t)off * CHECKERBOARD_TEXTURE_LENGTH;
                a.need(b, CHECKERBOARD_TEXTURE_LENGTH);
                t.type = DFR_TEX_CHECKERBOARD;
                t.angle = a.at(b + 0);
                sub(a.at(b + 2), t.a);
                sub(a.at(b + 3), t.b);
                t.scale[0] = a.at(b + 4);
                t.scale[1] = a.at(b + 5);
                break;
            }
            case TEX_BLEND: {  // :1551-1574 (A, B, t per component)
                const FA a = fa(s_.blend_textures, "blend_textures");
                const int64_t b = (int64_t)off * BLEND_TEXTURE_LENGTH;
                a.need(b, BLEND_TEXTURE_LENGTH);
                t.type = DFR_TEX_BLEND;
                sub(a.at(b + 0), t.a);
                sub(a.at(b + 1), t.b);
                for (int k = 0; k < 3; k++) t.t[k] = a.at(b + 2 + k);
                break;
            }
            case TEX_BULLSEYE: {  // :1641-1664 (origin, A, B, scale)
                const FA a = fa(s_.bullseye_textures, "bullseye_textures");
                const int64_t b = (int64_t)off * BULLSEYE_TEXTURE_LENGTH;
                a.need(b, BULLSEYE_TEXTURE_LENGTH);
                t.type = DFR_TEX_BULLSEYE;
                for (int k = 0; k < 3; k++) t.origin[k] = a.at(b + k);
                sub(a.at(b + 3), t.a);
                sub(a.at(b + 4), t.b);
                t.scale[0] = t.scale[1] = a.at(b + 5);
                break;
            }
            case TEX_POLKA_DOT: {  // :1956-1980 (angle degrees, radians, A, B, cell, radius)
                const FA a = fa(s_.polka_dot_textures, "polka_dot_textures");
                const int64_t b = (int64_t)off * POLKA_DOT_TEXTURE_LENGTH;
                a.need(b, POLKA_DOT_TEXTURE_LENGTH);
                t.type = DFR_TEX_POLKA_DOT;
                t.angle = a.at(b + 0);
                sub(a.at(b + 2), t.a);
                sub(a.at(b + 3), t.b);
                t.cell = a.at(b + 4);
                t.radius = a.at(b + 5);
                break;
            }
            case TEX_UV: t.type = DFR_TEX_UV; break;
            case TEX_SURFACE_NORMAL: t.type = DFR_TEX_SURFACE_NORMAL; break;
            case TEX_DOT_PRODUCT: t.type = DFR_TEX_DOT_PRODUCT; break;
            case TEX_LDR_IMAGE: {  // :1836-1862; packed 0xAARRGGBB ints stored as floats
                const IA offs = ia(s_.ldr_image_texture_offsets, "ldr_image_texture_offsets");
                const FA a = fa(s_.ldr_image_textures, "ldr_image_textures");
                const int64_t b = offs.at(off);
                const int w = (int)a.at(b + LDR_IMAGE_OFFSET_RESOLUTION_X), h = (int)a.at(b + LDR_IMAGE_OFFSET_RESOLUTION_Y);
                if (w < 1 || h < 1) fail(DFR_E_ARGUMENT, "LDR image texture with an empty image");
                a.need(b + LDR_IMAGE_OFFSET_IMAGE, (int64_t)w * h);
                std::vector<float> rgb(3 * (size_t)w * h);
                for (int64_t i = 0; i < (int64_t)w * h; i++) {
                    const uint32_t c = (uint32_t)(int32_t)a.at(b + LDR_IMAGE_OFFSET_IMAGE + i);
                    // colorRGBIntToRFloat and kin: byte / 255
                    rgb[3 * i + 0] = (float)((c >> 16) & 0xffu) / 255.0f;
                    rgb[3 * i + 1] = (float)((c >> 8) & 0xffu) / 255.0f;
                    rgb[3 * i + 2] = (float)(c & 0xffu) / 255.0f;
                }
                t.type = DFR_TEX_IMAGE;
                t.image = add_image(w, h, std::move(rgb));
                t.angle = degrees(a.at(b + LDR_IMAGE_OFFSET_ANGLE_RADIANS));
                t.scale[0] = a.at(b + LDR_IMAGE_OFFSET_SCALE);
                t.scale[1] = a.at(b + LDR_IMAGE_OFFSET_SCALE + 1);
                break;
            }
            case TEX_MARBLE: {  // :1885-1909 (colours A, B, C packed ARGB, frequency, scale, stripes, octaves)
                const FA a = fa(s_.marble_textures, "marble_textures");
                const int64_t b = (int64_t)off * MARBLE_TEXTURE_LENGTH;
                a.need(b, MARBLE_TEXTURE_LENGTH);
                t.type = DFR_TEX_MARBLE;
                argb(a.at(b + MARBLE_TEXTURE_OFFSET_COLOR_A), t.color);
                argb(a.at(b + MARBLE_TEXTURE_OFFSET_COLOR_B), t.origin);
                argb(a.at(b + MARBLE_TEXTURE_OFFSET_COLOR_C), t.t);
                t.scale[0] = a.at(b + MARBLE_TEXTURE_OFFSET_FREQUENCY);
                t.scale[1] = a.at(b + MARBLE_TEXTURE_OFFSET_SCALE);
                t.cell = a.at(b + MARBLE_TEXTURE_OFFSET_OCTAVES);
                break;
            }
            case TEX_SIMPLEX_FBM: {  // :2025-2043 (colour packed ARGB, frequency, gain, octaves)
                const FA a = fa(s_.simplex_fbm_textures, "simplex_fbm_textures");
                const int64_t b = (int64_t)off * SIMPLEX_FBM_TEXTURE_LENGTH;
                a.need(b, SIMPLEX_FBM_TEXTURE_LENGTH);
                t.type = DFR_TEX_SIMPLEX_FBM;
                argb(a.at(b + SIMPLEX_FBM_TEXTURE_OFFSET_COLOR), t.color);
                t.scale[0] = a.at(b + SIMPLEX_FBM_TEXTURE_OFFSET_FREQUENCY);
                t.scale[1] = a.at(b + SIMPLEX_FBM_TEXTURE_OFFSET_GAIN);
                t.cell = a.at(b + SIMPLEX_FBM_TEXTURE_OFFSET_OCTAVES);
                break;
            }
            case TEX_FUNCTION: fail(DFR_E_UNSUPPORTED, "FunctionTexture has no compiled form");
            default: fail(DFR_E_ARGUMENT, "unknown texture type ID " + std::to_string(id));
        }
        const int idx = (int)o_.textures.size();
        o_.textures.push_back(t);
        tex_memo_[key] = idx;
        return idx;
    }

    int add_image(int w, int h, std::vector<float> rgb) {
        o_.image_rgb.push_back(std::move(rgb));
        dfr_image im;
        im.width = w;
        im.height = h;
        im.rgb = nullptr;  // pointed at o_.image_rgb once the vector stops growing
        o_.images.push_back(im);
        return (int)o_.images.size() - 1;
    }

    // A colour slot: a constant fills the colour, anything else becomes a
    // texture of the slot.
    void color_slot(dfr_material& m, int slot, float* color, int id, int off) {
        if (id == TEX_CONSTANT) constant_color(id, off, color, "colour");
        else { m.textures[slot] = texture(id, off) + 1; color[0] = color[1] = color[2] = 0.0f; }
    }
    void value_color(float* color, int id, int off, const char* what) { constant_color(id, off, color, what); }

    // -- materials -------------------------------------------------------------
    int material(int id, int off) {
        auto key = std::make_pair(id, off);
        auto it = mat_memo_.find(key);
        if (it != mat_memo_.end()) return it->second;
        dfr_material m;
        std::memset(&m, 0, sizeof(m));
        const char* name = "material";
        auto rec = [&](const dfr_iarray& arr, const char* an, int len) {
            const IA a = ia(arr, an);
            std::vector<int32_t> w(len);
            for (int k = 0; k < len; k++) w[k] = a.at((int64_t)off + k);
            return w;
        };
        auto pair = [&](int32_t word, int k, int& tid, int& toff) { unpack8(word, k, tid, toff); };
        int tid, toff;
        auto emission_and_modifier = [&](int32_t word) {
            pair(word, 0, tid, toff);
            // textureEmission: a constant fills the colour, anything else is
            // the material's emission texture
            if (tid == TEX_CONSTANT) value_color(m.emission, tid, toff, "emission");
            else { m.emission_texture = texture(tid, toff) + 1; m.emission[0] = m.emission[1] = m.emission[2] = 0.0f; }
            int mid, moff;
            pair(word, 1, mid, moff);
            if (mid != MODIFIER_NO_OP && mid != 0)
                fail(DFR_E_UNSUPPORTED, "material modifier of type " + std::to_string(mid) +
                                            " (normal and parallax maps have no counterpart in the kernels)");
        };
        auto fval = [&](int32_t word, int k, const char* what) {
            pair(word, k, tid, toff);
            return constant_float(tid, toff, what);
        };
        // a roughness: a constant gives the value, anything else becomes the
        // slot's texture (its colour average, read at each hit)
        auto float_slot = [&](dfr_material& mm, int slot, int32_t word, int k, const char* what) {
            pair(word, k, tid, toff);
            if (tid == TEX_CONSTANT) return constant_float(tid, toff, what);
            mm.textures[slot] = texture(tid, toff) + 1;
            return 0.0f;
        };
        switch (id) {
            case MAT_MATTE: {  // CompiledMaterialCache.java:3010-3027: (emission, modifier) (angle, KD)
                auto w = rec(s_.matte_materials, "matte_materials", MATTE_LENGTH);
                m.type = DFR_MAT_MATTE;
                emission_and_modifier(w[0]);
                m.angle = fval(w[1], 0, "matte angle");
                pair(w[1], 1, tid, toff);
                color_slot(m, DFR_SLOT_KD, m.kd, tid, toff);
                break;
            }
            case MAT_MIRROR: {  // :3187-3203: (emission, modifier) (KR)
                auto w = rec(s_.mirror_materials, "mirror_materials", MIRROR_LENGTH);
                m.type = DFR_MAT_MIRROR;
                emission_and_modifier(w[0]);
                pair(w[1], 0, tid, toff);
                color_slot(m, DFR_SLOT_KR, m.kr, tid, toff);
                break;
            }
            case MAT_GLASS: {  // :2830-2857: (emission, modifier) (eta) (KR, KT) (roughness U, V) remap
                auto w = rec(s_.glass_materials, "glass_materials", GLASS_LENGTH);
                m.type = DFR_MAT_GLASS;
                emission_and_modifier(w[0]);
                m.eta[0] = float_slot(m, DFR_SLOT_ETA, w[1], 0, "glass eta");
                pair(w[2], 0, tid, toff);
                color_slot(m, DFR_SLOT_KR, m.kr, tid, toff);
                pair(w[2], 1, tid, toff);
                color_slot(m, DFR_SLOT_KT, m.kt, tid, toff);
                m.roughness_u = fval(w[3], 0, "glass roughness U");
                m.roughness_v = fval(w[3], 1, "glass roughness V");
                m.remap_roughness = w[4] != 0;
                break;
            }
            case MAT_METAL: {  // :3095-3119: (emission, modifier) (eta, K) (roughness U, V) remap
                auto w = rec(s_.metal_materials, "metal_materials", METAL_LENGTH);
                m.type = DFR_MAT_METAL;
                emission_and_modifier(w[0]);
                pair(w[1], 0, tid, toff);
                color_slot(m, DFR_SLOT_ETA, m.eta, tid, toff);
                pair(w[1], 1, tid, toff);
                color_slot(m, DFR_SLOT_K, m.k, tid, toff);
                m.roughness_u = float_slot(m, DFR_SLOT_ROUGHNESS_U, w[2], 0, "metal roughness U");
                m.roughness_v = float_slot(m, DFR_SLOT_ROUGHNESS_V, w[2], 1, "metal roughness V");
                m.remap_roughness = w[3] != 0;
                break;
            }
            case MAT_PLASTIC: {  // :3271-3294: (emission, modifier) (KD, KS) (roughness) remap
                auto w = rec(s_.plastic_materials, "plastic_materials", PLASTIC_LENGTH);
                m.type = DFR_MAT_PLASTIC;
                emission_and_modifier(w[0]);
                pair(w[1], 0, tid, toff);
                color_slot(m, DFR_SLOT_KD, m.kd, tid, toff);
                pair(w[1], 1, tid, toff);
                color_slot(m, DFR_SLOT_KS, m.ks, tid, toff);
                m.roughness_u = m.roughness_v = float_slot(m, DFR_SLOT_ROUGHNESS_U, w[2], 0, "plastic roughness");
                m.remap_roughness = w[3] != 0;
                break;
            }
            case MAT_DISNEY: {  // :2717-2762
                auto w = rec(s_.disney_materials, "disney_materials", DISNEY_LENGTH);
                m.type = DFR_MAT_DISNEY;
                emission_and_modifier(w[0]);
                m.disney_anisotropic = fval(w[1], 0, "disney anisotropic");
                m.disney_clearcoat = fval(w[2], 0, "disney clear coat");
                m.disney_clearcoat_gloss = fval(w[2], 1, "disney clear coat gloss");
                pair(w[3], 0, tid, toff);
                value_color(m.kd, tid, toff, "disney colour");
                m.disney_diff_trans = fval(w[3], 1, "disney diffuse transmission");
                m.eta[0] = fval(w[4], 0, "disney eta");
                m.disney_flatness = fval(w[4], 1, "disney flatness");
                m.disney_metallic = fval(w[5], 0, "disney metallic");
                m.disney_roughness = fval(w[5], 1, "disney roughness");
                m.disney_sheen = fval(w[6], 1, "disney sheen");
                m.disney_sheen_tint = fval(w[7], 0, "disney sheen tint");
                m.disney_specular_tint = fval(w[7], 1, "disney specular tint");
                m.disney_spec_trans = fval(w[8], 0, "disney specular transmission");
                m.thin = (int32_t)(((uint32_t)w[8] >> 16) & 0xffu);
                if (m.thin) fail(DFR_E_UNSUPPORTED, "thin Disney materials are not supported");
                break;
            }
            case MAT_GLOSSY: {  // :2925-2942: (emission, modifier) (KR, roughness)
                auto w = rec(s_.glossy_materials, "glossy_materials", GLOSSY_LENGTH);
                m.type = DFR_MAT_GLOSSY;
                emission_and_modifier(w[0]);
                pair(w[1], 0, tid, toff);
                color_slot(m, DFR_SLOT_KR, m.kr, tid, toff);
                m.roughness_u = m.roughness_v = float_slot(m, DFR_SLOT_ROUGHNESS_U, w[1], 1, "glossy roughness");
                break;
            }
            case MAT_SUBSTRATE: {  // :3362-3385: (emission, modifier) (KD, KS) (roughness U, V) remap
                auto w = rec(s_.substrate_materials, "substrate_materials", SUBSTRATE_LENGTH);
                m.type = DFR_MAT_SUBSTRATE;
                emission_and_modifier(w[0]);
                pair(w[1], 0, tid, toff);
                color_slot(m, DFR_SLOT_KD, m.kd, tid, toff);
                pair(w[1], 1, tid, toff);
                color_slot(m, DFR_SLOT_KS, m.ks, tid, toff);
                m.roughness_u = float_slot(m, DFR_SLOT_ROUGHNESS_U, w[2], 0, "substrate roughness U");
                m.roughness_v = float_slot(m, DFR_SLOT_ROUGHNESS_V, w[2], 1, "substrate roughness V");
                m.remap_roughness = w[3] != 0;
                break;
            }
            case MAT_TRANSLUCENT: {  // :3456-3480: (emission, modifier) (KD, KS) (reflectance, transmittance) (roughness, remap)
                auto w = rec(s_.translucent_materials, "translucent_materials", TRANSLUCENT_LENGTH);
                m.type = DFR_MAT_TRANSLUCENT;
                emission_and_modifier(w[0]);
                pair(w[1], 0, tid, toff);
                value_color(m.kd, tid, toff, "translucent KD");
                pair(w[1], 1, tid, toff);
                value_color(m.ks, tid, toff, "translucent KS");
                pair(w[2], 0, tid, toff);
                value_color(m.kr, tid, toff, "translucent reflectance");
                pair(w[2], 1, tid, toff);
                value_color(m.kt, tid, toff, "translucent transmittance");
                m.roughness_u = m.roughness_v = fval(w[3], 0, "translucent roughness");
                m.remap_roughness = (((uint32_t)w[3] >> 16) & 0xffu) != 0;
                break;
            }
            case MAT_CLEAR_COAT: {  // :2632-2649: (emission, modifier) (KD, KS)
                auto w = rec(s_.clear_coat_materials, "clear_coat_materials", CLEAR_COAT_LENGTH);
                m.type = DFR_MAT_CLEAR_COAT;
                emission_and_modifier(w[0]);
                pair(w[1], 0, tid, toff);
                color_slot(m, DFR_SLOT_KD, m.kd, tid, toff);
                pair(w[1], 1, tid, toff);
                color_slot(m, DFR_SLOT_KS, m.ks, tid, toff);
                break;
            }
            case MAT_UBER: {  // :3550-3580: (emission, modifier) (KD, KR) (KS, KT) (eta, opacity) (roughness U, V) remap
                auto w = rec(s_.uber_materials, "uber_materials", UBER_LENGTH);
                m.type = DFR_MAT_UBER;
                emission_and_modifier(w[0]);
                pair(w[1], 0, tid, toff);
                value_color(m.kd, tid, toff, "uber KD");
                pair(w[1], 1, tid, toff);
                value_color(m.kr, tid, toff, "uber KR");
                pair(w[2], 0, tid, toff);
                value_color(m.ks, tid, toff, "uber KS");
                pair(w[2], 1, tid, toff);
                value_color(m.kt, tid, toff, "uber KT");
                m.eta[0] = fval(w[3], 0, "uber eta");
                pair(w[3], 1, tid, toff);
                value_color(m.k, tid, toff, "uber opacity");
                m.roughness_u = fval(w[4], 0, "uber roughness U");
                m.roughness_v = fval(w[4], 1, "uber roughness V");
                m.remap_roughness = w[5] != 0;
                break;
            }
            case MAT_BULLSEYE: name = "BullseyeMaterial"; fail(DFR_E_UNSUPPORTED, std::string(name) + " has no counterpart in the kernels");
            case MAT_CHECKERBOARD: name = "CheckerboardMaterial"; fail(DFR_E_UNSUPPORTED, std::string(name) + " has no counterpart in the kernels");
            case MAT_POLKA_DOT: name = "PolkaDotMaterial"; fail(DFR_E_UNSUPPORTED, std::string(name) + " has no counterpart in the kernels");
            default: fail(DFR_E_ARGUMENT, "unknown material type ID " + std::to_string(id));
        }
        const int idx = (int)o_.mats.size();
        o_.mats.push_back(m);
        mat_memo_[key] = idx;
        return idx;
    }

    // -- shapes and primitives ---------------------------------------------------
    // One Triangle3F record (CompiledShape3FCache.java:315-360): positions,
    // texture coordinates and vertex normals (the orthonormal bases' w) of
    // A, B, C, appended as three vertices of their own.
    void triangle(int64_t off) {
        const FA a = fa(s_.triangles, "triangles");
        a.need(off, TRIANGLE_LENGTH);
        const int32_t base = (int32_t)(o_.positions.size() / 3);
        for (int v = 0; v < 3; v++) {
            for (int k = 0; k < 3; k++) o_.positions.push_back(a.at(off + TRIANGLE_OFFSET_POSITION + 3 * v + k));
            for (int k = 0; k < 3; k++) o_.normals.push_back(a.at(off + TRIANGLE_OFFSET_NORMAL + 3 * v + k));
            for (int k = 0; k < 2; k++) o_.uvs.push_back(a.at(off + TRIANGLE_OFFSET_TEXTURE_COORDINATES + 2 * v + k));
        }
        for (int v = 0; v < 3; v++) o_.indices.push_back(base + v);
    }

    // Every leaf of a mesh's threaded BVH, in node order: follow `left` from a
    // tree node and `next` from a leaf until -1, as the kernels' walk does
    // when every box is entered (AbstractShape3FKernel.java:3693-3722; node
    // layout CompiledShape3FCache.java:2652-2733).
    int64_t mesh(int64_t off) {
        const IA a = ia(s_.triangle_meshes, "triangle_meshes");
        int64_t count = 0, steps = 0;
        int64_t rel = 0;
        while (rel != -1) {
            if (++steps > a.n) fail(DFR_E_ARGUMENT, "triangle_meshes: the node links of a mesh do not end");
            const int64_t node = off + rel;
            const int id = a.at(node + MESH_NODE_OFFSET_ID);
            const int32_t next = a.at(node + MESH_NODE_OFFSET_NEXT);
            const int32_t third = a.at(node + MESH_NODE_OFFSET_LEFT_OR_COUNT);
            if (id == BVH_LEAF) {
                for (int32_t k = 0; k < third; k++) {
                    const int32_t t = a.at(node + MESH_NODE_OFFSET_LEFT_OR_COUNT + 1 + k);
                    triangle(t);
                    count++;
                }
                rel = next;
            } else if (id == BVH_TREE) {
                rel = third;
            } else {
                fail(DFR_E_ARGUMENT, "triangle_meshes: unknown node ID " + std::to_string(id));
            }
            if (rel < -1) fail(DFR_E_ARGUMENT, "triangle_meshes: negative node offset");
        }
        return count;
    }

    // Rectangles under a matrix of positive determinant, and planes under a
    // rotation, go to world space in their records: the kernels' texture
    // coordinates on them are then the object-space ones (a rectangle's are
    // affine coordinates in its corners, a plane's the projections on its
    // unit u and v), and the matrix becomes the identity.  Other maps stay
    // matrices.
    static void bake(dfr_primitive& p, float* m) {
        if (p.shape != DFR_SHAPE_RECTANGLE && p.shape != DFR_SHAPE_PLANE) return;
        double a[3][3], t[3];
        for (int r = 0; r < 3; r++) {
            for (int c = 0; c < 3; c++) a[r][c] = m[4 * r + c];
            t[r] = m[4 * r + 3];
        }
        if (m[12] != 0.0f || m[13] != 0.0f || m[14] != 0.0f || m[15] != 1.0f) return;
        const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                           a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                           a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
        if (!(det > 0.0)) return;
        auto lin = [&](const float* v, float* out) {
            for (int r = 0; r < 3; r++) out[r] = (float)(a[r][0] * v[0] + a[r][1] * v[1] + a[r][2] * v[2]);
        };
        auto pt = [&](const float* v, float* out) {
            for (int r = 0; r < 3; r++) out[r] = (float)(a[r][0] * v[0] + a[r][1] * v[1] + a[r][2] * v[2] + t[r]);
        };
        if (p.shape == DFR_SHAPE_PLANE) {
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) {
                    double dotp = 0.0;
                    for (int k = 0; k < 3; k++) dotp += a[k][r] * a[k][c];
                    if (std::fabs(dotp - (r == c ? 1.0 : 0.0)) > 1e-6) return;   // not a rotation
                }
            float d[12];
            std::memcpy(d, p.data, sizeof(d));
            pt(d, p.data);
            for (int k = 1; k < 4; k++) lin(d + 3 * k, p.data + 3 * k);
        } else {
            float d[9];
            std::memcpy(d, p.data, sizeof(d));
            for (int k = 0; k < 3; k++) pt(d + 3 * k, p.data + 3 * k);
        }
        for (int k = 0; k < 16; k++) m[k] = (k % 5 == 0) ? 1.0f : 0.0f;
    }

    void primitive(int32_t i) {
        const IA P = ia(s_.primitives, "primitives");
        const FA M = fa(s_.primitive_matrices, "primitive_matrices");
        const int64_t b = (int64_t)i * PRIMITIVE_LENGTH;
        const int shape_id = P.at(b + PRIMITIVE_OFFSET_SHAPE_ID);
        const int64_t shape_off = P.at(b + PRIMITIVE_OFFSET_SHAPE_OFFSET);
        dfr_primitive p;
        std::memset(&p, 0, sizeof(p));
        p.material = material(P.at(b + PRIMITIVE_OFFSET_MATERIAL_ID), P.at(b + PRIMITIVE_OFFSET_MATERIAL_OFFSET));
        p.area_light = -1;
        float m[16];
        for (int k = 0; k < 16; k++) m[k] = M.at((int64_t)i * MATRIX_PAIR_LENGTH + k);
        switch (shape_id) {
            case SHAPE_TRIANGLE:
                p.shape = DFR_SHAPE_TRIANGLE_MESH;
                p.first_triangle = (int64_t)o_.indices.size() / 3;
                triangle(shape_off);
                p.triangle_count = 1;
                break;
            case SHAPE_TRIANGLE_MESH: {
                // the compiler stores a mesh once however many primitives
                // name it (Shape3FCache.java:176-177, 298-300): those
                // primitives share its triangle range and become instances
                p.shape = DFR_SHAPE_TRIANGLE_MESH;
                auto it = mesh_memo_.find(shape_off);
                if (it != mesh_memo_.end()) {
                    p.first_triangle = it->second.first;
                    p.triangle_count = it->second.second;
                } else {
                    p.first_triangle = (int64_t)o_.indices.size() / 3;
                    p.triangle_count = mesh(shape_off);
                    mesh_memo_[shape_off] = {p.first_triangle, p.triangle_count};
                }
                break;
            }
            case SHAPE_SPHERE: {
                // the kernels' unit sphere at the origin (AbstractShape3FKernel
                // .java:2025-2038); a uniform scale s with no rotation is folded
                // into the radius, so a sphere light keeps a translation only
                p.shape = DFR_SHAPE_SPHERE;
                p.data[3] = 1.0f;
                const double sx = m[0], sy = m[5], sz = m[10];
                const bool axis_scale = m[1] == 0.0f && m[2] == 0.0f && m[4] == 0.0f && m[6] == 0.0f && m[8] == 0.0f &&
                                        m[9] == 0.0f && sx == sy && sy == sz && sx > 0.0;
                if (axis_scale) {
                    p.data[0] = m[3]; p.data[1] = m[7]; p.data[2] = m[11]; p.data[3] = m[0];
                    for (int k = 0; k < 16; k++) m[k] = (k % 5 == 0) ? 1.0f : 0.0f;
                    m[15] = 1.0f;
                }
                break;
            }
            case SHAPE_PLANE:
                // Plane3F: z = 0 with u = x, v = y, w = z (Plane3F.java:283)
                p.shape = DFR_SHAPE_PLANE;
                p.data[3] = 1.0f; p.data[7] = 1.0f; p.data[11] = 1.0f;
                break;
            case SHAPE_RECTANGLE: {
                // position A, side A = B - A, side B = C - B (CompiledShape3FCache.java:2280-2304)
                const FA a = fa(s_.rectangles, "rectangles");
                a.need(shape_off, RECTANGLE_LENGTH);
                p.shape = DFR_SHAPE_RECTANGLE;
                for (int k = 0; k < 3; k++) {
                    const float A = a.at(shape_off + RECTANGLE_OFFSET_POSITION + k);
                    const float B = A + a.at(shape_off + RECTANGLE_OFFSET_SIDE_A + k);
                    const float C = B + a.at(shape_off + RECTANGLE_OFFSET_SIDE_B + k);
                    p.data[k] = A; p.data[3 + k] = B; p.data[6 + k] = C;
                }
                break;
            }
            default: {
                const char* n = shape_name(shape_id);
                if (n) fail(DFR_E_UNSUPPORTED, std::string("primitive ") + std::to_string(i) + ": " + n +
                                                   " has no counterpart in the kernels");
                fail(DFR_E_ARGUMENT, "primitive " + std::to_string(i) + ": unknown shape type ID " + std::to_string(shape_id));
            }
        }
        bake(p, m);
        // the primitive's area light (Ints.pack(ID, offset); 0 when none)
        int lid, loff;
        unpack16(P.at(b + PRIMITIVE_OFFSET_AREA_LIGHT_ID_AND_OFFSET), lid, loff);
        if (lid == LIGHT_DIFFUSE_AREA) {
            auto it = area_light_index_.find(loff);
            if (it == area_light_index_.end())
                fail(DFR_E_ARGUMENT, "primitive " + std::to_string(i) + ": its area light is not in the light table");
            const int li = it->second;
            if (area_light_prim_[li] >= 0)
                fail(DFR_E_UNSUPPORTED, "an area light carried by two primitives");
            area_light_prim_[li] = (int)o_.prims.size();
            p.area_light = li;
        } else if (lid != 0) {
            fail(DFR_E_ARGUMENT, "primitive " + std::to_string(i) + ": area light of type " + std::to_string(lid));
        }
        o_.prims.push_back(p);
        o_.xforms.insert(o_.xforms.end(), m, m + 16);
    }

    // -- lights ------------------------------------------------------------------
    void lights_first() {
        const IA T = ia(s_.light_ids_and_offsets, "light_ids_and_offsets");
        for (int32_t l = 0; l < s_.light_count; l++) {
            int id, off;
            unpack16(T.at(l), id, off);
            dfr_light L;
            std::memset(&L, 0, sizeof(L));
            L.primitive = -1;
            L.triangle = -1;
            switch (id) {
                case LIGHT_POINT: {  // CompiledLightCache.java:1600-1620
                    const FA a = fa(s_.point_lights, "point_lights");
                    const int64_t b = (int64_t)off * POINT_LIGHT_LENGTH;
                    a.need(b, POINT_LIGHT_LENGTH);
                    L.type = DFR_LIGHT_POINT;
                    for (int k = 0; k < 3; k++) {
                        L.radiance[k] = a.at(b + POINT_LIGHT_OFFSET_INTENSITY + k);
                        L.position[k] = a.at(b + POINT_LIGHT_OFFSET_POSITION + k);
                    }
                    break;
                }
                case LIGHT_SPOT: {  // :1644-1690; the cone axis is the object-space +z
                    const FA a = fa(s_.spot_lights, "spot_lights");
                    const int64_t b = (int64_t)off * SPOT_LIGHT_LENGTH;
                    a.need(b, SPOT_LIGHT_LENGTH);
                    L.type = DFR_LIGHT_SPOT;
                    double w2o[3][3];
                    for (int r = 0; r < 3; r++)
                        for (int c = 0; c < 3; c++) w2o[r][c] = a.at(b + SPOT_LIGHT_OFFSET_WORLD_TO_OBJECT + 4 * r + c);
                    // object-to-world z axis: third column of the inverse
                    const double det = w2o[0][0] * (w2o[1][1] * w2o[2][2] - w2o[1][2] * w2o[2][1]) -
                                       w2o[0][1] * (w2o[1][0] * w2o[2][2] - w2o[1][2] * w2o[2][0]) +
                                       w2o[0][2] * (w2o[1][0] * w2o[2][1] - w2o[1][1] * w2o[2][0]);
                    if (!(std::fabs(det) > 0.0)) fail(DFR_E_ARGUMENT, "spot light with a singular transform");
                    const double ax = (w2o[0][1] * w2o[1][2] - w2o[0][2] * w2o[1][1]) / det;
                    const double ay = -(w2o[0][0] * w2o[1][2] - w2o[0][2] * w2o[1][0]) / det;
                    const double az = (w2o[0][0] * w2o[1][1] - w2o[0][1] * w2o[1][0]) / det;
                    L.direction[0] = (float)ax; L.direction[1] = (float)ay; L.direction[2] = (float)az;
                    for (int k = 0; k < 3; k++) {
                        L.radiance[k] = a.at(b + SPOT_LIGHT_OFFSET_INTENSITY + k);
                        L.position[k] = a.at(b + SPOT_LIGHT_OFFSET_POSITION + k);
                    }
                    const double c1 = a.at(b + SPOT_LIGHT_OFFSET_COS_CONE_ANGLE);
                    const double c2 = a.at(b + SPOT_LIGHT_OFFSET_COS_CONE_ANGLE_MINUS_DELTA);
                    const double deg = 180.0 / 3.14159265358979323846;
                    const double cone = std::acos(std::fmax(-1.0, std::fmin(1.0, c1))) * deg;
                    const double inner = std::acos(std::fmax(-1.0, std::fmin(1.0, c2))) * deg;
                    L.cone_angle = (float)cone;
                    L.cone_delta = (float)(cone - inner);
                    break;
                }
                case LIGHT_DIRECTIONAL: {  // :1349-1369 (direction towards the light, world space)
                    const FA a = fa(s_.directional_lights, "directional_lights");
                    const int64_t b = (int64_t)off * DIRECTIONAL_LIGHT_LENGTH;
                    a.need(b, DIRECTIONAL_LIGHT_LENGTH);
                    L.type = DFR_LIGHT_DIRECTIONAL;
                    for (int k = 0; k < 3; k++) {
                        L.radiance[k] = a.at(b + DIRECTIONAL_LIGHT_OFFSET_RADIANCE + k);
                        L.direction[k] = a.at(b + DIRECTIONAL_LIGHT_OFFSET_DIRECTION + k);
                    }
                    L.radius = a.at(b + DIRECTIONAL_LIGHT_OFFSET_RADIUS);
                    break;
                }
                case LIGHT_PEREZ: {  // :1495-1580; the sky's frame is fixed by PerezLight.set
                    const IA offs = ia(s_.perez_light_offsets, "perez_light_offsets");
                    const FA a = fa(s_.perez_lights, "perez_lights");
                    const int64_t b = offs.at(off);
                    L.type = DFR_LIGHT_PEREZ;
                    L.turbidity = a.at(b + PEREZ_LIGHT_OFFSET_TURBIDITY);
                    for (int k = 0; k < 3; k++) L.direction[k] = a.at(b + PEREZ_LIGHT_OFFSET_SUN_DIRECTION_WORLD_SPACE + k);
                    break;
                }
                case LIGHT_IMAGE: {  // :1393-1470; RGB floats, row 0 first; frame fixed by ImageLight's constructor
                    const IA offs = ia(s_.image_light_offsets, "image_light_offsets");
                    const FA a = fa(s_.image_lights, "image_lights");
                    const int64_t b = offs.at(off);
                    const int w = (int)a.at(b + IMAGE_LIGHT_OFFSET_RESOLUTION_X), h = (int)a.at(b + IMAGE_LIGHT_OFFSET_RESOLUTION_Y);
                    if (w < 1 || h < 1) fail(DFR_E_ARGUMENT, "image light with an empty image");
                    a.need(b + IMAGE_LIGHT_OFFSET_IMAGE, 3 * (int64_t)w * h);
                    std::vector<float> rgb(a.p + b + IMAGE_LIGHT_OFFSET_IMAGE, a.p + b + IMAGE_LIGHT_OFFSET_IMAGE + 3 * (int64_t)w * h);
                    L.type = DFR_LIGHT_IMAGE;
                    L.primitive = add_image(w, h, std::move(rgb));
                    L.cone_angle = degrees(a.at(b + IMAGE_LIGHT_OFFSET_ANGLE_RADIANS));
                    L.direction[0] = a.at(b + IMAGE_LIGHT_OFFSET_SCALE);
                    L.direction[1] = a.at(b + IMAGE_LIGHT_OFFSET_SCALE + 1);
                    L.radius = a.at(b + IMAGE_LIGHT_OFFSET_RADIUS);
                    break;
                }
                case LIGHT_DIFFUSE_AREA: {  // :1249-1302 (radiance, shape ID, shape offset, two-sided)
                    const FA a = fa(s_.diffuse_area_lights, "diffuse_area_lights");
                    const int64_t b = (int64_t)off * DIFFUSE_AREA_LIGHT_LENGTH;
                    a.need(b, DIFFUSE_AREA_LIGHT_LENGTH);
                    L.type = DFR_LIGHT_DIFFUSE_AREA;
                    for (int k = 0; k < 3; k++) L.radiance[k] = a.at(b + DIFFUSE_AREA_LIGHT_OFFSET_RADIANCE + k);
                    L.two_sided = a.at(b + DIFFUSE_AREA_LIGHT_OFFSET_IS_TWO_SIDED) != 0.0f;
                    area_light_shape_.resize(o_.lights.size() + 1, {0, 0});
                    area_light_shape_[o_.lights.size()] = {(int)a.at(b + DIFFUSE_AREA_LIGHT_OFFSET_SHAPE_ID),
                                                           (int)a.at(b + DIFFUSE_AREA_LIGHT_OFFSET_SHAPE_OFFSET)};
                    area_light_index_[off] = (int)o_.lights.size();
                    break;
                }
                default: fail(DFR_E_ARGUMENT, "light " + std::to_string(l) + ": unknown light type ID " + std::to_string(id));
            }
            o_.lights.push_back(L);
        }
        area_light_prim_.assign(o_.lights.size(), -1);
        area_light_shape_.resize(o_.lights.size(), {0, 0});
    }

    // Each DiffuseAreaLight emits from the primitive that carries it: a
    // Triangle3F (its one triangle) or a Sphere3F.
    void area_lights() {
        for (size_t l = 0; l < o_.lights.size(); l++) {
            dfr_light& L = o_.lights[l];
            if (L.type != DFR_LIGHT_DIFFUSE_AREA) continue;
            const int p = area_light_prim_[l];
            if (p < 0) fail(DFR_E_UNSUPPORTED, "light " + std::to_string(l) + ": a DiffuseAreaLight no primitive carries");
            const int shape = area_light_shape_[l].first;
            if (shape == SHAPE_TRIANGLE && o_.prims[p].shape == DFR_SHAPE_TRIANGLE_MESH && o_.prims[p].triangle_count == 1) {
                L.primitive = p;
                L.triangle = o_.prims[p].first_triangle;
            } else if (shape == SHAPE_SPHERE && o_.prims[p].shape == DFR_SHAPE_SPHERE) {
                L.primitive = p;
                L.triangle = -1;
            } else {
                fail(DFR_E_UNSUPPORTED, "light " + std::to_string(l) + ": an area light on a shape other than a "
                                                                     "Triangle3F or a Sphere3F");
            }
        }
    }
};

}  // namespace

// The transcoded scene owns every array its dfr_scene_desc points at.
struct Transcoded {
    Out out;
    dfr_scene_desc desc;
};

static int transcode(const dfr_compiled_scene* s, Transcoded& t, std::string& err) {
    if (!s) { err = "scene == null"; return DFR_E_NULL; }
    try {
        Transcoder(*s, t.out).run();
    } catch (const Failure& f) {
        err = f.msg;
        return f.code;
    } catch (const std::bad_alloc&) {
        err = "out of host memory while transcoding";
        return DFR_E_MEMORY;
    }
    Out& o = t.out;
    for (size_t i = 0; i < o.images.size(); i++) o.images[i].rgb = o.image_rgb[i].data();
    dfr_scene_desc& d = t.desc;
    std::memset(&d, 0, sizeof(d));
    d.n_vertices = (int64_t)o.positions.size() / 3;
    d.positions = o.positions.data();
    d.normals = o.normals.data();
    d.uvs = o.uvs.data();
    d.n_triangles = (int64_t)o.indices.size() / 3;
    d.indices = o.indices.data();
    d.n_primitives = (int32_t)o.prims.size();
    d.primitives = o.prims.data();
    d.n_materials = (int32_t)o.mats.size();
    d.materials = o.mats.data();
    d.n_lights = (int32_t)o.lights.size();
    d.lights = o.lights.data();
    std::memcpy(d.camera, o.camera, sizeof(d.camera));
    d.n_images = (int32_t)o.images.size();
    d.images = o.images.data();
    d.n_textures = (int32_t)o.textures.size();
    d.textures = o.textures.data();
    return DFR_OK;
}

// Used by dfr_set_compiled_scene (api.cpp): the scene and its matrices.
int transcode_compiled_scene(const dfr_compiled_scene* s, const dfr_scene_desc** desc, const float** xforms,
                             void** holder, std::string& err) {
    auto* t = new (std::nothrow) Transcoded();
    if (!t) { err = "out of host memory"; return DFR_E_MEMORY; }
    const int rc = transcode(s, *t, err);
    if (rc) { delete t; return rc; }
    *desc = &t->desc;
    *xforms = t->out.xforms.data();
    *holder = t;
    return DFR_OK;
}
void free_transcoded(void* holder) { delete static_cast<Transcoded*>(holder); }

}  // namespace dfr

struct dfr_transcoded {
    dfr::Transcoded t;
};

extern "C" {

int32_t dfr_transcode_compiled_scene(const dfr_compiled_scene* scene, dfr_transcoded** out, char* error,
                                     int32_t error_capacity) {
    if (error && error_capacity > 0) error[0] = 0;
    if (!out) return DFR_E_NULL;
    *out = nullptr;
    auto* r = new (std::nothrow) dfr_transcoded();
    if (!r) return DFR_E_MEMORY;
    std::string err;
    const int rc = dfr::transcode(scene, r->t, err);
    if (rc) {
        delete r;
        if (error && error_capacity > 0) {
            std::strncpy(error, err.c_str(), (size_t)error_capacity - 1);
            error[error_capacity - 1] = 0;
        }
        return rc;
    }
    *out = r;
    return DFR_OK;
}

const dfr_scene_desc* dfr_transcoded_desc(const dfr_transcoded* t) { return t ? &t->t.desc : nullptr; }
const float* dfr_transcoded_transforms(const dfr_transcoded* t) { return t ? t->t.out.xforms.data() : nullptr; }
void dfr_transcoded_free(dfr_transcoded* t) { delete t; }

}  // extern "C"
