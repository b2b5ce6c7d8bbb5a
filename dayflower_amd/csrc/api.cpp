// api.cpp — the C-ABI of include/dayflower_hip.h: renderer context, scene
// compiler (BVH + SoA upload), render loops, ray queries.  Batching and the
// serving loop live in serve.cpp.
#include "api_internal.h"

#include <algorithm>
#include <functional>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>

#include "bvh.h"
#include "perez_host.h"

using namespace dfr;

namespace dfr {

// ---------------------------------------------------------------------------
// material compiler: constant textures make every BSDF a fixed lobe list
// (Material.computeScatteringFunctions of M/scene/material/*Material.java,
// transport RADIANCE, multiple lobes allowed).
// ---------------------------------------------------------------------------
static float roughness_to_alpha(float r) {  // MicrofacetDistribution.convertRoughnessToAlpha
    float x = std::max(r, 1.0e-3f), y = (float)std::log((double)x);
    return 1.62142f + 0.819955f * y + 0.1734f * y * y + 0.0171201f * y * y * y + 0.000640711f * y * y * y * y;
}
static void set3(float* d, float a, float b, float c) { d[0] = a; d[1] = b; d[2] = c; }
static void sat3(float* d, const float* s) { for (int k = 0; k < 3; k++) d[k] = std::max(s[k], 0.0f); }
static bool black3(const float* c) { return c[0] == 0.0f && c[1] == 0.0f && c[2] == 0.0f; }

void compile_material(const dfr_material& m, DevMaterial& out, std::vector<DevLobe>& lobes) {
    out.first_lobe = (int32_t)lobes.size();
    out.n_lobes = 0;
    out.has_bsdf = 1;
    out.type = m.type;
    out.eta = 1.0f;
    for (int k = 0; k < 3; k++) out.emission[k] = m.emission[k];
    // a textured material keeps every lobe whatever its constant colours:
    // the lobes present at a hit follow from the textures there
    bool textured = false;
    for (int k = 0; k < 8; k++) { out.tex[k] = m.textures[k]; textured = textured || m.textures[k]; }
    out.tex_emission = m.emission_texture;
    out.remap = m.type == DFR_MAT_GLOSSY ? 0 : m.remap_roughness;  // Glossy uses its roughness as is
    auto add = [&](const DevLobe& l) { lobes.push_back(l); out.n_lobes++; };
    auto tr = [](DevLobe& l, float ax, float ay, bool sep) {
        l.alphaX = std::max(ax, 0.001f); l.alphaY = std::max(ay, 0.001f); l.separable = sep ? 1 : 0;
    };
    switch (m.type) {
        case DFR_MAT_MATTE: {
            DevLobe l{}; l.type = BX_DIFF | BX_REFL; sat3(l.R, m.kd);
            float deg = std::min(std::max(m.angle, 0.0f), 90.0f);
            if (deg == 0.0f) l.kind = LOBE_LAMBERT;
            else {
                float r = deg * (3.14159265358979323846f / 180.0f), r2 = r * r;
                l.kind = LOBE_OREN_NAYAR; l.A = 1.0f - (r2 / (2.0f * (r2 + 0.33f))); l.B = 0.45f * r2 / (r2 + 0.09f);
            }
            add(l);
            break;
        }
        case DFR_MAT_MIRROR: {
            DevLobe l{}; l.kind = LOBE_SPECULAR_R; l.type = BX_SPEC | BX_REFL; sat3(l.R, m.kr); l.fresnel = FR_CONST;
            if (black3(l.R) && !textured) { out.has_bsdf = 0; break; }
            add(l);
            break;
        }
        case DFR_MAT_GLASS: {
            float kr[3], kt[3]; sat3(kr, m.kr); sat3(kt, m.kt);
            float eta = m.eta[0];
            if (black3(kr) && black3(kt) && !textured) { out.has_bsdf = 0; break; }
            out.eta = eta;
            if (m.roughness_u == 0.0f && m.roughness_v == 0.0f) {
                DevLobe l{}; l.kind = LOBE_FRESNEL_SPECULAR; l.type = BX_SPEC | BX_REFL | BX_TRANS;
                std::memcpy(l.R, kr, 12); std::memcpy(l.T, kt, 12); l.etaA = 1.0f; l.etaB = eta;
                add(l);
                break;
            }
            float ru = m.remap_roughness ? roughness_to_alpha(m.roughness_u) : m.roughness_u;
            float rv = m.remap_roughness ? roughness_to_alpha(m.roughness_v) : m.roughness_v;
            if (!black3(kr) || textured) {
                DevLobe l{}; l.kind = LOBE_TS_R; l.type = BX_GLOSSY | BX_REFL; std::memcpy(l.R, kr, 12);
                tr(l, ru, rv, false); l.fresnel = FR_DIELECTRIC; l.frEtaI = 1.0f; l.frEtaT = eta;
                add(l);
            }
            if (!black3(kt) || textured) {
                DevLobe l{}; l.kind = LOBE_TS_T; l.type = BX_GLOSSY | BX_TRANS; std::memcpy(l.T, kt, 12);
                tr(l, ru, rv, false); l.etaA = 1.0f; l.etaB = eta;
                add(l);
            }
            break;
        }
        case DFR_MAT_METAL: {
            float ru = m.remap_roughness ? roughness_to_alpha(m.roughness_u) : m.roughness_u;
            float rv = m.remap_roughness ? roughness_to_alpha(m.roughness_v) : m.roughness_v;
            DevLobe l{}; l.kind = LOBE_TS_R; l.type = BX_GLOSSY | BX_REFL; set3(l.R, 1, 1, 1);
            l.fresnel = FR_CONDUCTOR; std::memcpy(l.condEta, m.eta, 12); std::memcpy(l.condK, m.k, 12);
            tr(l, ru, rv, false);
            add(l);
            break;
        }
        case DFR_MAT_PLASTIC: {
            float kd[3], ks[3]; sat3(kd, m.kd); sat3(ks, m.ks);
            if (!black3(kd) || textured) { DevLobe l{}; l.kind = LOBE_LAMBERT; l.type = BX_DIFF | BX_REFL; std::memcpy(l.R, kd, 12); add(l); }
            if (!black3(ks) || textured) {
                float r = m.remap_roughness ? roughness_to_alpha(m.roughness_u) : m.roughness_u;
                DevLobe l{}; l.kind = LOBE_TS_R; l.type = BX_GLOSSY | BX_REFL; std::memcpy(l.R, ks, 12);
                tr(l, r, r, false); l.fresnel = FR_DIELECTRIC; l.frEtaI = 1.5f; l.frEtaT = 1.0f;
                add(l);
            }
            if (out.n_lobes == 0) out.has_bsdf = 0;
            break;
        }
        case DFR_MAT_DISNEY: {  // thin = false, scatter distance = 0
            float color[3]; sat3(color, m.kd);
            float eta = m.eta[0], metal = m.disney_metallic, rough = m.disney_roughness, sheen = m.disney_sheen;
            float strans = m.disney_spec_trans;
            float lum = 0.212671f * color[0] + 0.715160f * color[1] + 0.072169f * color[2];
            float tint[3];
            for (int k = 0; k < 3; k++) tint[k] = lum > 0.0f ? color[k] / lum : 1.0f;
            float dw = (1.0f - metal) * (1.0f - strans);
            auto lerp = [](float a, float b, float t) { return (1.0f - t) * a + t * b; };
            if (dw > 0.0f) {
                DevLobe d{}; d.kind = LOBE_DISNEY_DIFFUSE; d.type = BX_DIFF | BX_REFL;
                for (int k = 0; k < 3; k++) d.R[k] = color[k] * dw;
                add(d);
                DevLobe r{}; r.kind = LOBE_DISNEY_RETRO; r.type = BX_DIFF | BX_REFL; r.roughness = rough;
                for (int k = 0; k < 3; k++) r.R[k] = color[k] * dw;
                add(r);
                if (sheen > 0.0f) {
                    DevLobe s{}; s.kind = LOBE_DISNEY_SHEEN; s.type = BX_DIFF | BX_REFL;
                    for (int k = 0; k < 3; k++) s.R[k] = lerp(1.0f, tint[k], m.disney_sheen_tint) * (dw * sheen);
                    add(s);
                }
            }
            float aspect = (float)std::sqrt((double)(1.0f - m.disney_anisotropic * 0.9f));
            float ax = std::max(0.001f, rough * rough / aspect), ay = std::max(0.001f, rough * rough * aspect);
            float r0 = ((eta - 1.0f) * (eta - 1.0f)) / ((eta + 1.0f) * (eta + 1.0f));
            DevLobe s{}; s.kind = LOBE_TS_R; s.type = BX_GLOSSY | BX_REFL; set3(s.R, 1, 1, 1);
            tr(s, ax, ay, true);
            s.fresnel = FR_DISNEY; s.disneyEta = eta; s.disneyMetallic = metal;
            for (int k = 0; k < 3; k++) s.r0[k] = lerp(lerp(1.0f, tint[k], m.disney_specular_tint) * r0, color[k], metal);
            add(s);
            if (m.disney_clearcoat > 0.0f) {
                DevLobe c{}; c.kind = LOBE_DISNEY_CLEARCOAT; c.type = BX_GLOSSY | BX_REFL;
                c.gloss = lerp(0.1f, 0.001f, m.disney_clearcoat_gloss); c.weight = m.disney_clearcoat;
                add(c);
            }
            if (strans > 0.0f) {
                DevLobe t{}; t.kind = LOBE_TS_T; t.type = BX_GLOSSY | BX_TRANS;
                for (int k = 0; k < 3; k++) t.T[k] = (float)std::sqrt((double)color[k]) * strans;
                tr(t, ax, ay, true); t.etaA = 1.0f; t.etaB = eta;
                add(t);
            }
            break;
        }
        case DFR_MAT_GLOSSY: {
            // TorranceSparrowBRDF(KR, ConductorFresnel(etaI = KR, etaT = 1,
            // k = 1)): the conductor term only sees etaT / etaI and k / etaI.
            DevLobe l{}; l.kind = LOBE_TS_R; l.type = BX_GLOSSY | BX_REFL; std::memcpy(l.R, m.kr, 12);
            l.fresnel = FR_CONDUCTOR;
            for (int k = 0; k < 3; k++) { l.condEta[k] = 1.0f / m.kr[k]; l.condK[k] = 1.0f / m.kr[k]; }
            tr(l, m.roughness_u, m.roughness_u, false);
            add(l);
            break;
        }
        case DFR_MAT_SUBSTRATE: {
            float kd[3], ks[3]; sat3(kd, m.kd); sat3(ks, m.ks);
            if (black3(kd) && black3(ks) && !textured) { out.has_bsdf = 0; break; }
            float ru = m.remap_roughness ? roughness_to_alpha(m.roughness_u) : m.roughness_u;
            float rv = m.remap_roughness ? roughness_to_alpha(m.roughness_v) : m.roughness_v;
            DevLobe l{}; l.kind = LOBE_FRESNEL_BLEND; l.type = BX_GLOSSY | BX_REFL;
            std::memcpy(l.R, kd, 12); std::memcpy(l.T, ks, 12); tr(l, ru, rv, false);
            add(l);
            break;
        }
        case DFR_MAT_TRANSLUCENT: {  // kr: reflectance, kt: transmittance
            float refl[3], trans[3], kd[3], ks[3]; sat3(refl, m.kr); sat3(trans, m.kt); sat3(kd, m.kd); sat3(ks, m.ks);
            if (black3(refl) && black3(trans)) { out.has_bsdf = 0; break; }
            if (!black3(kd)) {
                if (!black3(refl)) {
                    DevLobe l{}; l.kind = LOBE_LAMBERT; l.type = BX_DIFF | BX_REFL;
                    for (int k = 0; k < 3; k++) l.R[k] = kd[k] * refl[k];
                    add(l);
                }
                if (!black3(trans)) {
                    DevLobe l{}; l.kind = LOBE_LAMBERT_T; l.type = BX_DIFF | BX_TRANS;
                    for (int k = 0; k < 3; k++) l.T[k] = kd[k] * trans[k];
                    add(l);
                }
            }
            if (!black3(ks)) {
                float r = m.remap_roughness ? roughness_to_alpha(m.roughness_u) : m.roughness_u;
                if (!black3(refl)) {
                    DevLobe l{}; l.kind = LOBE_TS_R; l.type = BX_GLOSSY | BX_REFL;
                    for (int k = 0; k < 3; k++) l.R[k] = ks[k] * refl[k];
                    tr(l, r, r, false); l.fresnel = FR_DIELECTRIC; l.frEtaI = 1.0f; l.frEtaT = 1.5f;
                    add(l);
                }
                if (!black3(trans)) {
                    DevLobe l{}; l.kind = LOBE_TS_T; l.type = BX_GLOSSY | BX_TRANS;
                    for (int k = 0; k < 3; k++) l.T[k] = ks[k] * trans[k];
                    tr(l, r, r, false); l.etaA = 1.0f; l.etaB = 1.5f;
                    add(l);
                }
            }
            if (out.n_lobes == 0) out.has_bsdf = 0;
            break;
        }
        case DFR_MAT_CLEAR_COAT: {
            // Both layers; the shade kernel keeps one per hit (clear_coat_layer).
            DevLobe s{}; s.kind = LOBE_SPECULAR_R; s.type = BX_SPEC | BX_REFL; std::memcpy(s.R, m.ks, 12); s.fresnel = FR_CONST;
            add(s);
            DevLobe d{}; d.kind = LOBE_LAMBERT; d.type = BX_DIFF | BX_REFL; std::memcpy(d.R, m.kd, 12);
            add(d);
            break;
        }
        case DFR_MAT_UBER: {  // k: opacity
            float op[3], c[3];
            sat3(op, m.k);
            float kd[3], kr[3], ks[3], kt[3], ts[3];
            sat3(c, m.kd); for (int k = 0; k < 3; k++) kd[k] = op[k] * c[k];
            sat3(c, m.kr); for (int k = 0; k < 3; k++) kr[k] = op[k] * c[k];
            sat3(c, m.ks); for (int k = 0; k < 3; k++) ks[k] = op[k] * c[k];
            sat3(c, m.kt); for (int k = 0; k < 3; k++) kt[k] = op[k] * c[k];
            for (int k = 0; k < 3; k++) ts[k] = std::max(1.0f - op[k], 0.0f);
            const float eta = m.eta[0];
            float ru = m.remap_roughness ? roughness_to_alpha(m.roughness_u) : m.roughness_u;
            float rv = m.remap_roughness ? roughness_to_alpha(m.roughness_v) : m.roughness_v;
            if (!black3(ts)) {
                DevLobe l{}; l.kind = LOBE_SPECULAR_T; l.type = BX_SPEC | BX_TRANS; std::memcpy(l.T, ts, 12);
                l.etaA = 1.0f; l.etaB = 1.0f;
                add(l);
            }
            if (!black3(kd)) { DevLobe l{}; l.kind = LOBE_LAMBERT; l.type = BX_DIFF | BX_REFL; std::memcpy(l.R, kd, 12); add(l); }
            if (!black3(ks)) {
                DevLobe l{}; l.kind = LOBE_TS_R; l.type = BX_GLOSSY | BX_REFL; std::memcpy(l.R, ks, 12);
                tr(l, ru, rv, false); l.fresnel = FR_DIELECTRIC; l.frEtaI = 1.0f; l.frEtaT = eta;
                add(l);
            }
            if (!black3(kr)) {
                DevLobe l{}; l.kind = LOBE_SPECULAR_R; l.type = BX_SPEC | BX_REFL; std::memcpy(l.R, kr, 12);
                l.fresnel = FR_DIELECTRIC; l.frEtaI = 1.0f; l.frEtaT = eta;
                add(l);
            }
            if (!black3(kt)) {
                DevLobe l{}; l.kind = LOBE_SPECULAR_T; l.type = BX_SPEC | BX_TRANS; std::memcpy(l.T, kt, 12);
                l.etaA = 1.0f; l.etaB = eta;
                add(l);
            }
            // a BSDF with no lobe still ends the path (no pass-through)
            out.eta = black3(ts) ? eta : 1.0f;
            break;
        }
        default: out.has_bsdf = 0; break;
    }
}

// ---------------------------------------------------------------------------
// scene compiler
// ---------------------------------------------------------------------------
static float4 f4(float x, float y, float z, float w) { float4 v; v.x = x; v.y = y; v.z = z; v.w = w; return v; }

template <class T>
static int upload(DeviceBuffer& b, const std::vector<T>& v, hipStream_t s) {
    size_t bytes = std::max<size_t>(sizeof(T), v.size() * sizeof(T));
    if (b.alloc(bytes) != hipSuccess) return DFR_E_MEMORY;
    if (!v.empty() && hipMemcpyAsync(b.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess)
        return DFR_E_DEVICE;
    return DFR_OK;
}

// Object-to-world maps of the primitives (updateMatrix44Fs,
// AbstractSceneKernel.java:292-300; CompiledPrimitiveCache.java:323-350),
// validated: affine (last row 0 0 0 1) and invertible.  Spheres and
// rectangles take similarities (rotation or reflection, uniform scale,
// translation), which keep them spheres and rectangles; meshes and planes
// take any affine map.  A mesh that carries an area light takes rigid
// motions only and a sphere light translations only: the light samplers
// work in world space, where a scaled light's pdf would differ from the
// reference's object-space one.
int make_xforms(const SceneHost& in, const std::vector<float>& m, std::vector<Xform>& out, std::string& err) {
    out.clear();
    if (m.empty()) return DFR_OK;
    if (m.size() != 16 * in.prims.size()) { err = "one 4x4 matrix per primitive"; return DFR_E_ARGUMENT; }
    out.resize(in.prims.size());
    for (size_t p = 0; p < in.prims.size(); p++) {
        const float* M = &m[16 * p];
        Xform& x = out[p];
        for (int k = 0; k < 16; k++)
            if (!std::isfinite(M[k])) { err = "primitive " + std::to_string(p) + ": non-finite matrix"; return DFR_E_ARGUMENT; }
        if (M[12] != 0.0f || M[13] != 0.0f || M[14] != 0.0f || M[15] != 1.0f) {
            err = "primitive " + std::to_string(p) + ": projective matrices are not supported"; return DFR_E_UNSUPPORTED;
        }
        double A[3][3];
        for (int r = 0; r < 3; r++) { for (int c = 0; c < 3; c++) A[r][c] = M[4 * r + c]; x.t[r] = M[4 * r + 3]; }
        const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                           A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
        if (!(std::fabs(det) > 1e-30)) { err = "primitive " + std::to_string(p) + ": singular matrix"; return DFR_E_ARGUMENT; }
        // normal matrix: the inverse transpose, cofactors / det
        double C[3][3];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                C[r][c] = A[r1][c1] * A[r2][c2] - A[r1][c2] * A[r2][c1];
            }
        bool identity = true;
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                x.a[r][c] = A[r][c];
                x.n[r][c] = C[r][c] / det;
                identity = identity && A[r][c] == (r == c ? 1.0 : 0.0);
            }
        x.det = det;
        x.linear_identity = identity;
        x.identity = identity && x.t[0] == 0.0 && x.t[1] == 0.0 && x.t[2] == 0.0;
        // columns' lengths and orthogonality: A^T A = s^2 I for a similarity
        double g[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) g[i][j] = A[0][i] * A[0][j] + A[1][i] * A[1][j] + A[2][i] * A[2][j];
        const double s2 = (g[0][0] + g[1][1] + g[2][2]) / 3.0;
        x.scale = std::sqrt(s2);
        x.similarity = true;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                if (std::fabs(g[i][j] - (i == j ? s2 : 0.0)) > 1e-5 * s2) x.similarity = false;
        const int shape = in.prims[p].shape;
        if ((shape == DFR_SHAPE_SPHERE || shape == DFR_SHAPE_RECTANGLE) && !x.similarity) {
            err = "primitive " + std::to_string(p) + ": spheres and rectangles take rotation, uniform scale and translation only";
            return DFR_E_UNSUPPORTED;
        }
    }
    // area lights: the primitive that carries each one
    for (size_t l = 0; l < in.lights.size(); l++) {
        const dfr_light& L = in.lights[l];
        if (L.type != DFR_LIGHT_DIFFUSE_AREA) continue;
        int owner = -1;
        if (L.triangle < 0) {
            owner = L.primitive;
        } else {
            for (size_t p = 0; p < in.prims.size(); p++)
                if (in.prims[p].shape == DFR_SHAPE_TRIANGLE_MESH && L.triangle >= in.prims[p].first_triangle &&
                    L.triangle < in.prims[p].first_triangle + in.prims[p].triangle_count) owner = (int)p;
        }
        if (owner < 0 || owner >= (int)out.size()) continue;  // compile reports it
        const Xform& x = out[owner];
        const bool rigid = x.similarity && x.det > 0.0 && std::fabs(x.scale - 1.0) < 1e-5;
        if (L.triangle < 0 && !x.linear_identity) {
            err = "light " + std::to_string(l) + ": a sphere light takes translations only";
            return DFR_E_UNSUPPORTED;
        }
        if (L.triangle >= 0 && !rigid) {
            err = "light " + std::to_string(l) + ": a mesh that carries an area light takes rotations and translations only";
            return DFR_E_UNSUPPORTED;
        }
    }
    return DFR_OK;
}

namespace {
struct D3v { double x, y, z; };
inline D3v xpoint(const Xform* x, const float* p) {
    if (!x || x->identity) return {p[0], p[1], p[2]};
    D3v r;
    r.x = x->a[0][0] * p[0] + x->a[0][1] * p[1] + x->a[0][2] * p[2] + x->t[0];
    r.y = x->a[1][0] * p[0] + x->a[1][1] * p[1] + x->a[1][2] * p[2] + x->t[1];
    r.z = x->a[2][0] * p[0] + x->a[2][1] * p[1] + x->a[2][2] * p[2] + x->t[2];
    return r;
}
inline D3v xvec(const double (*m)[3], const float* v) {
    return {m[0][0] * v[0] + m[0][1] * v[1] + m[0][2] * v[2], m[1][0] * v[0] + m[1][1] * v[1] + m[1][2] * v[2],
            m[2][0] * v[0] + m[2][1] * v[1] + m[2][2] * v[2]};
}
inline std::array<float, 3> fl(D3v v) { return {(float)v.x, (float)v.y, (float)v.z}; }
inline std::array<float, 3> fl_unit(D3v v) {
    const double l = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    if (!(l > 0.0)) return {0.0f, 0.0f, 0.0f};
    return {(float)(v.x / l), (float)(v.y / l), (float)(v.z / l)};
}
}  // namespace

// World-space geometry of a scene under its transforms: the three vertex
// positions of every triangle (9 floats each), the world-space primitive
// records and the owning primitive of each triangle.
static int world_geometry(const SceneHost& in, const std::vector<Xform>& xf, std::vector<int32_t>& tri_prim,
                          const InstanceSet& is, std::vector<float>& wpos, std::vector<float>& wnrm, std::string& err) {
    const int64_t nt = (int64_t)in.indices.size() / 3;
    const int64_t nv = (int64_t)in.positions.size() / 3;
    // object-space vertex normals: given, or the normalised sum of the
    // adjacent faces' normals
    std::vector<float> vn;
    if (!in.normals.empty()) {
        vn = in.normals;
    } else {
        std::vector<double> acc(3 * nv, 0.0);
        for (int64_t t = 0; t < nt; t++) {
            const int32_t* ix = &in.indices[3 * t];
            auto P = [&](int k) { const float* p = &in.positions[3 * (int64_t)k]; return std::array<float, 3>{p[0], p[1], p[2]}; };
            auto nz = [](std::array<float, 3> v) {
                float l = (float)std::sqrt((double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
                if (l >= 0.99999982f && l <= 1.00000012f) return v;
                return std::array<float, 3>{v[0] / l, v[1] / l, v[2] / l};
            };
            auto A = P(ix[0]), B = P(ix[1]), C = P(ix[2]);
            auto ab = nz({B[0] - A[0], B[1] - A[1], B[2] - A[2]});
            auto ac = nz({C[0] - A[0], C[1] - A[1], C[2] - A[2]});
            auto n = nz({ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]});
            for (int k = 0; k < 3; k++) { acc[3 * ix[k]] += n[0]; acc[3 * ix[k] + 1] += n[1]; acc[3 * ix[k] + 2] += n[2]; }
        }
        vn.resize(3 * nv);
        for (int64_t v = 0; v < nv; v++) {
            double l = std::sqrt(acc[3 * v] * acc[3 * v] + acc[3 * v + 1] * acc[3 * v + 1] + acc[3 * v + 2] * acc[3 * v + 2]);
            if (l == 0) l = 1;
            for (int k = 0; k < 3; k++) vn[3 * v + k] = (float)(acc[3 * v + k] / l);
        }
    }
    wpos.resize(9 * nt);
    wnrm.resize(9 * nt);
    for (int64_t t = 0; t < nt; t++) {
        // a shared mesh's triangles stay in object space
        const Xform* x = xf.empty() || is.tri_group[t] >= 0 ? nullptr : &xf[tri_prim[t]];
        for (int k = 0; k < 3; k++) {
            const int64_t v = in.indices[3 * t + k];
            const auto p = fl(xpoint(x, &in.positions[3 * v]));
            std::memcpy(&wpos[9 * t + 3 * k], p.data(), 12);
            // N n, not normalised: the kernels normalise the barycentric
            // blend, which then equals the normal matrix applied to the
            // object-space blend (the reference transforms the interpolated
            // object-space normal)
            if (!x || x->linear_identity) std::memcpy(&wnrm[9 * t + 3 * k], &vn[3 * v], 12);
            else { const auto n = fl(xvec(x->n, &vn[3 * v])); std::memcpy(&wnrm[9 * t + 3 * k], n.data(), 12); }
        }
    }
    (void)err;
    return DFR_OK;
}

// Scenes with at least this many spheres and rectangles walk a BVH over them;
// fewer are tested linearly with the infinite planes (a handful of shape
// tests costs less than the tree's root node).
constexpr size_t ITEM_TREE_MIN = 8;

static bool items_in_tree(const SceneHost& in) {
    size_t n = 0;
    for (const dfr_primitive& pr : in.prims) n += pr.shape == DFR_SHAPE_SPHERE || pr.shape == DFR_SHAPE_RECTANGLE;
    return n >= ITEM_TREE_MIN;
}

// The spheres and rectangles of a scene, as items of their own BVH: each
// primitive's world-space box, padded by a few ulps of its coordinates (the
// box test must not cull a hit the shape's own test reports).  `prim[k]` is
// the primitive of item k, `box[6k..6k+5]` its lo.xyz, hi.xyz.
static void bounded_items(const SceneHost& in, const std::vector<Xform>& xf, std::vector<int32_t>& prim,
                          std::vector<float>& box) {
    prim.clear();
    box.clear();
    for (size_t p = 0; p < in.prims.size(); p++) {
        const dfr_primitive& pr = in.prims[p];
        if (pr.shape != DFR_SHAPE_SPHERE && pr.shape != DFR_SHAPE_RECTANGLE) continue;
        const Xform* x = xf.empty() || xf[p].identity ? nullptr : &xf[p];
        double lo[3], hi[3];
        if (pr.shape == DFR_SHAPE_SPHERE) {
            const D3v c = xpoint(x, pr.data);
            const double r = std::fabs((double)pr.data[3] * (x ? x->scale : 1.0));
            const double cc[3] = {c.x, c.y, c.z};
            for (int k = 0; k < 3; k++) { lo[k] = cc[k] - r; hi[k] = cc[k] + r; }
        } else {
            // corners A, B, C and D = A + C - B
            D3v q[4];
            for (int k = 0; k < 3; k++) q[k] = xpoint(x, pr.data + 3 * k);
            q[3] = {q[0].x + q[2].x - q[1].x, q[0].y + q[2].y - q[1].y, q[0].z + q[2].z - q[1].z};
            for (int k = 0; k < 3; k++) { lo[k] = INFINITY; hi[k] = -INFINITY; }
            for (const D3v& v : q) {
                const double vv[3] = {v.x, v.y, v.z};
                for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], vv[k]); hi[k] = std::max(hi[k], vv[k]); }
            }
        }
        prim.push_back((int32_t)p);
        float b[6];
        for (int k = 0; k < 3; k++) {
            const double pad = 4.0 * 1.1920928955078125e-07 * std::max(std::fabs(lo[k]), std::fabs(hi[k])) + 1e-30;
            b[k] = std::nextafter((float)(lo[k] - pad), -INFINITY);
            b[3 + k] = std::nextafter((float)(hi[k] + pad), INFINITY);
        }
        box.insert(box.end(), b, b + 6);
    }
}

// Refits a four-wide tree in place: each child box becomes the union of its
// subtree, leaves bounded by `leaf(first, count, box6)`; children are
// numbered after their parents, so one backward pass suffices.
int refit_nodes(std::vector<Node4>& nodes, const std::function<void(uint32_t, uint32_t, float*)>& leaf,
                std::string& err) {
    const int64_t nn = (int64_t)nodes.size();
    std::vector<std::array<float, 6>> bound(nn);
    for (int64_t q = nn - 1; q >= 0; q--) {
        Node4& w = nodes[q];
        std::array<float, 6> nb = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < 4; k++) {
            uint32_t ref;
            std::memcpy(&ref, &w.f[24 + k], 4);
            if (ref == EMPTY_REF) continue;
            std::array<float, 6> cb = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
            if (ref & LEAF_FLAG) {
                leaf((ref & 0x7fffffffu) >> 3, (ref & 7u) + 1u, cb.data());
            } else {
                if ((int64_t)ref <= q || (int64_t)ref >= nn) { err = "BVH child numbered before its parent"; return DFR_E_STATE; }
                cb = bound[ref];
            }
            w.f[0 + k] = cb[0]; w.f[4 + k] = cb[3];
            w.f[8 + k] = cb[1]; w.f[12 + k] = cb[4];
            w.f[16 + k] = cb[2]; w.f[20 + k] = cb[5];
            for (int a = 0; a < 3; a++) { nb[a] = std::min(nb[a], cb[a]); nb[3 + a] = std::max(nb[3 + a], cb[3 + a]); }
        }
        bound[q] = nb;
    }
    return DFR_OK;
}

// Bakes the per-triangle and per-primitive device records of a scene from
// its world-space geometry (`wpos`, `wnrm`: three vertices per triangle) in
// the BVH's triangle order: traversal triangles, shading lines, primitives,
// lights.  Fills `out`'s record vectors.
struct Baked {
    std::vector<float4> tris, tshade;
    std::vector<DevPrim> prims;
    std::vector<int32_t> analytic;
    std::vector<DevLight> lights;
    std::vector<DevPerez> perez;
    std::vector<float4> xforms;   // per primitive: DevScene::xforms
    std::vector<DevImageLight> images;
    std::vector<float> texels;
    std::vector<DevTexture> textures;
};

static int bake_scene(const SceneHost& in, const std::vector<Xform>& xf, const std::vector<int32_t>& tri_prim,
                      const InstanceSet& is, const std::vector<float>& wpos, const std::vector<float>& wnrm,
                      const std::vector<uint32_t>& tri_order, Baked& b, std::string& err) {
    const int64_t nt = (int64_t)in.indices.size() / 3;
    // Triangle3F surface normal: normalize(cross(normalize(b-a), normalize(c-a)))
    auto nz = [](std::array<float, 3> v) {
        float l = (float)std::sqrt((double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
        if (l >= 0.99999982f && l <= 1.00000012f) return v;
        return std::array<float, 3>{v[0] / l, v[1] / l, v[2] / l};
    };
    b.tshade.assign(8 * nt, f4(0.0f, 0.0f, 0.0f, 0.0f));
    for (int64_t t = 0; t < nt; t++) {
        const float* P = &wpos[9 * t];
        const float* N = &wnrm[9 * t];
        auto ab = nz({P[3] - P[0], P[4] - P[1], P[5] - P[2]});
        auto ac = nz({P[6] - P[0], P[7] - P[1], P[8] - P[2]});
        auto ng = nz({ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]});
        // a reflection turns the world-space winding: the normal matrix
        // keeps the object-space side
        if (!xf.empty() && is.tri_group[t] < 0 && xf[tri_prim[t]].det < 0.0) ng = {-ng[0], -ng[1], -ng[2]};
        float4* r = &b.tshade[8 * t];
        r[0] = f4(P[0], P[1], P[2], N[0]);
        r[1] = f4(P[3], P[4], P[5], N[1]);
        r[2] = f4(P[6], P[7], P[8], N[2]);
        r[3] = f4(N[3], N[4], N[5], ng[0]);
        r[4] = f4(N[6], N[7], N[8], ng[1]);
        // texture coordinates of a, b, c (zero without them)
        float uv[6] = {0, 0, 0, 0, 0, 0};
        if (!in.uvs.empty())
            for (int k = 0; k < 3; k++) {
                const int64_t v = in.indices[3 * t + k];
                uv[2 * k] = in.uvs[2 * v]; uv[2 * k + 1] = in.uvs[2 * v + 1];
            }
        r[5] = f4(ng[2], uv[0], uv[1], uv[2]);
        r[6] = f4(uv[3], uv[4], uv[5], 0.0f);
    }
    // the BVH's triangles (those of unshared meshes), in its order
    b.tris.resize(3 * tri_order.size());
    for (size_t i = 0; i < tri_order.size(); i++) {
        const uint32_t t = tri_order[i];
        const float* P = &wpos[9 * (int64_t)t];
        float tid, pid;
        int32_t ti = (int32_t)t, pi = tri_prim[t];
        std::memcpy(&tid, &ti, 4); std::memcpy(&pid, &pi, 4);
        b.tris[3 * i + 0] = f4(P[0], P[1], P[2], tid);
        b.tris[3 * i + 1] = f4(P[3], P[4], P[5], pid);
        b.tris[3 * i + 2] = f4(P[6], P[7], P[8], 0.0f);
    }
    b.prims.assign(in.prims.size(), DevPrim{});
    b.xforms.assign(6 * std::max<size_t>(1, in.prims.size()), f4(0.0f, 0.0f, 0.0f, 0.0f));
    b.analytic.clear();
    const bool tree = items_in_tree(in);
    for (size_t p = 0; p < in.prims.size(); p++) {
        const dfr_primitive& pr = in.prims[p];
        DevPrim& d = b.prims[p];
        std::memset(&d, 0, sizeof(d));
        d.shape = pr.shape; d.material = pr.material; d.area_light = pr.area_light;
        std::memcpy(d.data, pr.data, sizeof(d.data));
        const Xform* x = xf.empty() || xf[p].identity ? nullptr : &xf[p];
        // infinite planes, and spheres and rectangles too few for a tree of their own
        if (pr.shape == DFR_SHAPE_PLANE || (!tree && (pr.shape == DFR_SHAPE_SPHERE || pr.shape == DFR_SHAPE_RECTANGLE)))
            b.analytic.push_back((int32_t)p);
        if (x && !x->linear_identity && pr.shape != DFR_SHAPE_PLANE) {
            // surface frames are built in object space and carried to world
            // space: rows 0-2 map a world normal to its object-space
            // direction (A^T), rows 3-5 an object-space frame vector to world
            // space (the normal matrix, OrthonormalBasis33F.transformTranspose
            // of the inverse); for a sphere both are the rotation itself
            const double k0 = pr.shape == DFR_SHAPE_SPHERE ? 1.0 / x->scale : 1.0;
            const double k1 = pr.shape == DFR_SHAPE_SPHERE ? x->scale : 1.0;
            float4* X = &b.xforms[6 * p];
            for (int r = 0; r < 3; r++) {
                X[r] = f4((float)(x->a[0][r] * k0), (float)(x->a[1][r] * k0), (float)(x->a[2][r] * k0), 0.0f);
                X[3 + r] = f4((float)(x->n[r][0] * k1), (float)(x->n[r][1] * k1), (float)(x->n[r][2] * k1), 0.0f);
            }
            d.xformed = PRIM_FRAME_XFORM;
        }
        if (pr.shape == DFR_SHAPE_TRIANGLE_MESH && is.prim_group[p] >= 0) {
            // an instance: its mesh's records are in object space, the
            // kernels carry them to world space through this matrix
            d.xformed |= PRIM_INSTANCE;
            for (int r = 0; r < 3; r++) {
                for (int c = 0; c < 3; c++) d.data[4 * r + c] = x ? (float)x->a[r][c] : (r == c ? 1.0f : 0.0f);
                d.data[4 * r + 3] = x ? (float)x->t[r] : 0.0f;
            }
        }
        if (pr.shape == DFR_SHAPE_SPHERE && x) {
            const auto c = fl(xpoint(x, pr.data));
            d.data[0] = c[0]; d.data[1] = c[1]; d.data[2] = c[2];
            d.data[3] = (float)(pr.data[3] * x->scale);
        } else if (pr.shape == DFR_SHAPE_RECTANGLE && x) {
            for (int k = 0; k < 3; k++) {
                const auto q = fl(xpoint(x, pr.data + 3 * k));
                std::memcpy(d.data + 3 * k, q.data(), 12);
            }
        } else if (pr.shape == DFR_SHAPE_PLANE && x) {
            // Plane3F's frame: all three vectors go through the normal
            // matrix and are normalised, as the reference transforms an
            // intersection's basis (AbstractSceneKernel.java:1104-1217)
            const auto o = fl(xpoint(x, pr.data));
            const auto u = fl_unit(xvec(x->n, pr.data + 3));
            const auto v = fl_unit(xvec(x->n, pr.data + 6));
            const auto w = fl_unit(xvec(x->n, pr.data + 9));
            std::memcpy(d.data, o.data(), 12); std::memcpy(d.data + 3, u.data(), 12);
            std::memcpy(d.data + 6, v.data(), 12); std::memcpy(d.data + 9, w.data(), 12);
        }
        if (pr.shape == DFR_SHAPE_RECTANGLE) {
            auto nz3 = [](float* v) {
                float l = (float)std::sqrt((double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
                if (l >= 0.99999982f && l <= 1.00000012f) return;
                v[0] /= l; v[1] /= l; v[2] /= l;
            };
            float ab[3] = {d.data[3] - d.data[0], d.data[4] - d.data[1], d.data[5] - d.data[2]};
            float ac[3] = {d.data[6] - d.data[0], d.data[7] - d.data[1], d.data[8] - d.data[2]};
            nz3(ab); nz3(ac);
            float n[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
            nz3(n);
            if (x && x->det < 0.0) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; }
            d.normal[0] = n[0]; d.normal[1] = n[1]; d.normal[2] = n[2];
        }
    }
    b.lights.assign(in.lights.size(), DevLight{});
    b.perez.clear();
    b.images.clear();
    b.texels.clear();
    // textures: image textures' texels go into the shared texel array
    b.textures.assign(std::max<size_t>(1, in.textures.size()), DevTexture{});
    for (size_t t = 0; t < in.textures.size(); t++) {
        const dfr_texture& T = in.textures[t];
        DevTexture& d = b.textures[t];
        std::memset(&d, 0, sizeof(d));
        d.type = T.type; d.a = T.a; d.b = T.b;
        std::memcpy(d.color, T.color, 12);
        std::memcpy(d.origin, T.origin, 12);
        std::memcpy(d.t, T.t, 12);
        d.cell = T.cell; d.radius = T.radius;
        d.scale_u = T.scale[0]; d.scale_v = T.scale[1];
        // AngleF.degrees: wrapped into [0, 360], radians, cos and sin in double
        float deg = T.angle;
        if (!(deg >= 0.0f && deg <= 360.0f)) {
            const float span = std::nextafter(360.0f, 1e9f);
            deg = deg < 0.0f ? 360.0f - std::fmod(0.0f - deg, span) : 0.0f + std::fmod(deg - 360.0f, span);
        }
        const float rad = (float)(deg * (3.14159265358979323846 / 180.0));
        d.cos_angle = (float)std::cos((double)rad);
        d.sin_angle = (float)std::sin((double)rad);
        if (T.type == DFR_TEX_IMAGE) {
            const SceneHost::Image& im = in.images[T.image];
            d.width = im.width; d.height = im.height;
            d.texel = (int64_t)b.texels.size();
            b.texels.insert(b.texels.end(), im.rgb.begin(), im.rgb.end());
        }
    }
    // a textured material's coordinates on a transformed rectangle or plane,
    // and a bullseye's object-space distance, need the inverse map: refused
    for (size_t p = 0; p < in.prims.size(); p++) {
        if (xf.empty() || xf[p].identity) continue;
        const dfr_material& M = in.mats[in.prims[p].material];
        if (std::none_of(M.textures, M.textures + 8, [](int32_t t) { return t != 0; })) continue;
        const int shape = in.prims[p].shape;
        if (shape == DFR_SHAPE_RECTANGLE || shape == DFR_SHAPE_PLANE) {
            err = "primitive " + std::to_string(p) + ": a textured rectangle or plane takes no transform";
            return DFR_E_UNSUPPORTED;
        }
        for (int k = 0; k < 8; k++) {
            std::vector<int> todo;
            if (M.textures[k]) todo.push_back(M.textures[k] - 1);
            while (!todo.empty()) {
                const dfr_texture& T = in.textures[todo.back()];
                todo.pop_back();
                if (T.type == DFR_TEX_BULLSEYE) {
                    err = "primitive " + std::to_string(p) + ": a bullseye texture on a transformed primitive";
                    return DFR_E_UNSUPPORTED;
                }
                if (T.type == DFR_TEX_CHECKERBOARD || T.type == DFR_TEX_POLKA_DOT || T.type == DFR_TEX_BLEND) {
                    todo.push_back(T.a); todo.push_back(T.b);
                }
            }
        }
    }
    for (size_t l = 0; l < in.lights.size(); l++) {
        const dfr_light& L = in.lights[l];
        DevLight& d = b.lights[l];
        std::memset(&d, 0, sizeof(d));
        d.type = L.type; d.prim = L.primitive; d.tri = (int32_t)L.triangle; d.two_sided = L.two_sided;
        std::memcpy(d.radiance, L.radiance, 12);
        std::memcpy(d.position, L.position, 12);
        d.perez = -1;
        if (L.type == DFR_LIGHT_DIFFUSE_AREA && L.triangle < 0) {
            // a sphere primitive (Sphere3F.sample, Sphere3F.java:156-226)
            const int sp = L.primitive;
            if (sp < 0 || sp >= (int)in.prims.size() || in.prims[sp].shape != DFR_SHAPE_SPHERE) {
                err = "light " + std::to_string(l) + ": an area light is a triangle of a mesh or a sphere primitive";
                return DFR_E_UNSUPPORTED;
            }
            if (in.prims[sp].area_light != (int32_t)l) {
                err = "light " + std::to_string(l) + ": its sphere primitive must name it as its area light";
                return DFR_E_ARGUMENT;
            }
            if (!(in.prims[sp].data[3] > 0.0f)) { err = "light " + std::to_string(l) + ": sphere radius must be positive"; return DFR_E_ARGUMENT; }
            const double r = b.prims[sp].data[3];
            d.area = (float)(4.0 * 3.14159265358979323846 * r * r);
            d.prim = sp;
            d.tri = -1;
        } else if (L.type == DFR_LIGHT_DIFFUSE_AREA) {
            if (L.triangle >= nt) { err = "area light triangle out of range"; return DFR_E_ARGUMENT; }
            const float* P = &wpos[9 * L.triangle];
            float e1[3] = {P[3] - P[0], P[4] - P[1], P[5] - P[2]}, e2[3] = {P[6] - P[0], P[7] - P[1], P[8] - P[2]};
            float x = e1[1] * e2[2] - e1[2] * e2[1], y = e1[2] * e2[0] - e1[0] * e2[2], z = e1[0] * e2[1] - e1[1] * e2[0];
            d.area = 0.5f * (float)std::sqrt((double)(x * x + y * y + z * z));
            d.prim = tri_prim[L.triangle];
        } else if (L.type == DFR_LIGHT_PEREZ) {
            DevPerez pz;
            perez_build(pz, L.turbidity, L.direction);
            d.perez = (int32_t)b.perez.size();
            b.perez.push_back(pz);
        } else if (L.type == DFR_LIGHT_IMAGE) {
            if (L.primitive < 0 || L.primitive >= (int)in.images.size()) {
                err = "light " + std::to_string(l) + ": no such image"; return DFR_E_ARGUMENT;
            }
            const SceneHost::Image& im = in.images[L.primitive];
            const int64_t off = (int64_t)b.texels.size();
            b.texels.insert(b.texels.end(), im.rgb.begin(), im.rgb.end());
            float su = L.direction[0], sv = L.direction[1];
            if (su == 0.0f && sv == 0.0f) su = sv = 1.0f;
            DevImageLight il;
            image_light_build(il, im.width, im.height, off, b.texels.data(), L.cone_angle, su, sv,
                              L.radius > 0.0f ? L.radius : 10.0f);
            d.perez = (int32_t)b.images.size();
            b.images.push_back(il);
        } else if (L.type == DFR_LIGHT_SPOT || L.type == DFR_LIGHT_DIRECTIONAL) {
            const double x = L.direction[0], y = L.direction[1], z = L.direction[2], n = std::sqrt(x * x + y * y + z * z);
            if (!(n > 0.0)) { err = "light " + std::to_string(l) + ": zero direction"; return DFR_E_ARGUMENT; }
            for (int k = 0; k < 3; k++) d.dir[k] = (float)(L.direction[k] / n);
            if (L.type == DFR_LIGHT_SPOT) {  // SpotLight.java:205-206
                const double rad = 3.14159265358979323846 / 180.0;
                d.area = (float)std::cos((double)(float)(L.cone_angle * rad));
                d.param = (float)std::cos((double)(float)((L.cone_angle - L.cone_delta) * rad));
            } else {
                d.param = L.radius;
            }
        } else if (L.type != DFR_LIGHT_POINT) {
            err = "light " + std::to_string(l) + ": type out of range"; return DFR_E_ARGUMENT;
        }
    }
    return DFR_OK;
}

static int upload_baked(const Baked& bk, CompiledScene& out, hipStream_t s) {
    int rc;
    if ((rc = upload(out.tris, bk.tris, s))) return rc;
    if ((rc = upload(out.tshade, bk.tshade, s))) return rc;
    if ((rc = upload(out.prims, bk.prims, s))) return rc;
    if ((rc = upload(out.analytic, bk.analytic, s))) return rc;
    if ((rc = upload(out.lights, bk.lights, s))) return rc;
    if ((rc = upload(out.perez, bk.perez, s))) return rc;
    if ((rc = upload(out.xforms, bk.xforms, s))) return rc;
    if ((rc = upload(out.images, bk.images, s))) return rc;
    if ((rc = upload(out.texels, bk.texels, s))) return rc;
    if ((rc = upload(out.textures, bk.textures, s))) return rc;
    return DFR_OK;
}

static int validate_scene(const SceneHost& in, std::vector<int32_t>& tri_prim, InstanceSet& is, std::string& err) {
    const int64_t nt = (int64_t)in.indices.size() / 3;
    const int64_t nv = (int64_t)in.positions.size() / 3;
    for (int64_t i = 0; i < 3 * nt; i++)
        if (in.indices[i] < 0 || in.indices[i] >= nv) { err = "triangle index out of range"; return DFR_E_ARGUMENT; }
    tri_prim.assign(nt, -1);
    for (size_t p = 0; p < in.prims.size(); p++) {
        const dfr_primitive& pr = in.prims[p];
        if (pr.material < 0 || pr.material >= (int)in.mats.size()) { err = "primitive material out of range"; return DFR_E_ARGUMENT; }
        if (pr.shape == DFR_SHAPE_TRIANGLE_MESH) {
            if (pr.first_triangle < 0 || pr.triangle_count < 0 || pr.first_triangle + pr.triangle_count > nt) {
                err = "primitive triangle range out of bounds"; return DFR_E_ARGUMENT;
            }
            for (int64_t t = pr.first_triangle; t < pr.first_triangle + pr.triangle_count; t++) tri_prim[t] = (int32_t)p;
        } else if (pr.shape < DFR_SHAPE_TRIANGLE_MESH || pr.shape > DFR_SHAPE_PLANE) {
            err = "unknown shape"; return DFR_E_ARGUMENT;
        }
    }
    for (int64_t t = 0; t < nt; t++)
        if (tri_prim[t] < 0) { err = "triangle not owned by any mesh primitive"; return DFR_E_ARGUMENT; }
    return find_instances(in, is, err);
}

int compile_scene(const SceneHost& in, const std::vector<Xform>& xf, CompiledScene& out, int light_pdf_mode,
                  hipStream_t s, std::string& err) {
    const int64_t nt = (int64_t)in.indices.size() / 3;
    std::vector<int32_t> tri_prim;
    InstanceSet& is = out.inst;
    int rc = validate_scene(in, tri_prim, is, err);
    if (rc) return rc;
    if (is.any() && BVH_WIDTH != 4) { err = "shared meshes need four-wide trees"; return DFR_E_UNSUPPORTED; }
    std::vector<float> wpos, wnrm;
    if ((rc = world_geometry(in, xf, tri_prim, is, wpos, wnrm, err))) return rc;

    auto t0 = std::chrono::steady_clock::now();
    int threads = (int)std::max(1u, std::thread::hardware_concurrency());
    // the scene's BVH holds the triangles of unshared meshes (all of them
    // when no mesh is shared)
    std::vector<int32_t> flat_idx;
    std::vector<uint32_t> own;  // BVH input position -> triangle
    flat_idx.reserve(3 * nt);
    for (int64_t t = 0; t < nt; t++) {
        if (is.tri_group[t] >= 0) continue;
        own.push_back((uint32_t)t);
        for (int j = 0; j < 3; j++) flat_idx.push_back((int32_t)(3 * t + j));
    }
    const int64_t n_own = (int64_t)own.size();
    BvhResult bvh = build_bvh(wpos.data(), flat_idx.data(), n_own, threads, BVH_WIDTH);
    if (n_own < nt)
        for (uint32_t& t : bvh.tri_order) t = own[t];
    if (is.any()) build_instances(wpos, xf, is);
    out.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // deepest traversal stack (entries) the kernels can need
    out.max_depth = BVH_WIDTH == 4 ? bvh.max_stack4 + 1 : bvh.max_depth + 1;
    if (is.any()) out.max_depth = std::max(out.max_depth, is.max_stack + 1);
    out.n_nodes = BVH_WIDTH == 4 ? (int64_t)bvh.nodes4.size() : (int64_t)bvh.nodes.size();
    out.n_tris = nt;
    out.n_bvh_tris = n_own;
    out.sah_cost = bvh.sah_cost;

    // spheres and rectangles: a BVH of their own over their boxes, each box
    // given to the builder as a degenerate triangle (lo, hi, lo)
    {
        std::vector<int32_t> item_prim;
        std::vector<float> item_box;
        if (items_in_tree(in)) bounded_items(in, xf, item_prim, item_box);
        const int64_t ni = (int64_t)item_prim.size();
        out.item_nodes_host.clear();
        out.item_prim_order.clear();
        if (ni > 0) {
            std::vector<float> ipos(9 * ni);
            std::vector<int32_t> iidx(3 * ni);
            for (int64_t k = 0; k < ni; k++) {
                const float* B = &item_box[6 * k];
                const float v[9] = {B[0], B[1], B[2], B[3], B[4], B[5], B[0], B[1], B[2]};
                std::memcpy(&ipos[9 * k], v, sizeof(v));
                for (int j = 0; j < 3; j++) iidx[3 * k + j] = (int32_t)(3 * k + j);
            }
            BvhResult ib = build_bvh(ipos.data(), iidx.data(), ni, 1, 4);
            out.item_nodes_host = std::move(ib.nodes4);
            out.item_prim_order.resize(ni);
            for (int64_t k = 0; k < ni; k++) out.item_prim_order[k] = item_prim[ib.tri_order[k]];
            out.max_depth = std::max(out.max_depth, ib.max_stack4 + 1);
            if ((rc = upload(out.item_nodes, out.item_nodes_host, s))) return rc;
            if ((rc = upload(out.item_prims, out.item_prim_order, s))) return rc;
        }
        out.n_items = ni;
    }
    Baked bk;
    if ((rc = bake_scene(in, xf, tri_prim, is, wpos, wnrm, bvh.tri_order, bk, err))) return rc;
    // (host sources of the asynchronous uploads live until the final synchronisation)
    std::vector<Node4> inodes;
    std::vector<DevInstance> irecs;
    if (is.any()) {
        inodes = instance_nodes(is);
        irecs = instance_records(xf, is);
        if (inodes.size() * sizeof(Node4) >= (size_t(1) << 32) || is.tris.size() * sizeof(float4) >= (size_t(1) << 32)) {
            err = "scene too large: shared meshes' nodes and triangles must each stay below 4 GiB";
            return DFR_E_UNSUPPORTED;
        }
        if ((rc = upload(out.inst_nodes, inodes, s))) return rc;
        if ((rc = upload(out.inst_tris, is.tris, s))) return rc;
        if ((rc = upload(out.instances, irecs, s))) return rc;
    }
    std::vector<DevMaterial> mats(in.mats.size());
    std::vector<DevLobe> lobes;
    for (size_t m = 0; m < in.mats.size(); m++) compile_material(in.mats[m], mats[m], lobes);
    // the extended materials and every textured material run in the EXT kernels
    for (const dfr_material& m : in.mats)
        out.ext_features = out.ext_features || m.type >= DFR_MAT_GLOSSY || m.emission_texture ||
                           std::any_of(m.textures, m.textures + 8, [](int32_t t) { return t != 0; });
    // traversal addresses nodes and triangles with 32-bit byte offsets
    const size_t node_bytes = BVH_WIDTH == 4 ? bvh.nodes4.size() * sizeof(bvh.nodes4[0]) : bvh.nodes.size() * sizeof(bvh.nodes[0]);
    if (node_bytes >= (size_t(1) << 32) || bk.tris.size() * sizeof(float4) >= (size_t(1) << 32)) {
        err = "scene too large: BVH nodes and triangles must each stay below 4 GiB";
        return DFR_E_UNSUPPORTED;
    }
    if ((rc = BVH_WIDTH == 4 ? upload(out.nodes, bvh.nodes4, s) : upload(out.nodes, bvh.nodes, s))) return rc;
    if ((rc = upload_baked(bk, out, s))) return rc;
    // 4 x material sort class per primitive (see sort_key_in)
    std::vector<int32_t> prim_class(std::max<size_t>(1, bk.prims.size()), 0);
    for (size_t i = 0; i < bk.prims.size(); i++) {
        const int m = bk.prims[i].material;
        const int t = (m >= 0 && m < (int)in.mats.size()) ? in.mats[m].type : 0;
        prim_class[i] = 4 * std::min(t, 6);
    }
    if ((rc = upload(out.prim_class, prim_class, s))) return rc;
    if ((rc = upload(out.mats, mats, s))) return rc;
    if ((rc = upload(out.lobes, lobes, s))) return rc;
    // host copies kept for transform updates (the tree's topology stays, its boxes are refitted)
    out.nodes4_host = std::move(bvh.nodes4);
    out.tri_order = std::move(bvh.tri_order);
    DevScene& ds = out.dev;
    ds.nodes = (const float4*)out.nodes.ptr;
    ds.tris = (const float4*)out.tris.ptr;
    ds.tshade = (const float4*)out.tshade.ptr;
    ds.prims = (const DevPrim*)out.prims.ptr;
    ds.prim_class = (const int32_t*)out.prim_class.ptr;
    ds.analytic = (const int32_t*)out.analytic.ptr;
    ds.mats = (const DevMaterial*)out.mats.ptr;
    ds.lobes = (const DevLobe*)out.lobes.ptr;
    ds.lights = (const DevLight*)out.lights.ptr;
    ds.perez = (const DevPerez*)out.perez.ptr;
    ds.xforms = (const float4*)out.xforms.ptr;
    ds.images = (const DevImageLight*)out.images.ptr;
    ds.texels = (const float*)out.texels.ptr;
    ds.textures = (const DevTexture*)out.textures.ptr;
    ds.n_analytic = (int32_t)bk.analytic.size();
    ds.item_nodes = (const float4*)out.item_nodes.ptr;
    ds.item_prims = (const int32_t*)out.item_prims.ptr;
    ds.n_items = (int32_t)out.n_items;
    ds.inst_nodes = (const float4*)out.inst_nodes.ptr;
    ds.inst_tris = (const float4*)out.inst_tris.ptr;
    ds.instances = (const DevInstance*)out.instances.ptr;
    ds.n_instances = (int32_t)is.order.size();
    ds.prepass = (ds.n_items > 0 || ds.n_instances > 0) ? 1 : 0;
    ds.n_lights = (int32_t)bk.lights.size();
    for (const DevLight& L : bk.lights) {
        out.area_lights = out.area_lights || L.type == DFR_LIGHT_DIFFUSE_AREA;
        out.ext_features = out.ext_features || L.type == DFR_LIGHT_SPOT || L.type == DFR_LIGHT_DIRECTIONAL ||
                           L.type == DFR_LIGHT_IMAGE;
    }
    ds.has_bvh = n_own > 0 ? 1 : 0;
    ds.light_pdf_mode = light_pdf_mode;
    std::memcpy(out.camera, in.camera, sizeof(out.camera));
    out.bytes = out.nodes.bytes + out.tris.bytes + out.tshade.bytes + out.prims.bytes + out.analytic.bytes +
                out.item_nodes.bytes + out.item_prims.bytes + out.inst_nodes.bytes + out.inst_tris.bytes +
                out.instances.bytes +
                out.mats.bytes + out.lobes.bytes + out.lights.bytes + out.perez.bytes + out.xforms.bytes +
                out.images.bytes + out.texels.bytes + out.textures.bytes;
    if (hipStreamSynchronize(s) != hipSuccess) { err = "upload failed"; return DFR_E_DEVICE; }
    return DFR_OK;
}

// Moves a compiled scene's primitives to new transforms without a rebuild:
// world-space geometry is recomputed, the BVH keeps its topology and its
// boxes are refitted bottom-up (children are numbered after their parents),
// and every per-triangle and per-primitive record is uploaded again.
int refit_scene(const SceneHost& in, const std::vector<Xform>& xf, CompiledScene& out, hipStream_t s, std::string& err) {
    if (BVH_WIDTH != 4) { err = "transform updates refit four-wide trees"; return DFR_E_UNSUPPORTED; }
    const int64_t nt = (int64_t)in.indices.size() / 3;
    if (nt != out.n_tris) { err = "scene changed since it was compiled"; return DFR_E_STATE; }
    std::vector<int32_t> tri_prim;
    InstanceSet is;
    int rc = validate_scene(in, tri_prim, is, err);
    if (rc) return rc;
    if (is.prim_group != out.inst.prim_group) { err = "scene changed since it was compiled"; return DFR_E_STATE; }
    std::vector<float> wpos, wnrm;
    if ((rc = world_geometry(in, xf, tri_prim, is, wpos, wnrm, err))) return rc;
    Baked bk;
    if ((rc = bake_scene(in, xf, tri_prim, is, wpos, wnrm, out.tri_order, bk, err))) return rc;
    std::vector<Node4> inodes;
    std::vector<DevInstance> irecs;
    if (out.inst.any()) {
        // the shared meshes' trees are object space: only the top level moves
        refit_instances(xf, out.inst);
        inodes = instance_nodes(out.inst);
        irecs = instance_records(xf, out.inst);
        if ((rc = upload(out.inst_nodes, inodes, s))) return rc;
        if ((rc = upload(out.instances, irecs, s))) return rc;
    }
    // refit: each node's bound is the union of its children's
    std::vector<Node4>& nodes = out.nodes4_host;
    rc = refit_nodes(nodes, [&](uint32_t first, uint32_t count, float* cb) {
        for (uint32_t i = first; i < first + count; i++)
            for (int v = 0; v < 3; v++) {
                const float4& P = bk.tris[3 * (size_t)i + v];
                const float c[3] = {P.x, P.y, P.z};
                for (int a = 0; a < 3; a++) { cb[a] = std::min(cb[a], c[a]); cb[3 + a] = std::max(cb[3 + a], c[a]); }
            }
    }, err);
    if (rc) return rc;
    if (out.n_items > 0) {
        std::vector<int32_t> item_prim;
        std::vector<float> item_box;
        bounded_items(in, xf, item_prim, item_box);
        std::vector<int64_t> box_of(in.prims.size(), -1);
        for (size_t k = 0; k < item_prim.size(); k++) box_of[item_prim[k]] = (int64_t)k;
        rc = refit_nodes(out.item_nodes_host, [&](uint32_t first, uint32_t count, float* cb) {
            for (uint32_t i = first; i < first + count; i++) {
                const float* B = &item_box[6 * box_of[out.item_prim_order[i]]];
                for (int a = 0; a < 3; a++) { cb[a] = std::min(cb[a], B[a]); cb[3 + a] = std::max(cb[3 + a], B[3 + a]); }
            }
        }, err);
        if (rc) return rc;
        if ((rc = upload(out.item_nodes, out.item_nodes_host, s))) return rc;
    }
    // same sizes as before: the device buffers are rewritten in place, so the
    // scene record's pointers stay valid
    if ((rc = upload(out.nodes, nodes, s))) return rc;
    if ((rc = upload_baked(bk, out, s))) return rc;
    if (hipStreamSynchronize(s) != hipSuccess) { err = "upload failed"; return DFR_E_DEVICE; }
    return DFR_OK;
}

int copy_scene_desc(const dfr_scene_desc* d, SceneHost& h, std::string& err) {
    if (!d) { err = "scene == null"; return DFR_E_NULL; }
    if (d->n_vertices < 0 || d->n_triangles < 0 || d->n_primitives < 0 || d->n_materials < 0 || d->n_lights < 0) {
        err = "negative count"; return DFR_E_ARGUMENT;
    }
    if ((d->n_vertices && !d->positions) || (d->n_triangles && !d->indices) || (d->n_primitives && !d->primitives) ||
        (d->n_materials && !d->materials) || (d->n_lights && !d->lights)) {
        err = "null array"; return DFR_E_NULL;
    }
    for (int32_t i = 0; i < d->n_materials; i++) {
        if (d->materials[i].type < DFR_MAT_MATTE || d->materials[i].type > DFR_MAT_UBER) {
            err = "material " + std::to_string(i) + ": type out of range"; return DFR_E_ARGUMENT;
        }
        // DisneyMaterial.java:1140-1146, 1185-1197: the thin form's fake
        // subsurface and diffuse transmission lobes are not compiled
        if (d->materials[i].type == DFR_MAT_DISNEY && d->materials[i].thin) {
            err = "material " + std::to_string(i) + ": thin Disney materials are not supported"; return DFR_E_UNSUPPORTED;
        }
    }
    h.positions.assign(d->positions, d->positions + 3 * d->n_vertices);
    if (d->normals) h.normals.assign(d->normals, d->normals + 3 * d->n_vertices); else h.normals.clear();
    h.indices.assign(d->indices, d->indices + 3 * d->n_triangles);
    h.prims.assign(d->primitives, d->primitives + d->n_primitives);
    h.mats.assign(d->materials, d->materials + d->n_materials);
    h.lights.assign(d->lights, d->lights + d->n_lights);
    std::memcpy(h.camera, d->camera, sizeof(h.camera));
    h.uvs.clear();
    if (d->uvs && d->n_vertices > 0) h.uvs.assign(d->uvs, d->uvs + 2 * d->n_vertices);
    if (d->n_textures < 0 || (d->n_textures && !d->textures)) { err = "textures == null"; return DFR_E_NULL; }
    h.textures.assign(d->textures, d->textures + d->n_textures);
    if (d->n_images < 0 || (d->n_images && !d->images)) { err = "images == null"; return DFR_E_NULL; }
    h.images.assign(d->n_images, SceneHost::Image{});
    for (int32_t i = 0; i < d->n_images; i++) {
        const dfr_image& im = d->images[i];
        if (im.width < 1 || im.height < 1) { err = "image " + std::to_string(i) + ": empty"; return DFR_E_ARGUMENT; }
        if (!im.rgb) { err = "image " + std::to_string(i) + ": rgb == null"; return DFR_E_NULL; }
        if ((int64_t)im.width * im.height > (int64_t(1) << 28)) { err = "image " + std::to_string(i) + ": too large"; return DFR_E_ARGUMENT; }
        h.images[i].width = im.width;
        h.images[i].height = im.height;
        h.images[i].rgb.assign(im.rgb, im.rgb + 3 * (size_t)im.width * im.height);
    }
    // textures: indices in range, blends of non-blends, no cycles; material
    // slots only where the kernels colour a lobe from them
    const int nt = (int)h.textures.size();
    for (int t = 0; t < nt; t++) {
        const dfr_texture& T = h.textures[t];
        const std::string who = "texture " + std::to_string(t);
        if (T.type < DFR_TEX_CONSTANT || T.type > DFR_TEX_SIMPLEX_FBM) { err = who + ": type out of range"; return DFR_E_UNSUPPORTED; }
        // noise octaves: a whole number the kernels loop over
        if ((T.type == DFR_TEX_MARBLE || T.type == DFR_TEX_SIMPLEX_FBM) &&
            !(T.cell >= 0.0f && T.cell <= 64.0f && T.cell == std::floor(T.cell))) {
            err = who + ": octaves must be a whole number from 0 to 64"; return DFR_E_ARGUMENT;
        }
        const bool two = T.type == DFR_TEX_CHECKERBOARD || T.type == DFR_TEX_BULLSEYE || T.type == DFR_TEX_POLKA_DOT ||
                         T.type == DFR_TEX_BLEND;
        if (two && (T.a < 0 || T.a >= nt || T.b < 0 || T.b >= nt)) { err = who + ": sub-texture out of range"; return DFR_E_ARGUMENT; }
        if (T.type == DFR_TEX_BLEND && (h.textures[T.a].type == DFR_TEX_BLEND || h.textures[T.b].type == DFR_TEX_BLEND)) {
            err = who + ": a blend of blends"; return DFR_E_UNSUPPORTED;
        }
        if (T.type == DFR_TEX_IMAGE && (T.image < 0 || T.image >= (int)h.images.size())) { err = who + ": no such image"; return DFR_E_ARGUMENT; }
    }
    {   // a chain of sub-textures must end: depth-first search for cycles
        std::vector<int> state(nt, 0);
        std::function<bool(int)> cyclic = [&](int t) {
            if (state[t] == 1) return true;
            if (state[t] == 2) return false;
            state[t] = 1;
            const dfr_texture& T = h.textures[t];
            const bool two = T.type == DFR_TEX_CHECKERBOARD || T.type == DFR_TEX_BULLSEYE || T.type == DFR_TEX_POLKA_DOT ||
                             T.type == DFR_TEX_BLEND;
            if (two && (cyclic(T.a) || cyclic(T.b))) return true;
            state[t] = 2;
            return false;
        };
        for (int t = 0; t < nt; t++)
            if (cyclic(t)) { err = "texture " + std::to_string(t) + ": its sub-textures form a cycle"; return DFR_E_ARGUMENT; }
    }
    for (size_t m = 0; m < h.mats.size(); m++) {
        const dfr_material& M = h.mats[m];
        if (M.emission_texture < 0 || M.emission_texture > nt) {
            err = "material " + std::to_string(m) + ": emission texture out of range"; return DFR_E_ARGUMENT;
        }
        for (int k = 0; k < 8; k++) {
            if (!M.textures[k]) continue;
            if (M.textures[k] < 0 || M.textures[k] > nt) { err = "material " + std::to_string(m) + ": texture out of range"; return DFR_E_ARGUMENT; }
            const bool ok = (M.type == DFR_MAT_MATTE && k == DFR_SLOT_KD) ||
                            (M.type == DFR_MAT_PLASTIC && (k == DFR_SLOT_KD || k == DFR_SLOT_KS || k == DFR_SLOT_ROUGHNESS_U)) ||
                            (M.type == DFR_MAT_MIRROR && k == DFR_SLOT_KR) ||
                            (M.type == DFR_MAT_GLASS && (k == DFR_SLOT_KR || k == DFR_SLOT_KT || k == DFR_SLOT_ETA)) ||
                            (M.type == DFR_MAT_METAL && (k == DFR_SLOT_ROUGHNESS_U || k == DFR_SLOT_ROUGHNESS_V ||
                                                         k == DFR_SLOT_ETA || k == DFR_SLOT_K)) ||
                            (M.type == DFR_MAT_SUBSTRATE && (k == DFR_SLOT_KD || k == DFR_SLOT_KS ||
                                                             k == DFR_SLOT_ROUGHNESS_U || k == DFR_SLOT_ROUGHNESS_V)) ||
                            (M.type == DFR_MAT_GLOSSY && (k == DFR_SLOT_KR || k == DFR_SLOT_ROUGHNESS_U)) ||
                            (M.type == DFR_MAT_CLEAR_COAT && (k == DFR_SLOT_KD || k == DFR_SLOT_KS));
            if (!ok) { err = "material " + std::to_string(m) + ": this slot takes no texture"; return DFR_E_UNSUPPORTED; }
        }
    }
    return DFR_OK;
}

// ---------------------------------------------------------------------------
// context helpers
// ---------------------------------------------------------------------------
int Ctx::fail(int code, const std::string& msg) {
    error = msg;
    return code;
}

int Ctx::hip_fail(hipError_t e, const char* what) {
    error = std::string(what) + ": " + hipGetErrorString(e);
    return DFR_E_DEVICE;
}

int Ctx::sync_scene_table() {
    std::vector<DevScene> table(scenes.size());
    for (size_t i = 0; i < scenes.size(); i++)
        if (scenes[i]) table[i] = scenes[i]->dev;
    if (scene_table.alloc(std::max<size_t>(1, table.size()) * sizeof(DevScene)) != hipSuccess) return fail(DFR_E_MEMORY, "scene table");
    hipError_t e = hipMemcpyAsync(scene_table.ptr, table.data(), table.size() * sizeof(DevScene), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_fail(e, "scene table upload");
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "scene table upload");
    scene_gen++;
    return DFR_OK;
}

// The device record of the scene all uploaded jobs share, else null.
const DevScene* Ctx::uniform_scene() const {
    if (jobs_scene < 0 || jobs_scene >= (int)scenes.size() || !scenes[jobs_scene]) return nullptr;
    return (const DevScene*)scene_table.ptr + jobs_scene;
}

bool Ctx::any_area_lights() const {
    for (auto& s : scenes) if (s && s->area_lights) return true;
    return false;
}

bool Ctx::any_items() const {
    for (auto& s : scenes) if (s && s->dev.prepass) return true;
    return false;
}

bool Ctx::any_ext_features() const {
    for (auto& s : scenes) if (s && s->ext_features) return true;
    return false;
}

int Ctx::max_depth() const {
    int d = 1;
    for (auto& s : scenes) if (s) d = std::max(d, s->max_depth);
    return d;
}

int Ctx::ensure_overflow(int64_t threads) {
    size_t need = (size_t)overflow_entries_per_thread(max_depth()) * (size_t)threads * sizeof(int);
    if (overflow.bytes >= need) return DFR_OK;
    if (overflow.alloc(need) != hipSuccess) return fail(DFR_E_MEMORY, "traversal overflow stack");
    return DFR_OK;
}

// ColorSpaceF.S_R_G_B (M/color/ColorSpaceF.java:85, 116-127, 328-392) and
// GaussianFilter2F(2, 2, 2)'s 16 x 16 table (M/filter/Filter2F.java:119-136,
// M/filter/GaussianFilter2F.java:51-53, 117-119, 142-144), in float as the
// reference evaluates them (exp and pow in double, rounded).
static void make_film_consts(FilmConsts& f) {
    auto fpow = [](float a, float b) { return (float)std::pow((double)a, (double)b); };
    const float bp = 0.00304f, g = 2.4f;
    const float xR = 0.6400f, yR = 0.3300f, xG = 0.3000f, yG = 0.6000f, xB = 0.1500f, yB = 0.0600f;
    const float xW = 0.31271f, yW = 0.32902f;
    f.break_point = bp;
    f.gamma_recip = 1.0f / g;
    f.slope = 1.0f / (g / fpow(bp, 1.0f / g - 1.0f) - g * bp + bp);
    f.slope_match = g * f.slope / fpow(bp, 1.0f / g - 1.0f);
    f.segment_offset = f.slope_match * fpow(bp, 1.0f / g) - f.slope * bp;
    const float zR = 1.0f - (xR + yR), zG = 1.0f - (xG + yG), zB = 1.0f - (xB + yB), zW = 1.0f - (xW + yW);
    // rows of XYZ -> RGB: each primary's row is the cross product of the other
    // two primaries' chromaticities, scaled to map the white point to one
    const float row[3][3] = {{yG * zB - yB * zG, xB * zG - xG * zB, xG * yB - xB * yG},
                             {yB * zR - yR * zB, xR * zB - xB * zR, xB * yR - xR * yB},
                             {yR * zG - yG * zR, xG * zR - xR * zG, xR * yG - xG * yR}};
    float* m = f.xyz_to_rgb;
    for (int r = 0; r < 3; r++) {
        const float wr = (row[r][0] * xW + row[r][1] * yW + row[r][2] * zW) / yW;
        for (int k = 0; k < 3; k++) m[3 * r + k] = row[r][k] / wr;
        m[9 + r] = wr;
    }
    // RGB -> XYZ, the inverse by cofactors, stored by primary (X, Y, Z of red, ...)
    const float a = m[0] * (m[4] * m[8] - m[7] * m[5]);
    const float b = m[1] * (m[3] * m[8] - m[6] * m[5]);
    const float cc = m[2] * (m[3] * m[7] - m[6] * m[4]);
    const float sc = 1.0f / (a - b + cc);
    float* r = f.rgb_to_xyz;
    r[0] = sc * (m[4] * m[8] - m[5] * m[7]); r[1] = sc * (m[5] * m[6] - m[3] * m[8]); r[2] = sc * (m[3] * m[7] - m[4] * m[6]);
    r[3] = sc * (m[2] * m[7] - m[1] * m[8]); r[4] = sc * (m[0] * m[8] - m[2] * m[6]); r[5] = sc * (m[1] * m[6] - m[0] * m[7]);
    r[6] = sc * (m[1] * m[5] - m[2] * m[4]); r[7] = sc * (m[2] * m[3] - m[0] * m[5]); r[8] = sc * (m[0] * m[4] - m[1] * m[3]);
    r[9] = xW; r[10] = yW; r[11] = zW;
    const float radius = 2.0f, falloff = 2.0f;
    const float edge = (float)std::exp((double)(-falloff * radius * radius));
    auto gauss = [&](float v) { return std::max(0.0f, (float)std::exp((double)(-falloff * v * v)) - edge); };
    for (int i = 0, y = 0; y < 16; y++)
        for (int x = 0; x < 16; x++)
            f.table[i++] = gauss((x + 0.5f) * radius * (1.0f / 16)) * gauss((y + 0.5f) * radius * (1.0f / 16));
}

int Ctx::ensure_filtered_film() {
    if (!film_consts.ptr) {
        FilmConsts fc;
        make_film_consts(fc);
        if (film_consts.alloc(sizeof(FilmConsts)) != hipSuccess) return fail(DFR_E_MEMORY, "film constants");
        hipError_t e = hipMemcpyAsync(film_consts.ptr, &fc, sizeof(fc), hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail(e, "film constants");
    }
    const int64_t n = (int64_t)width * height;
    if (xyz_pixels != n) {
        if (xyz_film.alloc((size_t)n * sizeof(float4)) != hipSuccess) return fail(DFR_E_MEMORY, "XYZ film");
        hipError_t e = hipMemsetAsync(xyz_film.ptr, 0, (size_t)n * sizeof(float4), stream);
        if (e != hipSuccess) return hip_fail(e, "XYZ film clear");
        xyz_pixels = n;
    }
    return DFR_OK;
}

int Ctx::ensure_film(int64_t pixels) {
    size_t need = (size_t)pixels * sizeof(float4);
    if (film.bytes < need && film.alloc(need) != hipSuccess) return fail(DFR_E_MEMORY, "film");
    return DFR_OK;
}

bool alloc_wavefront_buffers(WavefrontBuffers& wf, DeviceBuffer* mem, int capacity) {
    const size_t c = (size_t)capacity;
    bool ok = true;
    auto A = [&](DeviceBuffer& b, size_t bytes) { ok = ok && b.alloc(bytes) == hipSuccess; return b.ptr; };
    wf.a.o = (float4*)A(mem[0], c * 16); wf.a.d = (float4*)A(mem[1], c * 16);
    wf.a.beta = (float4*)A(mem[2], c * 16); wf.a.info = (uint4*)A(mem[3], c * 16);
    wf.b.o = (float4*)A(mem[4], c * 16); wf.b.d = (float4*)A(mem[5], c * 16);
    wf.b.beta = (float4*)A(mem[6], c * 16); wf.b.info = (uint4*)A(mem[7], c * 16);
    wf.hit = (float4*)A(mem[8], c * 16);
    wf.perm = (int32_t*)A(mem[9], c * 4);
    wf.keys = (int32_t*)A(mem[10], c * 4);
    wf.bhist = (int32_t*)A(mem[16], 32 * (c / 4096 + 2) * 4);  // tile histograms + key totals
    wf.sh.o = (float4*)A(mem[11], 2 * c * 16); wf.sh.d = (float4*)A(mem[12], 2 * c * 16);
    wf.sh.c = (float4*)A(mem[13], 2 * c * 16); wf.sh.aux = (int4*)A(mem[14], 2 * c * 16);
    wf.ctr = (Counters*)A(mem[15], sizeof(Counters));
    if (!ok) { wf.ctr = nullptr; return false; }
    wf.capacity = capacity;
    return true;
}

int Ctx::ensure_wavefront(int capacity) {
    // buffers only grow; a run may use fewer slots than are allocated
    if (wf_capacity >= capacity && wf.ctr) { wf.capacity = capacity; return DFR_OK; }
    if (!alloc_wavefront_buffers(wf, wf_mem, capacity)) return fail(DFR_E_MEMORY, "wavefront buffers");
    wf_capacity = capacity;
    return DFR_OK;
}

size_t wavefront_bytes(int capacity) {
    return (size_t)capacity * (9 * 16 + 8 + 2 * 4 * 16) + 32 * ((size_t)capacity / 4096 + 2) * 4 + sizeof(Counters);
}

// Runs `jobs` (already on the device in job_buf) to completion through the
// wavefront engine.  Counters are polled every `poll` iterations.
// Runs `iterations` (an even number, so the ping-pong queues end where they
// started) as one graph launch, capturing the graph first when `key` differs
// from the cached one.
int Ctx::replay_iterations(const std::vector<uintptr_t>& key, const DevScene* scenes, const DevJob* jobs, int n_jobs,
                           float4* film, const WavefrontParams& p, int iterations) {
    GraphEntry* g = nullptr;
    for (GraphEntry& e : graphs)
        if (e.key == key) { g = &e; break; }
    if (!g) {
        if (graphs.size() >= MAX_GRAPHS) {
            auto lru = std::min_element(graphs.begin(), graphs.end(),
                                        [](const GraphEntry& a, const GraphEntry& b) { return a.used < b.used; });
            hipGraphExecDestroy(lru->exec);
            graphs.erase(lru);
        }
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return hip_fail(e, "graph capture");
        for (int k = 0; k < iterations; k++) wavefront_iteration(scenes, jobs, n_jobs, film, wf, p, nullptr, stream);
        e = hipStreamEndCapture(stream, &gr);
        if (e != hipSuccess) return hip_fail(e, "graph capture");
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
        hipGraphDestroy(gr);
        if (e != hipSuccess) return hip_fail(e, "graph instantiate");
        graphs.push_back(GraphEntry{key, exec, 0});
        g = &graphs.back();
    }
    g->used = ++graph_clock;
    hipError_t e = hipGraphLaunch(g->exec, stream);
    if (e != hipSuccess) return hip_fail(e, "graph launch");
    return DFR_OK;
}

int Ctx::run_wavefront(int n_jobs, int64_t work_total, float4* film_ptr, Counters* out_ctr) {
    // a run never needs more slots than it has work items: smaller runs get a
    // smaller pool (and smaller launch grids)
    int64_t want = 65536;
    while (want < work_total && want < wf_params.capacity) want <<= 1;
    const int capacity = (int)std::min<int64_t>(want, wf_params.capacity);
    int rc = ensure_wavefront(capacity);
    if (rc) return rc;
    int64_t threads = wavefront_stack_threads(capacity);
    if ((rc = ensure_overflow(threads))) return rc;
    wf.overflow = (int*)overflow.ptr;
    wavefront_init(wf, (const DevJob*)job_buf.ptr, n_jobs, work_total, stream);
    WavefrontParams p{wf_params.block, wf_params.sort_keys, (capacity + 255) / 256, any_area_lights(), uniform_scene(),
                      any_ext_features()};
    p.items = any_items();
    // After a poll that finds every work item generated, the next
    // iterations' queues hold at most the paths waiting then: size their
    // grids to that count, rounded up to capacity / 4^k (a few captured
    // graphs serve every run).
    auto bound_grids = [&](const Counters& h) {
        if (h.work_next < h.work_total) { p.queue_bound = 0; p.generate = true; return; }
        int b = capacity;
        while (b / 4 >= std::max<int64_t>(h.n_next, 1) && b / 4 >= 16384) b /= 4;
        p.queue_bound = b;
        p.generate = false;
    };
    Counters host{};
    const int poll = 4;
    const int max_iterations = 1 << 20;  // far above max_bounce * (work / capacity + 1)
    unsigned long long* trav = nullptr;
    wf.sample = RaySample{};
    if (count_traversal) {
        if (counters.alloc(128) != hipSuccess) return fail(DFR_E_MEMORY, "counters");
        trav = (unsigned long long*)counters.ptr + 4;
        hipMemsetAsync(trav, 0, 40, stream);
        if (ray_sample_stride > 0) {
            constexpr int cap = 1 << 20;  // records of 48 bytes
            if (sample_buf.alloc((size_t)cap * 3 * sizeof(float4)) != hipSuccess) return fail(DFR_E_MEMORY, "ray sample");
            wf.sample.rec = (float4*)sample_buf.ptr;
            wf.sample.cap = cap;
            wf.sample.stride = ray_sample_stride;
        }
    }
    int iterations = 0;
    const DevScene* scenes = (const DevScene*)scene_table.ptr;
    const DevJob* jobs = (const DevJob*)job_buf.ptr;
    if (use_graphs && !profile && !trav) {
        // everything the captured launches bake in
        const std::vector<uintptr_t> key = {
            (uintptr_t)scenes, (uintptr_t)jobs, (uintptr_t)n_jobs, (uintptr_t)film_ptr, (uintptr_t)wf.a.o,
            (uintptr_t)wf.a.d, (uintptr_t)wf.a.beta, (uintptr_t)wf.a.info, (uintptr_t)wf.b.o, (uintptr_t)wf.b.d,
            (uintptr_t)wf.b.beta, (uintptr_t)wf.b.info, (uintptr_t)wf.hit, (uintptr_t)wf.perm, (uintptr_t)wf.keys,
            (uintptr_t)wf.bhist, (uintptr_t)wf.sh.o, (uintptr_t)wf.sh.d, (uintptr_t)wf.sh.c, (uintptr_t)wf.sh.aux,
            (uintptr_t)wf.overflow, (uintptr_t)wf.ctr, (uintptr_t)wf.job_done, (uintptr_t)wf.capacity,
            (uintptr_t)p.block, (uintptr_t)p.sort_keys, (uintptr_t)p.closest_shadow, (uintptr_t)p.uscene,
            (uintptr_t)p.ext_features, (uintptr_t)p.items};
        for (int it = 0;; it += poll) {
            if (it >= max_iterations) return fail(DFR_E_DEVICE, "wavefront did not drain");
            std::vector<uintptr_t> k2 = key;
            k2.push_back((uintptr_t)p.queue_bound);
            k2.push_back((uintptr_t)p.generate);
            if ((rc = replay_iterations(k2, scenes, jobs, n_jobs, film_ptr, p, poll))) return rc;
            iterations = it + poll;
            hipError_t e = hipMemcpyAsync(&host, wf.ctr, sizeof(Counters), hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return hip_fail(e, "wavefront counters");
            if (host.n_next == 0 && host.work_next >= host.work_total) break;
            if (stop.load()) return fail(DFR_E_STOPPED, "render stopped");
            bound_grids(host);
        }
    }
    for (int it = 0; iterations == 0; it++) {
        if (it >= max_iterations) return fail(DFR_E_DEVICE, "wavefront did not drain");
        IterProfile prof{};
        prof.timing = profile != 0;
        prof.trav = trav;
        if (prof.timing) {
            while (prof_events.size() < (size_t)6 * (it + 1)) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) return fail(DFR_E_DEVICE, "event create");
                prof_events.push_back(e);
            }
            for (int k = 0; k < 6; k++) prof.ev[k] = prof_events[6 * it + k];
        }
        wavefront_iteration(scenes, jobs, n_jobs, film_ptr, wf, p, (prof.timing || trav) ? &prof : nullptr, stream);
        if ((it + 1) % poll == 0) {
            hipError_t e = hipMemcpyAsync(&host, wf.ctr, sizeof(Counters), hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return hip_fail(e, "wavefront counters");
            // after an iteration, the paths of the next queue wait in n_next
            if (host.n_next == 0 && host.work_next >= host.work_total) { iterations = it + 1; break; }
            if (stop.load()) return fail(DFR_E_STOPPED, "render stopped");
            bound_grids(host);
        }
    }
    prof_stats = dfr_stats{};
    if (profile) {
        hipStreamSynchronize(stream);
        for (int it = 0; it < iterations; it++) {
            float a = 0, b = 0, d = 0;
            hipEventElapsedTime(&a, prof_events[6 * it + 0], prof_events[6 * it + 1]);
            hipEventElapsedTime(&b, prof_events[6 * it + 2], prof_events[6 * it + 3]);
            hipEventElapsedTime(&d, prof_events[6 * it + 4], prof_events[6 * it + 5]);
            prof_stats.extend_seconds += a * 1e-3;
            prof_stats.shade_seconds += b * 1e-3;
            prof_stats.shadow_seconds += d * 1e-3;
        }
        prof_stats.extend_launches = iterations;
    }
    if (trav) {
        unsigned long long v[5];
        hipMemcpy(v, trav, 40, hipMemcpyDeviceToHost);
        prof_stats.extend_nodes = (int64_t)v[0];
        prof_stats.extend_tris = (int64_t)v[1];
        prof_stats.shadow_nodes = (int64_t)v[2];
        prof_stats.shadow_tris = (int64_t)v[3];
        sample_count = wf.sample.rec ? std::min<int64_t>((int64_t)v[4], wf.sample.cap) : 0;
        wf.sample = RaySample{};
    }
    prof_stats.extend_rays = host.bounces;
    prof_stats.shadow_rays = host.rays_closest + host.rays_any - host.bounces;
    if (out_ctr) *out_ctr = host;
    return DFR_OK;
}

int Ctx::upload_jobs(const std::vector<DevJob>& jobs) {
    jobs_scene = jobs.empty() ? -1 : jobs[0].scene;
    for (const DevJob& j : jobs)
        if (j.scene != jobs_scene) jobs_scene = -1;
    if (job_buf.alloc(std::max<size_t>(1, jobs.size()) * sizeof(DevJob)) != hipSuccess) return fail(DFR_E_MEMORY, "job table");
    hipError_t e = hipMemcpyAsync(job_buf.ptr, jobs.data(), jobs.size() * sizeof(DevJob), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_fail(e, "job upload");
    return DFR_OK;
}

DevJob Ctx::main_job(int spp, int sample_base) const {
    DevJob j{};
    j.seed = seed;
    j.width = width; j.height = height;
    j.spp = spp; j.sample_base = sample_base;
    j.max_bounce = max_bounce; j.min_rr = min_rr;
    j.algorithm = algorithm; j.ao_samples = ao_samples;
    j.max_distance = max_distance;
    j.preview = preview;
    j.film_offset = 0;
    j.work_begin = 0;
    j.work_count = (int64_t)width * height * spp;
    std::memcpy(j.cam, camera, sizeof(j.cam));
    j.cam[DFR_CAMERA_RES_X] = (float)width;
    j.cam[DFR_CAMERA_RES_Y] = (float)height;
    j.scene = main_scene;
    return j;
}

}  // namespace dfr

// ===========================================================================
// extern "C" entry points
// ===========================================================================
#define CTX_GUARD(ctx)                                  \
    if (!(ctx)) return DFR_E_NULL;                      \
    std::lock_guard<std::mutex> _lock((ctx)->mutex);     \
    hipSetDevice((ctx)->device)

extern "C" {

dfr_ctx* dfr_create(int32_t device) {
    int dev = device;
    if (dev < 0 && hipGetDevice(&dev) != hipSuccess) return nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || dev >= count) return nullptr;
    if (hipSetDevice(dev) != hipSuccess) return nullptr;
    Ctx* c = new Ctx();
    c->device = dev;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return nullptr; }
    // the address of this device's copy of the no-allocation scratch array
    // (dfr_trace_rays_noalloc reports a failure here when it is called)
    c->scratch_status = noalloc_scratch_init(dev);
    hipEventCreate(&c->ev0);
    hipEventCreate(&c->ev1);
    return reinterpret_cast<dfr_ctx*>(c);
}

void dfr_destroy(dfr_ctx* h) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    if (!c) return;
    {
        std::lock_guard<std::mutex> lock(c->mutex);
        hipSetDevice(c->device);
        hipStreamSynchronize(c->stream);
    }
    hipSetDevice(c->device);
    c->scenes.clear();
    for (hipEvent_t e : c->prof_events) hipEventDestroy(e);
    for (Ctx::GraphEntry& g : c->graphs) hipGraphExecDestroy(g.exec);
    hipEventDestroy(c->ev0);
    hipEventDestroy(c->ev1);
    hipStreamDestroy(c->stream);
    delete c;
}

const char* dfr_last_error(const dfr_ctx* h) {
    const Ctx* c = reinterpret_cast<const Ctx*>(h);
    return c ? c->error.c_str() : "context == null";
}

// setScene with the context's lock held
static int32_t set_scene_locked(Ctx* c, const dfr_scene_desc* scene) {
    if (!scene) return c->fail(DFR_E_NULL, "scene == null");
    int rc = copy_scene_desc(scene, c->scene_host, c->error);
    if (rc) return rc;
    c->xform_host.clear();
    c->has_scene = true;
    std::memcpy(c->camera, scene->camera, sizeof(c->camera));
    c->scene_dirty = true;
    return DFR_OK;
}

int32_t dfr_set_scene(dfr_ctx* h, const dfr_scene_desc* scene) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    return set_scene_locked(c, scene);
}

// updateTransforms with the context's lock held
static int32_t update_transforms_locked(Ctx* c, const float* object_to_world, int32_t n_primitives, int32_t rebuild) {
    if (!c->has_scene) return c->fail(DFR_E_STATE, "setScene before updateTransforms");
    if (!object_to_world) return c->fail(DFR_E_NULL, "object_to_world == null");
    if (n_primitives != (int32_t)c->scene_host.prims.size())
        return c->fail(DFR_E_ARGUMENT, "one matrix per primitive of the scene: " + std::to_string(c->scene_host.prims.size()));
    std::vector<float> m(object_to_world, object_to_world + 16 * (size_t)n_primitives);
    std::vector<Xform> xf;
    int rc = make_xforms(c->scene_host, m, xf, c->error);
    if (rc) return rc;
    const bool compiled = !c->scene_dirty && c->main_scene >= 0 && c->scenes[c->main_scene];
    if (compiled) {
        const int ms = c->main_scene;
        if (ms < (int)c->scene_refs.size() && c->scene_refs[ms] > 0)
            return c->fail(DFR_E_STATE, "a serving loop still has requests on this scene");
        if (hipStreamSynchronize(c->stream) != hipSuccess) return c->fail(DFR_E_DEVICE, "stream synchronisation failed");
        if (rebuild) {
            auto cs = std::make_unique<CompiledScene>();
            if ((rc = compile_scene(c->scene_host, xf, *cs, c->light_pdf_mode, c->stream, c->error))) return rc;
            c->scenes[ms] = std::move(cs);
            if ((rc = c->sync_scene_table())) return rc;
        } else if ((rc = refit_scene(c->scene_host, xf, *c->scenes[ms], c->stream, c->error))) {
            // the device records may be half rewritten: compile afresh at the next setup
            c->scene_dirty = true;
            return rc;
        }
    }
    c->xform_host = std::move(m);
    return DFR_OK;
}

int32_t dfr_update_transforms(dfr_ctx* h, const float* object_to_world, int32_t n_primitives, int32_t rebuild) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    return update_transforms_locked(c, object_to_world, n_primitives, rebuild);
}

int32_t dfr_update_matrix44fs(dfr_ctx* h, const float* matrices, int64_t length) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (!c->has_scene) return c->fail(DFR_E_STATE, "setScene before updateMatrix44Fs");
    if (!matrices) return c->fail(DFR_E_NULL, "matrices == null");
    if (length != 32 * (int64_t)c->scene_host.prims.size())
        return c->fail(DFR_E_ARGUMENT, "32 floats per primitive of the scene: " +
                                           std::to_string(32 * c->scene_host.prims.size()));
    std::vector<float> o2w(16 * (size_t)(length / 32));
    for (int64_t p = 0; p < length / 32; p++) std::memcpy(&o2w[16 * p], matrices + 32 * p, 16 * sizeof(float));
    return update_transforms_locked(c, o2w.data(), (int32_t)(length / 32), 0);
}

int32_t dfr_set_compiled_scene(dfr_ctx* h, const dfr_compiled_scene* scene) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (!scene) return c->fail(DFR_E_NULL, "scene == null");
    const dfr_scene_desc* desc = nullptr;
    const float* xf = nullptr;
    void* holder = nullptr;
    std::string err;
    int rc = transcode_compiled_scene(scene, &desc, &xf, &holder, err);
    if (rc) return c->fail(rc, err);
    rc = set_scene_locked(c, desc);
    if (!rc && desc->n_primitives > 0) rc = update_transforms_locked(c, xf, desc->n_primitives, 0);
    free_transcoded(holder);
    return rc;
}

int32_t dfr_set_resolution(dfr_ctx* h, int32_t w, int32_t hh) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (w < 1 || hh < 1) return c->fail(DFR_E_ARGUMENT, "resolution must be positive");
    if (w != c->width || hh != c->height) { c->width = w; c->height = hh; c->film_dirty = true; }
    return DFR_OK;
}

int32_t dfr_set_i32(dfr_ctx* h, int32_t key, int32_t v) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    switch (key) {
        case DFR_KEY_MAXIMUM_BOUNCE: c->max_bounce = v; break;
        case DFR_KEY_MINIMUM_BOUNCE_RR: c->min_rr = v; break;
        case DFR_KEY_RENDERING_ALGORITHM:
            if (v < 0 || v > 4) return c->fail(DFR_E_ARGUMENT, "rendering algorithm ordinal out of range");
            c->algorithm = v; break;
        case DFR_KEY_SAMPLES: c->ao_samples = v; break;
        case DFR_KEY_TONE_MAPPER: c->tone_mapper = v; break;
        case DFR_KEY_PREVIEW_MODE: c->preview = v; break;
        case DFR_KEY_SEED: c->seed = (c->seed & 0xffffffff00000000ull) | (uint32_t)v; break;
        case DFR_KEY_SEED_HI: c->seed = (c->seed & 0xffffffffull) | ((uint64_t)(uint32_t)v << 32); break;
        case DFR_KEY_KERNEL: c->kernel = v; break;
        case DFR_KEY_LIGHT_PDF_MODE:
            if (v < 0 || v > 1) return c->fail(DFR_E_ARGUMENT, "light pdf mode is 0 or 1");
            if (v != c->light_pdf_mode) { c->light_pdf_mode = v; c->scene_dirty = true; }
            break;
        case DFR_KEY_SPP_PER_PASS:
            if (v < 1) return c->fail(DFR_E_ARGUMENT, "spp per pass must be >= 1");
            c->spp_per_pass = v; break;
        case DFR_KEY_PROFILE: c->profile = v != 0; break;
        case DFR_KEY_COUNT_TRAVERSAL: c->count_traversal = v != 0; break;
        case DFR_KEY_WF_BLOCK:
            if (v != 128 && v != 256) return c->fail(DFR_E_ARGUMENT, "wavefront block is 128 or 256");
            c->wf_params.block = v; break;
        case DFR_KEY_WF_POOL_PATHS:
            if (v < 1024 || v > MAX_POOL_PATHS) return c->fail(DFR_E_ARGUMENT, "path pool must hold 1024 .. 2^26 paths");
            c->wf_params.capacity = v; break;
        case DFR_KEY_WF_SORT_KEYS:
            if (v != 0 && v != 1) return c->fail(DFR_E_ARGUMENT, "sort is 0 (off) or 1 (on)");
            c->wf_params.sort_keys = v; break;
        case DFR_KEY_USE_GRAPHS: c->use_graphs = v != 0; break;
        case DFR_KEY_RAY_SAMPLE:
            if (v < 0) return c->fail(DFR_E_ARGUMENT, "ray sample stride must be >= 0");
            c->ray_sample_stride = v; break;
        case DFR_KEY_FILM_FILTER:
            if (v != 0 && v != 1) return c->fail(DFR_E_ARGUMENT, "film filter is 0 (running mean) or 1 (Gaussian XYZ)");
            if (v != c->film_filter) c->clearing = true;  // both films start again
            c->film_filter = v; break;
        default: return c->fail(DFR_E_ARGUMENT, "unknown integer key");
    }
    return DFR_OK;
}

int32_t dfr_get_i32(const dfr_ctx* h, int32_t key, int32_t* v) {
    const Ctx* c = reinterpret_cast<const Ctx*>(h);
    if (!c || !v) return DFR_E_NULL;
    switch (key) {
        case DFR_KEY_MAXIMUM_BOUNCE: *v = c->max_bounce; break;
        case DFR_KEY_MINIMUM_BOUNCE_RR: *v = c->min_rr; break;
        case DFR_KEY_RENDERING_ALGORITHM: *v = c->algorithm; break;
        case DFR_KEY_SAMPLES: *v = c->ao_samples; break;
        case DFR_KEY_TONE_MAPPER: *v = c->tone_mapper; break;
        case DFR_KEY_PREVIEW_MODE: *v = c->preview; break;
        case DFR_KEY_SEED: *v = (int32_t)(uint32_t)c->seed; break;
        case DFR_KEY_SEED_HI: *v = (int32_t)(uint32_t)(c->seed >> 32); break;
        case DFR_KEY_KERNEL: *v = c->kernel; break;
        case DFR_KEY_LIGHT_PDF_MODE: *v = c->light_pdf_mode; break;
        case DFR_KEY_SPP_PER_PASS: *v = c->spp_per_pass; break;
        case DFR_KEY_PROFILE: *v = c->profile; break;
        case DFR_KEY_COUNT_TRAVERSAL: *v = c->count_traversal; break;
        case DFR_KEY_WF_BLOCK: *v = c->wf_params.block; break;
        case DFR_KEY_WF_POOL_PATHS: *v = c->wf_params.capacity; break;
        case DFR_KEY_WF_SORT_KEYS: *v = c->wf_params.sort_keys; break;
        case DFR_KEY_BVH_WIDTH: *v = BVH_WIDTH; break;
        case DFR_KEY_USE_GRAPHS: *v = c->use_graphs; break;
        case DFR_KEY_RAY_SAMPLE: *v = c->ray_sample_stride; break;
        case DFR_KEY_FILM_FILTER: *v = c->film_filter; break;
        default: return DFR_E_ARGUMENT;
    }
    return DFR_OK;
}

int32_t dfr_set_f32(dfr_ctx* h, int32_t key, float v) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (key != DFR_KEY_MAXIMUM_DISTANCE) return c->fail(DFR_E_ARGUMENT, "unknown float key");
    c->max_distance = v;
    return DFR_OK;
}

int32_t dfr_get_f32(const dfr_ctx* h, int32_t key, float* v) {
    const Ctx* c = reinterpret_cast<const Ctx*>(h);
    if (!c || !v) return DFR_E_NULL;
    if (key != DFR_KEY_MAXIMUM_DISTANCE) return DFR_E_ARGUMENT;
    *v = c->max_distance;
    return DFR_OK;
}

static int32_t setup_impl(Ctx* c, int32_t compile_scene_flag) {
    if (compile_scene_flag || c->scene_dirty) {
        if (!c->has_scene) return c->fail(DFR_E_STATE, "setScene before setup");
        std::vector<Xform> xf;
        int rc = make_xforms(c->scene_host, c->xform_host, xf, c->error);
        if (rc) return rc;
        auto cs = std::make_unique<CompiledScene>();
        rc = compile_scene(c->scene_host, xf, *cs, c->light_pdf_mode, c->stream, c->error);
        if (rc) return rc;
        if (c->main_scene < 0) {
            c->main_scene = (int)c->scenes.size();
            c->scenes.push_back(std::move(cs));
        } else {
            c->scenes[c->main_scene] = std::move(cs);
        }
        if ((rc = c->sync_scene_table())) return rc;
        c->scene_dirty = false;
    }
    int rc = c->ensure_film((int64_t)c->width * c->height);
    if (rc) return rc;
    hipError_t e = hipMemsetAsync(c->film.ptr, 0, (size_t)c->width * c->height * sizeof(float4), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "film clear");
    c->xyz_pixels = -1;
    c->render_pass = 0;
    c->clearing = false;
    c->film_dirty = false;
    c->is_setup = true;
    return DFR_OK;
}

int32_t dfr_setup(dfr_ctx* h, int32_t compile_scene_flag) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    return setup_impl(c, compile_scene_flag);
}

int32_t dfr_render(dfr_ctx* h, int32_t passes) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (passes < 1) return c->fail(DFR_E_ARGUMENT, "passes must be >= 1");
    if (!c->is_setup || c->main_scene < 0) return c->fail(DFR_E_STATE, "setup before render");
    if (c->scene_dirty || c->film_dirty) {
        int rc = setup_impl(c, 0);
        if (rc) return rc;
    }
    const int64_t npix = (int64_t)c->width * c->height;
    hipError_t e;
    if (c->clearing) {
        e = hipMemsetAsync(c->film.ptr, 0, (size_t)npix * sizeof(float4), c->stream);
        if (e != hipSuccess) return c->hip_fail(e, "film clear");
        c->xyz_pixels = -1;  // the XYZ film is cleared when next used
        c->render_pass = 0;
        c->clearing = false;
    }
    c->stop.store(false);
    const int spp = passes * c->spp_per_pass;
    const int first_sample = c->render_pass * c->spp_per_pass;
    if (c->counters.alloc(128) != hipSuccess) return c->fail(DFR_E_MEMORY, "counters");
    hipMemsetAsync(c->counters.ptr, 0, 64, c->stream);
    dfr_stats st{};
    bool plain = c->kernel == 1 || c->algorithm != DFR_ALGO_PATH_TRACING;
    int rc;
    if (plain) {
        int64_t threads = ((npix + 255) / 256) * 256;
        if ((rc = c->ensure_overflow(threads))) return rc;
    }
    // Renders samples [base, base + n) into `target`: the (sum, count) film,
    // or per-sample records (sample_film) of a chunk starting at base.
    auto run = [&](int n, int base, float4* target, bool sample_film) -> int {
        if (plain) {
            // one launch per pass: each launch adds spp_per_pass samples to every pixel
            for (int p = 0; p < n / c->spp_per_pass; p++) {
                DevJob jp = c->main_job(c->spp_per_pass, base + p * c->spp_per_pass);
                jp.sample_film = sample_film ? 1 : 0;
                jp.film_sample_base = base;
                hipError_t ee = hipMemcpyAsync(c->job_buf.ptr, &jp, sizeof(DevJob), hipMemcpyHostToDevice, c->stream);
                if (ee != hipSuccess) return c->hip_fail(ee, "job upload");
                launch_megakernel((const DevScene*)c->scene_table.ptr, (const DevJob*)c->job_buf.ptr, target,
                                  (int*)c->overflow.ptr, (int)npix, 256, c->any_ext_features(), (int64_t*)c->counters.ptr,
                                  c->stream);
                if (c->stop.load()) break;
            }
            return DFR_OK;
        }
        DevJob job = c->main_job(n, base);
        job.sample_film = sample_film ? 1 : 0;
        job.film_sample_base = base;
        int r2 = c->upload_jobs({job});
        if (r2) return r2;
        Counters out{};
        if ((r2 = c->run_wavefront(1, job.work_count, target, &out))) return r2;
        st.rays_closest += out.rays_closest; st.rays_any += out.rays_any; st.bounces += out.bounces;
        st.extend_seconds += c->prof_stats.extend_seconds; st.shade_seconds += c->prof_stats.shade_seconds;
        st.shadow_seconds += c->prof_stats.shadow_seconds; st.extend_launches += c->prof_stats.extend_launches;
        st.extend_rays += c->prof_stats.extend_rays; st.extend_nodes += c->prof_stats.extend_nodes;
        st.extend_tris += c->prof_stats.extend_tris; st.shadow_rays += c->prof_stats.shadow_rays;
        st.shadow_nodes += c->prof_stats.shadow_nodes; st.shadow_tris += c->prof_stats.shadow_tris;
        return DFR_OK;
    };
    if (c->upload_jobs({c->main_job(c->spp_per_pass, first_sample)})) return c->fail(DFR_E_MEMORY, "job table");
    hipEventRecord(c->ev0, c->stream);
    if (c->film_filter) {
        // CPURenderer's Gaussian film: render chunks of samples into
        // per-sample records, then gather each chunk into the XYZ film
        if ((rc = c->ensure_filtered_film())) return rc;
        const int64_t rec = npix * (int64_t)sizeof(float4);
        int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(spp, SAMPLE_FILM_BYTES / rec));
        chunk = std::max(c->spp_per_pass, chunk / c->spp_per_pass * c->spp_per_pass);
        if (c->sample_records.alloc((size_t)rec * chunk) != hipSuccess) return c->fail(DFR_E_MEMORY, "sample records");
        for (int done = 0; done < spp && !c->stop.load(); done += chunk) {
            const int n = std::min(chunk, spp - done);
            hipError_t ee = hipMemsetAsync(c->sample_records.ptr, 0, (size_t)rec * n, c->stream);
            if (ee != hipSuccess) return c->hip_fail(ee, "sample records clear");
            if ((rc = run(n, first_sample + done, (float4*)c->sample_records.ptr, true))) return rc;
            launch_film_gauss((const float4*)c->sample_records.ptr, c->width, c->height, n, c->seed, first_sample + done,
                              (const FilmConsts*)c->film_consts.ptr, (float4*)c->xyz_film.ptr, (float4*)c->film.ptr,
                              c->stream);
        }
    } else if ((rc = run(spp, first_sample, (float4*)c->film.ptr, false))) {
        return rc;
    }
    hipEventRecord(c->ev1, c->stream);
    e = hipEventSynchronize(c->ev1);
    if (e != hipSuccess) return c->hip_fail(e, plain ? "megakernel" : "wavefront");
    if (plain) {
        int64_t cnt[2];
        hipMemcpy(cnt, c->counters.ptr, sizeof(cnt), hipMemcpyDeviceToHost);
        st.rays_closest = cnt[0]; st.rays_any = cnt[1];
    }
    float ms = 0.0f;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    st.seconds = ms * 1e-3;
    st.path_samples = npix * spp;
    c->stats = st;
    c->render_pass += passes;
    return DFR_OK;
}

int32_t dfr_clear(dfr_ctx* h) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    c->clearing = true;
    return DFR_OK;
}

int32_t dfr_get_render_pass(const dfr_ctx* h) {
    const Ctx* c = reinterpret_cast<const Ctx*>(h);
    return c ? c->render_pass : DFR_E_NULL;
}

int32_t dfr_request_stop(dfr_ctx* h) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    if (!c) return DFR_E_NULL;
    c->stop.store(true);
    return DFR_OK;
}

int32_t dfr_update_camera(dfr_ctx* h, const float* cam) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (!cam) return c->fail(DFR_E_NULL, "camera == null");
    std::memcpy(c->camera, cam, sizeof(c->camera));
    return DFR_OK;
}

int32_t dfr_read_film(dfr_ctx* h, float* rgb, int32_t* count) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (!c->is_setup) return c->fail(DFR_E_STATE, "setup before read");
    const int64_t n = (int64_t)c->width * c->height;
    std::vector<float4> f(n);
    hipError_t e = hipMemcpyAsync(f.data(), c->film.ptr, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "film readback");
    for (int64_t i = 0; i < n; i++) {
        float w = f[i].w;
        float inv = w > 0.0f ? 1.0f / w : 0.0f;
        if (rgb) { rgb[3 * i] = f[i].x * inv; rgb[3 * i + 1] = f[i].y * inv; rgb[3 * i + 2] = f[i].z * inv; }
        if (count) count[i] = (int32_t)w;
    }
    return DFR_OK;
}

int32_t dfr_read_film_xyz(dfr_ctx* h, float* xyzw) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (!xyzw) return c->fail(DFR_E_NULL, "xyzw == null");
    if (!c->is_setup) return c->fail(DFR_E_STATE, "setup before read");
    if (!c->film_filter) return c->fail(DFR_E_STATE, "the Gaussian film is off (DFR_KEY_FILM_FILTER)");
    int rc = c->ensure_filtered_film();
    if (rc) return rc;
    const int64_t n = (int64_t)c->width * c->height;
    hipError_t e = hipMemcpyAsync(xyzw, c->xyz_film.ptr, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "XYZ film readback");
    return DFR_OK;
}

int32_t dfr_read_film_filtered(dfr_ctx* h, float* rgb) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (!rgb) return c->fail(DFR_E_NULL, "rgb == null");
    if (!c->is_setup) return c->fail(DFR_E_STATE, "setup before read");
    if (!c->film_filter) return c->fail(DFR_E_STATE, "the Gaussian film is off (DFR_KEY_FILM_FILTER)");
    int rc = c->ensure_filtered_film();
    if (rc) return rc;
    const int64_t n = (int64_t)c->width * c->height;
    if (c->film_rgb.alloc((size_t)n * 3 * sizeof(float)) != hipSuccess) return c->fail(DFR_E_MEMORY, "film RGB");
    launch_film_rgb((const float4*)c->xyz_film.ptr, (float*)c->film_rgb.ptr, (int)n, (const FilmConsts*)c->film_consts.ptr,
                    c->stream);
    hipError_t e = hipMemcpyAsync(rgb, c->film_rgb.ptr, n * 3 * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "filtered film readback");
    return DFR_OK;
}

int32_t dfr_read_rgba8(dfr_ctx* h, uint8_t* rgba) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (!rgba) return c->fail(DFR_E_NULL, "rgba == null");
    if (!c->is_setup) return c->fail(DFR_E_STATE, "setup before read");
    const int64_t n = (int64_t)c->width * c->height;
    if (c->image.alloc(n * 4) != hipSuccess) return c->fail(DFR_E_MEMORY, "rgba8");
    // GPURenderer.java:312-335: only doRunPathTracing tone-maps
    const int tone = c->algorithm == DFR_ALGO_PATH_TRACING ? c->tone_mapper : DFR_TONE_MAPPER_NONE;
    launch_rgba8((const float4*)c->film.ptr, (uint8_t*)c->image.ptr, (int)n, tone, c->stream);
    hipError_t e = hipMemcpyAsync(rgba, c->image.ptr, n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "rgba8 readback");
    return DFR_OK;
}

int32_t dfr_get_stats(const dfr_ctx* h, dfr_stats* st) {
    const Ctx* c = reinterpret_cast<const Ctx*>(h);
    if (!c || !st) return DFR_E_NULL;
    *st = c->stats;
    return DFR_OK;
}

int32_t dfr_trace_rays(dfr_ctx* h, const float* rays, int64_t n, int32_t any_hit, int32_t* prim, int64_t* tri, float* t,
                       int32_t device_ptrs) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (n < 0) return c->fail(DFR_E_ARGUMENT, "n < 0");
    if (n == 0) return DFR_OK;
    if (!rays || !prim || !tri || !t) return c->fail(DFR_E_NULL, "null pointer");
    if (c->main_scene < 0) return c->fail(DFR_E_STATE, "setup before trace");
    const int grid = std::min<int64_t>((n + 255) / 256, 256 * 64);
    int rc = c->ensure_overflow((int64_t)grid * 256);
    if (rc) return rc;
    const DevScene* sc = (const DevScene*)c->scene_table.ptr + c->main_scene;
    hipError_t e;
    hipEventRecord(c->ev0, c->stream);
    if (device_ptrs) {
        launch_trace_rays(sc, (const float4*)rays, n, any_hit, prim, tri, t, (int*)c->overflow.ptr, 256, grid, c->stream);
        hipEventRecord(c->ev1, c->stream);
        e = hipEventSynchronize(c->ev1);
        if (e != hipSuccess) return c->hip_fail(e, "trace");
    } else {
        DeviceBuffer r, p, tr, tt;
        if (r.alloc(n * 32) || p.alloc(n * 4) || tr.alloc(n * 8) || tt.alloc(n * 4)) return c->fail(DFR_E_MEMORY, "ray buffers");
        e = hipMemcpyAsync(r.ptr, rays, n * 32, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) return c->hip_fail(e, "ray upload");
        hipEventRecord(c->ev0, c->stream);
        launch_trace_rays(sc, (const float4*)r.ptr, n, any_hit, (int32_t*)p.ptr, (int64_t*)tr.ptr, (float*)tt.ptr,
                          (int*)c->overflow.ptr, 256, grid, c->stream);
        hipEventRecord(c->ev1, c->stream);
        hipMemcpyAsync(prim, p.ptr, n * 4, hipMemcpyDeviceToHost, c->stream);
        hipMemcpyAsync(tri, tr.ptr, n * 8, hipMemcpyDeviceToHost, c->stream);
        hipMemcpyAsync(t, tt.ptr, n * 4, hipMemcpyDeviceToHost, c->stream);
        e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return c->hip_fail(e, "trace");
    }
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->stats = dfr_stats{};
    c->stats.seconds = ms * 1e-3;
    if (any_hit) c->stats.rays_any = n; else c->stats.rays_closest = n;
    return DFR_OK;
}

int32_t dfr_read_ray_sample(dfr_ctx* h, float* records, int64_t max_records, int64_t* n_records) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (!n_records) return c->fail(DFR_E_NULL, "n_records == null");
    const int64_t n = records ? std::min(max_records, c->sample_count) : c->sample_count;
    *n_records = n;
    if (!records || n <= 0) return DFR_OK;
    const hipError_t e = hipMemcpy(records, c->sample_buf.ptr, (size_t)n * 3 * sizeof(float4), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, "ray sample readback");
    return DFR_OK;
}

int32_t dfr_read_bvh(dfr_ctx* h, float* nodes, int64_t* n_nodes, float* tris, int64_t* n_tris) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (!n_nodes || !n_tris) return c->fail(DFR_E_NULL, "count pointer == null");
    if (c->main_scene < 0 || !c->scenes[c->main_scene]) return c->fail(DFR_E_STATE, "setup before reading the BVH");
    const CompiledScene& cs = *c->scenes[c->main_scene];
    if (BVH_WIDTH != 4) return c->fail(DFR_E_UNSUPPORTED, "the BVH export describes four-wide nodes");
    const int64_t want_nodes = *n_nodes, want_tris = *n_tris;
    *n_nodes = cs.n_nodes;
    *n_tris = cs.n_bvh_tris;
    hipError_t e = hipSuccess;
    if (nodes && cs.n_nodes > 0) {
        if (want_nodes < cs.n_nodes) return c->fail(DFR_E_ARGUMENT, "node buffer too small");
        e = hipMemcpy(nodes, cs.nodes.ptr, (size_t)cs.n_nodes * 128, hipMemcpyDeviceToHost);
    }
    if (e == hipSuccess && tris && cs.n_bvh_tris > 0) {
        if (want_tris < cs.n_bvh_tris) return c->fail(DFR_E_ARGUMENT, "triangle buffer too small");
        e = hipMemcpy(tris, cs.tris.ptr, (size_t)cs.n_bvh_tris * 48, hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) return c->hip_fail(e, "BVH readback");
    return DFR_OK;
}

int32_t dfr_trace_rays_noalloc(dfr_ctx* h, const float* rays, int64_t n, int32_t any_hit, int32_t* prim, int64_t* tri,
                               float* t) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (n < 0) return c->fail(DFR_E_ARGUMENT, "n < 0");
    if (n == 0) return DFR_OK;
    if (!rays || !prim || !tri || !t) return c->fail(DFR_E_NULL, "null pointer");
    if (c->main_scene < 0) return c->fail(DFR_E_STATE, "setup before trace");
    if (c->scratch_status != (int)hipSuccess) return c->hip_fail((hipError_t)c->scratch_status, "scratch address");
    int* scratch_ptr = noalloc_scratch(c->device);
    if (!scratch_ptr) return c->fail(DFR_E_DEVICE, "scratch address unknown on this device");
    // the deepest traversal stack of the scene must fit the static overflow
    const int need = overflow_entries_per_thread(c->scenes[c->main_scene]->max_depth);
    if (need > noalloc_stack_entries())
        return c->fail(DFR_E_UNSUPPORTED, "BVH too deep for the no-allocation trace: needs " + std::to_string(need) +
                                              " overflow stack entries per thread, the static scratch holds " +
                                              std::to_string(noalloc_stack_entries()));
    const DevScene* sc = (const DevScene*)c->scene_table.ptr + c->main_scene;
    hipEventRecord(c->ev0, c->stream);
    launch_trace_rays_noalloc(sc, (const float4*)rays, n, any_hit, prim, tri, t, scratch_ptr, c->stream);
    hipEventRecord(c->ev1, c->stream);
    const hipError_t e = hipEventSynchronize(c->ev1);
    if (e != hipSuccess) return c->hip_fail(e, "trace");
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->stats = dfr_stats{};
    c->stats.seconds = ms * 1e-3;
    if (any_hit) c->stats.rays_any = n; else c->stats.rays_closest = n;
    return DFR_OK;
}

int32_t dfr_cast_primary(dfr_ctx* h, int32_t sample_index, float* rays, int32_t* prim, int64_t* tri, float* t) {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_GUARD(c);
    if (!rays || !prim || !tri || !t) return c->fail(DFR_E_NULL, "null pointer");
    if (c->main_scene < 0 || !c->is_setup) return c->fail(DFR_E_STATE, "setup before cast");
    const int64_t n = (int64_t)c->width * c->height;
    int rc = c->ensure_overflow(((n + 255) / 256) * 256);
    if (rc) return rc;
    DevJob job = c->main_job(1, sample_index);
    if ((rc = c->upload_jobs({job}))) return rc;
    DeviceBuffer r, p, tr, tt;
    if (r.alloc(n * 32) || p.alloc(n * 4) || tr.alloc(n * 8) || tt.alloc(n * 4)) return c->fail(DFR_E_MEMORY, "ray buffers");
    launch_cast_primary((const DevScene*)c->scene_table.ptr + c->main_scene, (const DevJob*)c->job_buf.ptr, (float*)r.ptr,
                        (int32_t*)p.ptr, (int64_t*)tr.ptr, (float*)tt.ptr, (int*)c->overflow.ptr, (int)n, c->stream);
    hipMemcpyAsync(rays, r.ptr, n * 32, hipMemcpyDeviceToHost, c->stream);
    hipMemcpyAsync(prim, p.ptr, n * 4, hipMemcpyDeviceToHost, c->stream);
    hipMemcpyAsync(tri, tr.ptr, n * 8, hipMemcpyDeviceToHost, c->stream);
    hipMemcpyAsync(t, tt.ptr, n * 4, hipMemcpyDeviceToHost, c->stream);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "cast primary");
    return DFR_OK;
}

}  // extern "C"
