// perez_host.cpp — PerezLight setup on the host: sun colour, Perez
// coefficients, zenith values, daylight basis and the 32x32 importance
// distribution of the sky (M/scene/light/PerezLight.java set(...),
// doInitializeDistribution; M/sampler/Distribution1F.java, Distribution2F.java).
#include "perez_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "spectral_tables.inc"

namespace dfr {
namespace {

const float PI_F = 3.14159265358979323846f;

// Linear interpolation on a regular wavelength grid (RegularSpectralCurveF).
float regular_sample(const float* a, int n, float lmin, float lmax, float l) {
    if (l < lmin || l > lmax) return 0.0f;
    float delta = (lmax - lmin) / (float)(n - 1);
    float x = (l - lmin) * (1.0f / delta);
    int i0 = (int)x;
    int i1 = std::min(i0 + 1, n - 1);
    float dx = x - (float)i0;
    return (1.0f - dx) * a[i0] + dx * a[i1];
}

// Piecewise linear on given wavelengths (IrregularSpectralCurveF).
float irregular_sample(const float* a, const float* wl, int n, float l) {
    if (n == 0) return 0.0f;
    if (n == 1 || l <= wl[0]) return a[0];
    for (int i = 1; i < n; i++)
        if (l < wl[i]) {
            float dx = (l - wl[i - 1]) / (wl[i] - wl[i - 1]);
            return (1.0f - dx) * a[i - 1] + dx * a[i];
        }
    return a[n - 1];
}

// SpectralCurveF.toColorXYZ: sum over the 1 nm CIE grid.
template <class F>
void to_xyz(F sample, float out[3]) {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    const int n = (int)(sizeof(kCieLambda) / sizeof(kCieLambda[0]));
    for (int i = 0; i < n; i++) {
        float s = sample(kCieLambda[i]);
        x = (float)((double)x + (double)s * kCieX[i]);
        y = (float)((double)y + (double)s * kCieY[i]);
        z = (float)((double)z + (double)s * kCieZ[i]);
    }
    float scale = (kCieLambda[n - 1] - kCieLambda[0]) / (kCieYIntegral * (float)n);
    out[0] = x * scale; out[1] = y * scale; out[2] = z * scale;
}

// sRGB XYZ->RGB matrix from primaries and white point (ColorSpaceF).
void srgb_xyz_to_rgb(float m[9]) {
    const float xR = 0.64f, yR = 0.33f, xG = 0.30f, yG = 0.60f, xB = 0.15f, yB = 0.06f, xW = 0.31271f, yW = 0.32902f;
    float zR = 1 - (xR + yR), zG = 1 - (xG + yG), zB = 1 - (xB + yB), zW = 1 - (xW + yW);
    float rX = yG * zB - yB * zG, rY = xB * zG - xG * zB, rZ = xG * yB - xB * yG;
    float rW = (rX * xW + rY * yW + rZ * zW) / yW;
    float gX = yB * zR - yR * zB, gY = xR * zB - xB * zR, gZ = xB * yR - xR * yB;
    float gW = (gX * xW + gY * yW + gZ * zW) / yW;
    float bX = yR * zG - yG * zR, bY = xG * zR - xR * zG, bZ = xR * yG - xG * yR;
    float bW = (bX * xW + bY * yW + bZ * zW) / yW;
    float e[9] = {rX / rW, rY / rW, rZ / rW, gX / gW, gY / gW, gZ / gW, bX / bW, bY / bW, bZ / bW};
    std::memcpy(m, e, sizeof(e));
}

double perez_fn(const double* lam, double theta, double gamma, double lvz, double sun_theta) {
    double den = (1.0 + lam[0] * std::exp((double)lam[1])) *
                 (1.0 + lam[2] * std::exp(lam[3] * sun_theta) + lam[4] * std::cos(sun_theta) * std::cos(sun_theta));
    double num = (1.0 + lam[0] * std::exp(lam[1] / std::cos(theta))) *
                 (1.0 + lam[2] * std::exp(lam[3] * gamma) + lam[4] * std::cos(gamma) * std::cos(gamma));
    return lvz * num / den;
}

// doRadianceSky for an object-space direction (before minTo0).
void sky(const DevPerez& p, const float M[9], const float d[3], float rgb[3]) {
    if (d[2] < 0.0f) { rgb[0] = rgb[1] = rgb[2] = 0.0f; return; }
    float v[3] = {d[0], d[1], std::max(d[2], 0.001f)};
    float l = (float)std::sqrt((double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    if (!(l >= 0.99999982f && l <= 1.00000012f)) { v[0] /= l; v[1] /= l; v[2] /= l; }
    double theta = std::acos(std::min(std::max((double)v[2], -1.0), 1.0));
    float dp = v[0] * p.sun_dir_object[0] + v[1] * p.sun_dir_object[1] + v[2] * p.sun_dir_object[2];
    double gamma = std::acos(std::min(std::max((double)dp, -1.0), 1.0));
    double rl = perez_fn(p.perez_lum, theta, gamma, p.zenith[0], p.theta) * 0.0001;
    double x = perez_fn(p.perez_x, theta, gamma, p.zenith[1], p.theta);
    double y = perez_fn(p.perez_y, theta, gamma, p.zenith[2], p.theta);
    float fx = (float)x, fy = (float)y;
    float den = 0.0241f + 0.2562f * fx - 0.7341f * fy;
    float m1 = (-1.3515f - 1.7703f * fx + 5.9114f * fy) / den;
    float m2 = (0.03f - 31.4424f * fx + 30.0717f * fy) / den;
    float cx = p.s0[0] + m1 * p.s1[0] + m2 * p.s2[0];
    float cy = p.s0[1] + m1 * p.s1[1] + m2 * p.s2[1];
    float cz = p.s0[2] + m1 * p.s1[2] + m2 * p.s2[2];
    float X = (float)(cx * rl / cy), Y = (float)rl, Z = (float)(cz * rl / cy);
    rgb[0] = M[0] * X + M[1] * Y + M[2] * Z;
    rgb[1] = M[3] * X + M[4] * Y + M[5] * Z;
    rgb[2] = M[6] * X + M[7] * Y + M[8] * Z;
}

// Distribution1F over `f` (n entries): cdf (n+1) and integral.
float dist1(const float* f, int n, float* cdf) {
    cdf[0] = 0.0f;
    for (int i = 1; i <= n; i++) cdf[i] = cdf[i - 1] + f[i - 1] / (float)n;
    float integral = cdf[n];
    if (integral == 0.0f) {
        for (int i = 1; i <= n; i++) cdf[i] = (float)i / (float)n;
    } else {
        for (int i = 1; i <= n; i++) cdf[i] /= integral;
    }
    return integral;
}

}  // namespace

void perez_build(DevPerez& p, float turbidity, const float sun_direction_world[3]) {
    std::memset(&p, 0, sizeof(p));
    float M[9];
    srgb_xyz_to_rgb(M);
    std::memcpy(p.xyz_to_rgb, M, sizeof(M));
    p.radius = 10.0f;
    float sw[3] = {sun_direction_world[0], sun_direction_world[1], sun_direction_world[2]};
    float l = (float)std::sqrt((double)(sw[0] * sw[0] + sw[1] * sw[1] + sw[2] * sw[2]));
    if (!(l >= 0.99999982f && l <= 1.00000012f)) { sw[0] /= l; sw[1] /= l; sw[2] /= l; }
    std::memcpy(p.sun_dir_world, sw, 12);
    // object frame: Matrix44F.rotate(OrthonormalBasis33F(y, x)) has columns
    // u = (0,0,1), v = (1,0,0), w = (0,1,0): object (x,y,z) -> world (y,z,x)
    const float o2w[9] = {0, 1, 0, 0, 0, 1, 1, 0, 0};
    const float w2o[9] = {0, 0, 1, 1, 0, 0, 0, 1, 0};
    std::memcpy(p.object_to_world, o2w, sizeof(o2w));
    std::memcpy(p.world_to_object, w2o, sizeof(w2o));
    float so[3] = {w2o[0] * sw[0] + w2o[1] * sw[1] + w2o[2] * sw[2], w2o[3] * sw[0] + w2o[4] * sw[1] + w2o[5] * sw[2],
                   w2o[6] * sw[0] + w2o[7] * sw[1] + w2o[8] * sw[2]};
    l = (float)std::sqrt((double)(so[0] * so[0] + so[1] * so[1] + so[2] * so[2]));
    if (!(l >= 0.99999982f && l <= 1.00000012f)) { so[0] /= l; so[1] /= l; so[2] /= l; }
    std::memcpy(p.sun_dir_object, so, 12);
    float theta = (float)std::acos((double)std::min(std::max(so[2], -1.0f), 1.0f));
    p.theta = theta;
    const double T = turbidity;
    // sun colour: attenuated solar spectrum (Preetham et al.), 350-800 nm
    if (so[2] > 0.0f) {
        float spectrum[91];
        const double alpha = 1.3, lozone = 0.35, w = 2.0;
        const double beta = 0.04608365822050 * T - 0.04586025928522;
        const double m = 1.0 / (std::cos((double)theta) + 0.000940 * std::pow(1.6386 - theta, -1.253));
        for (int i = 0, lam = 350; lam <= 800; i++, lam += 5) {
            float fl = (float)lam;
            double tr = std::exp(-m * 0.008735 * std::pow(lam / 1000.0, -4.08));
            double ta = std::exp(-m * beta * std::pow(lam / 1000.0, -alpha));
            double to = std::exp(-m * irregular_sample(kOzoneAmp, kOzoneWl, 64, fl) * lozone);
            float kg = irregular_sample(kGasAmp, kGasWl, 4, fl);
            double tg = std::exp(-1.41 * kg * m / std::pow(1.0 + 118.93 * kg * m, 0.45));
            float kw = irregular_sample(kWaterAmp, kWaterWl, 13, fl);
            double tw = std::exp(-0.2385 * kw * w * m / std::pow(1.0 + 20.07 * kw * w * m, 0.45));
            double amp = regular_sample(kSol, 38, 380.0f, 750.0f, fl) * tr * ta * to * tg * tw;
            spectrum[i] = (float)amp;
        }
        float xyz[3];
        to_xyz([&](float lam) { return regular_sample(spectrum, 91, 350.0f, 800.0f, lam); }, xyz);
        for (int k = 0; k < 3; k++) xyz[k] *= 0.0001f;
        for (int k = 0; k < 3; k++)
            p.sun_color[k] = std::max(0.0f, M[3 * k] * xyz[0] + M[3 * k + 1] * xyz[1] + M[3 * k + 2] * xyz[2]);
    }
    const double th = theta;
    p.zenith[0] = ((4.0453 * T - 4.9710) * std::tan((4.0 / 9.0 - T / 120.0) * (M_PI - 2.0 * th)) -
                   0.2155 * T + 2.4192) * 1000.0;
    p.zenith[1] = (0.00165 * th * th * th - 0.00374 * th * th + 0.00208 * th + 0.0) * T * T +
                  (-0.02902 * th * th * th + 0.06377 * th * th - 0.03202 * th + 0.00394) * T +
                  (0.11693 * th * th * th - 0.21196 * th * th + 0.06052 * th + 0.25885);
    p.zenith[2] = (0.00275 * th * th * th - 0.00610 * th * th + 0.00316 * th + 0.0) * T * T +
                  (-0.04212 * th * th * th + 0.08970 * th * th - 0.04153 * th + 0.00515) * T +
                  (0.15346 * th * th * th - 0.26756 * th * th + 0.06669 * th + 0.26688);
    const double lum[5] = {0.17872 * T - 1.46303, -0.35540 * T + 0.42749, -0.02266 * T + 5.32505,
                           0.12064 * T - 2.57705, -0.06696 * T + 0.37027};
    const double px[5] = {-0.01925 * T - 0.25922, -0.06651 * T + 0.00081, -0.00041 * T + 0.21247,
                          -0.06409 * T - 0.89887, -0.00325 * T + 0.04517};
    const double py[5] = {-0.01669 * T - 0.26078, -0.09495 * T + 0.00921, -0.00792 * T + 0.21023,
                          -0.04405 * T - 1.65369, -0.01092 * T + 0.05291};
    std::memcpy(p.perez_lum, lum, sizeof(lum));
    std::memcpy(p.perez_x, px, sizeof(px));
    std::memcpy(p.perez_y, py, sizeof(py));
    // single-precision copies for the device and the sun-only denominator
    // (1 + A e^B)(1 + C e^(D theta_s) + E cos^2 theta_s) of each set
    auto den = [&](const double* l, float* c) {
        for (int k = 0; k < 5; k++) c[k] = (float)l[k];
        const double cs = std::cos((double)p.theta);
        return (float)((1.0 + l[0] * std::exp(l[1])) * (1.0 + l[2] * std::exp(l[3] * (double)p.theta) + l[4] * cs * cs));
    };
    p.den_lum = den(lum, p.coef_lum);
    p.den_x = den(px, p.coef_x);
    p.den_y = den(py, p.coef_y);
    p.zd_lum = (float)(p.zenith[0] / (double)p.den_lum);
    p.zd_x = (float)(p.zenith[1] / (double)p.den_x);
    p.zd_y = (float)(p.zenith[2] / (double)p.den_y);
    to_xyz([&](float lam) { return regular_sample(kS0, 54, 300.0f, 830.0f, lam); }, p.s0);
    to_xyz([&](float lam) { return regular_sample(kS1, 54, 300.0f, 830.0f, lam); }, p.s1);
    to_xyz([&](float lam) { return regular_sample(kS2, 54, 300.0f, 830.0f, lam); }, p.s2);
    // importance distribution over (u = phi / 2pi, v = theta / pi)
    const int R = PEREZ_RES;
    float fn[PEREZ_RES];
    for (int u = 0; u < R; u++) {
        float su = (u + 0.5f) * (1.0f / R);
        for (int v = 0; v < R; v++) {
            float sv = (v + 0.5f) * (1.0f / R);
            float st = (float)std::sin((double)(PI_F * sv));
            float phi = su * 2.0f * PI_F, t = sv * PI_F;
            float sth = (float)std::sin((double)t);
            float d[3] = {sth * (float)std::cos((double)phi), sth * (float)std::sin((double)phi), (float)std::cos((double)t)};
            float rgb[3];
            sky(p, M, d, rgb);
            for (int k = 0; k < 3; k++) rgb[k] = std::max(rgb[k], 0.0f);
            float lumv = rgb[0] * 0.212671f + rgb[1] * 0.715160f + rgb[2] * 0.072169f;
            p.cond_func[u][v] = lumv * st;
        }
        p.cond_int[u] = dist1(p.cond_func[u], R, p.cond_cdf[u]);
        fn[u] = p.cond_int[u];
    }
    std::memcpy(p.marginal_func, fn, sizeof(fn));
    p.func_int = dist1(p.marginal_func, R, p.marginal_cdf);
}

// AngleF.degrees(d): wrapped into [0, 360] (Floats.wrapAround), then radians.
static float wrap_degrees(float d) {
    const float span = std::nextafter(360.0f, 1e9f);
    if (d >= 0.0f && d <= 360.0f) return d;
    if (d < 0.0f) return 360.0f - std::fmod(0.0f - d, span);
    return 0.0f + std::fmod(d - 360.0f, span);
}

// Floats.positiveModulo
static float pmodf(float x, float y) { return x < 0.0f ? std::fmod(std::fmod(x, y) + y, y) : std::fmod(x, y); }
static int pmodi(int x, int y) { return x < 0 ? (x % y + y) % y : x % y; }

// ImageLight.doRadianceSky for an object-space direction: its spherical
// coordinates, rotated counterclockwise, scaled, moved to image space and
// looked up bilinearly with wrap-around (ImageF.getColorRGB(x, y,
// WRAP_AROUND)).
void image_light_radiance(const DevImageLight& L, const float* texels, const float d[3], float rgb[3]) {
    float phi = (float)std::atan2((double)d[1], (double)d[0]);
    if (phi < 0.0f) phi += 2.0f * PI_F;
    const float u = phi * (1.0f / (2.0f * PI_F));
    const float v = (float)std::acos((double)std::min(std::max(d[2], -1.0f), 1.0f)) * (1.0f / PI_F);
    const float ru = (u * L.cos_angle - v * L.sin_angle) * L.scale_u;
    const float rv = (v * L.cos_angle + u * L.sin_angle) * L.scale_v;
    const float x = pmodf(ru * (float)L.width - 0.5f, (float)L.width);
    const float y = pmodf(rv * (float)L.height - 0.5f, (float)L.height);
    const int x0 = (int)std::floor(x), x1 = (int)std::ceil(x), y0 = (int)std::floor(y), y1 = (int)std::ceil(y);
    auto texel = [&](int tx, int ty, float* c) {
        const float* t = texels + L.texel + 3 * ((int64_t)pmodi(ty, L.height) * L.width + pmodi(tx, L.width));
        c[0] = t[0]; c[1] = t[1]; c[2] = t[2];
    };
    if (x0 == x1 && y0 == y1) { texel(x0, y0, rgb); return; }
    float c00[3], c01[3], c10[3], c11[3];
    texel(x0, y0, c00); texel(x1, y0, c01); texel(x0, y1, c10); texel(x1, y1, c11);
    const float fx = x - (float)x0, fy = y - (float)y0;
    for (int k = 0; k < 3; k++) {
        const float a = (1.0f - fx) * c00[k] + fx * c01[k];
        const float b = (1.0f - fx) * c10[k] + fx * c11[k];
        rgb[k] = (1.0f - fy) * a + fy * b;
    }
}

void image_light_build(DevImageLight& L, int width, int height, int64_t texel_offset, const float* texels,
                       float angle_degrees, float scale_u, float scale_v, float radius) {
    std::memset(&L, 0, sizeof(L));
    L.width = width;
    L.height = height;
    L.texel = texel_offset;
    const float rad = (float)(wrap_degrees(angle_degrees) * (3.14159265358979323846 / 180.0));
    L.cos_angle = (float)std::cos((double)rad);
    L.sin_angle = (float)std::sin((double)rad);
    L.scale_u = scale_u;
    L.scale_v = scale_v;
    L.radius = radius;
    // the same object frame as PerezLight's: Matrix44F.rotate(OrthonormalBasis33F(y, x))
    const float o2w[9] = {0, 1, 0, 0, 0, 1, 1, 0, 0};
    const float w2o[9] = {0, 0, 1, 1, 0, 0, 0, 1, 0};
    std::memcpy(L.object_to_world, o2w, sizeof(o2w));
    std::memcpy(L.world_to_object, w2o, sizeof(w2o));
    // doCreateDistribution: luminance x sin(theta) at the cell centres
    const int R = PEREZ_RES;
    float fn[PEREZ_RES];
    for (int u = 0; u < R; u++) {
        const float su = (u + 0.5f) * (1.0f / R);
        for (int v = 0; v < R; v++) {
            const float sv = (v + 0.5f) * (1.0f / R);
            const float st = (float)std::sin((double)(PI_F * sv));
            const float phi = su * 2.0f * PI_F, t = sv * PI_F;
            const float sth = (float)std::sin((double)t);
            const float d[3] = {sth * (float)std::cos((double)phi), sth * (float)std::sin((double)phi), (float)std::cos((double)t)};
            float rgb[3];
            image_light_radiance(L, texels, d, rgb);
            L.cond_func[u][v] = (rgb[0] * 0.212671f + rgb[1] * 0.715160f + rgb[2] * 0.072169f) * st;
        }
        L.cond_int[u] = dist1(L.cond_func[u], R, L.cond_cdf[u]);
        fn[u] = L.cond_int[u];
    }
    std::memcpy(L.marginal_func, fn, sizeof(fn));
    L.func_int = dist1(L.marginal_func, R, L.marginal_cdf);
}

}  // namespace dfr
