// oracle_math.h — TEST INFRASTRUCTURE ONLY (the oracle).  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the oracle;
// nothing in the product links it.
//
// Float32 value types restated from the reference with Java float semantics
// (IEEE single, no FMA contraction, transcendentals evaluated in double and
// rounded, as Floats.sin/cos/pow/exp/log do).  Citations: M/ =
// src/main/java/org/dayflower/ of the reference tree.
#pragma once
#include <cmath>
#include <cstdint>
#include <cfloat>
#include <algorithm>

namespace orc {

// M/utility/Floats.java:1027 (sin), :849 (pow), :1109 (sqrt): computed in
// double, rounded to float.
inline float fsin(float x) { return (float)std::sin((double)x); }
inline float fcos(float x) { return (float)std::cos((double)x); }
inline float ftan(float x) { return (float)std::tan((double)x); }
inline float fatan(float x) { return (float)std::atan((double)x); }
inline float facos(float x) { return (float)std::acos((double)x); }
inline float fasin(float x) { return (float)std::asin((double)x); }
inline float fatan2(float y, float x) { return (float)std::atan2((double)y, (double)x); }
inline float fsqrt(float x) { return (float)std::sqrt((double)x); }
inline float fpow(float b, float e) { return (float)std::pow((double)b, (double)e); }
inline float fexp(float x) { return (float)std::exp((double)x); }
inline float flog(float x) { return (float)std::log((double)x); }
inline bool isZero(float x) { return x == 0.0f; }

const float PI = 3.14159265358979323846f;
const float PI_RECIPROCAL = 1.0f / PI;
// M/utility/Floats.java:44, :565-567
const float MACHINE_EPSILON = FLT_EPSILON * 0.5f;
inline float gamma(int n) { return (n * MACHINE_EPSILON) / (1.0f - n * MACHINE_EPSILON); }
inline float nextUp(float v) { return std::nextafter(v, INFINITY); }
inline float nextDown(float v) { return std::nextafter(v, -INFINITY); }

struct V3 {
    float x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return V3(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator/(V3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float dotAbs(V3 a, V3 b) { return std::fabs(dot(a, b)); }
inline V3 cross(V3 a, V3 b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float lengthSquared(V3 a) { return dot(a, a); }
inline float length(V3 a) { return fsqrt(lengthSquared(a)); }
// M/geometry/Vector3F.java:1000-1011: returned unchanged when the length is
// within [nextDown(1)^3-ish, nextUp(1)] of one, else divided by the length.
const float NEXT_DOWN_1_3 = 0.99999982f;  // nextDown(nextDown(nextDown(1)))
const float NEXT_UP_1_1 = 1.00000012f;    // nextUp(1)
inline V3 normalize(V3 v) {
    float l = length(v);
    if (l >= NEXT_DOWN_1_3 && l <= NEXT_UP_1_1) return v;
    return v / l;
}
inline V3 negateZ(V3 v) { return V3(v.x, v.y, -v.z); }
inline float cosTheta(V3 v) { return v.z; }
inline float cosThetaAbs(V3 v) { return std::fabs(v.z); }
inline float cosThetaQuartic(V3 v) { return v.z * v.z * v.z * v.z; }
inline float sinThetaSquared(V3 v) { return std::max(0.0f, 1.0f - v.z * v.z); }
inline float sinTheta(V3 v) { return fsqrt(sinThetaSquared(v)); }
inline float tanThetaSquared(V3 v) { return sinThetaSquared(v) / (v.z * v.z); }
inline float tanThetaAbs(V3 v) { return std::fabs(sinTheta(v) / v.z); }
inline float tanThetaSquaredRef(V3 v) { return sinThetaSquared(v) / (v.z * v.z); }
inline float cosPhi(V3 v) {
    float s = sinTheta(v);
    return s == 0.0f ? 1.0f : std::min(std::max(v.x / s, -1.0f), 1.0f);
}
inline float sinPhi(V3 v) {
    float s = sinTheta(v);
    return s == 0.0f ? 0.0f : std::min(std::max(v.y / s, -1.0f), 1.0f);
}
inline float cosPhiSquared(V3 v) { float c = cosPhi(v); return c * c; }
inline float sinPhiSquared(V3 v) { float s = sinPhi(v); return s * s; }
inline bool sameHemisphereZ(V3 a, V3 b) { return a.z * b.z > 0.0f; }
inline V3 faceForward(V3 a, V3 b) { return dot(a, b) < 0.0f ? -a : a; }
inline V3 faceForwardZ(V3 a, V3 b) { return a.z < 0.0f ? negateZ(b) : b; }
// Vector3F.reflection(direction, normal) = direction - normal*2*dot, negated
// form used by pbrt Reflect(wo, n) = -wo + 2*dot(wo,n)*n.
inline V3 reflectionPBRT(V3 wo, V3 n) { return -wo + n * (2.0f * dot(wo, n)); }
// M/geometry/Vector3F.java:454-466
inline bool refraction(V3 d, V3 n, float eta, V3* out) {
    float cosI = dot(d, n);
    float sin2I = std::max(0.0f, 1.0f - cosI * cosI);
    float sin2T = eta * eta * sin2I;
    if (sin2T >= 1.0f) return false;
    float cosT = fsqrt(1.0f - sin2T);
    *out = (-d) * eta + n * (eta * cosI - cosT);
    return true;
}
// M/geometry/Vector3F.java:527-545
inline V3 computeV(V3 w) {
    V3 n = normalize(w);
    float ax = std::fabs(n.x), ay = std::fabs(n.y), az = std::fabs(n.z);
    if (ax < ay && ax < az) return normalize(V3(0.0f, n.z, -n.y));
    if (ay < az) return normalize(V3(n.z, 0.0f, -n.x));
    return normalize(V3(n.y, -n.x, 0.0f));
}

// M/geometry/OrthonormalBasis33F.java
struct ONB { V3 u, v, w; };
inline ONB onbFromW(V3 w) {  // :88-97
    ONB o; o.w = normalize(w); o.v = computeV(w); o.u = cross(o.v, o.w); return o;
}
inline ONB onbFromWV(V3 w, V3 v) {  // :110-114
    ONB o; o.w = normalize(w); o.u = normalize(cross(normalize(v), o.w)); o.v = cross(o.w, o.u); return o;
}
inline V3 toLocal(V3 v, const ONB& b) {  // Vector3F.transformReverse + normalize
    return normalize(V3(dot(v, b.u), dot(v, b.v), dot(v, b.w)));
}
inline V3 toWorld(V3 v, const ONB& b) {  // Vector3F.transform + normalize
    return normalize(V3(v.x * b.u.x + v.y * b.v.x + v.z * b.w.x,
                        v.x * b.u.y + v.y * b.v.y + v.z * b.w.y,
                        v.x * b.u.z + v.y * b.v.z + v.z * b.w.z));
}

struct C3 {
    float r, g, b;
    C3() : r(0), g(0), b(0) {}
    C3(float a) : r(a), g(a), b(a) {}
    C3(float a, float b_, float c) : r(a), g(b_), b(c) {}
};
inline C3 operator+(C3 a, C3 b) { return C3(a.r + b.r, a.g + b.g, a.b + b.b); }
inline C3 operator-(C3 a, C3 b) { return C3(a.r - b.r, a.g - b.g, a.b - b.b); }
inline C3 operator*(C3 a, C3 b) { return C3(a.r * b.r, a.g * b.g, a.b * b.b); }
inline C3 operator*(C3 a, float s) { return C3(a.r * s, a.g * s, a.b * s); }
inline C3 operator/(C3 a, float s) { return C3(a.r / s, a.g / s, a.b / s); }
inline C3 operator/(C3 a, C3 b) { return C3(a.r / b.r, a.g / b.g, a.b / b.b); }
inline bool isBlack(C3 c) { return c.r == 0.0f && c.g == 0.0f && c.b == 0.0f; }
inline float maxc(C3 c) { return std::max(c.r, std::max(c.g, c.b)); }
inline C3 saturate0(C3 c) {  // Color3F.saturate(c, 0, MAX_VALUE)
    return C3(std::max(c.r, 0.0f), std::max(c.g, 0.0f), std::max(c.b, 0.0f));
}
inline bool hasNaN(C3 c) { return std::isnan(c.r) || std::isnan(c.g) || std::isnan(c.b); }
inline bool hasInf(C3 c) { return std::isinf(c.r) || std::isinf(c.g) || std::isinf(c.b); }
// sRGB relative luminance (ColorSpaceF default: sRGB, Color3F.relativeLuminance)
inline float relativeLuminance(C3 c) { return 0.212671f * c.r + 0.715160f * c.g + 0.072169f * c.b; }

}  // namespace orc
