// device.h — device functions of the path-tracing hot path (gfx950).
//
// Semantics follow the reference CPURenderer path (M/ = src/main/java/
// org/dayflower/ of the reference tree); the data layout is this build's:
// a two-wide BVH over all triangles with child-pair boxes, triangles as three
// float4 in BVH order, shading data indexed by the original triangle.
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

#include "../../include/dayflower_hip.h"
#include "device_types.h"

namespace dfr {

#define DEV __device__ __forceinline__

// Address-space qualified pointers.  Scene buffers are reached through
// pointers stored in DevScene records, which the compiler cannot place in an
// address space by itself: without these it emits flat accesses (which wait
// on both the vector-memory and the LDS counters).
#define AS_GLOBAL __attribute__((address_space(1)))
#define AS_LDS __attribute__((address_space(3)))
typedef float native_float4 __attribute__((ext_vector_type(4)));
// p[i] of a float4 array in HBM, loaded as a global (not flat) access
DEV float4 ld4g(const float4* p, size_t i) {
    const native_float4 v = ((const AS_GLOBAL native_float4*)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}

// ---------------------------------------------------------------------------
// small vector type
// ---------------------------------------------------------------------------
struct V {
    float x, y, z;
};
DEV V mk(float x, float y, float z) { V v; v.x = x; v.y = y; v.z = z; return v; }
DEV V operator+(V a, V b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
DEV V operator-(V a, V b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
DEV V operator-(V a) { return mk(-a.x, -a.y, -a.z); }
DEV V operator*(V a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
DEV V operator*(V a, V b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
// one correctly rounded reciprocal, three products (within an ulp of three
// quotients)
DEV V operator/(V a, float s) { const float r = __builtin_amdgcn_rcpf(s); return mk(a.x * r, a.y * r, a.z * r); }
DEV float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DEV V cross(V a, V b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
DEV float comp(V v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
DEV float len(V v) { return sqrtf(dot(v, v)); }
// Vector3F.normalize (M/geometry/Vector3F.java:1000-1011): unchanged when the
// length is already within [nextDown^3(1), nextUp(1)] (tested on the squared
// length); otherwise scaled by the hardware reciprocal square root (about an
// ulp), which replaces a square root and three divisions.
DEV V nrm(V v) {
    const float d = dot(v, v);
    if (d >= 0.99999964f && d <= 1.00000024f) return v;
    return v * __builtin_amdgcn_rsqf(d);
}
DEV V ld3(const float* p) { return mk(p[0], p[1], p[2]); }
DEV V f4v(float4 a) { return mk(a.x, a.y, a.z); }
DEV bool black(V c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }
DEV float maxc(V c) { return fmaxf(c.x, fmaxf(c.y, c.z)); }

constexpr float PI_F = 3.14159265358979323846f;
constexpr float INV_PI = 1.0f / 3.14159265358979323846f;
constexpr float T_MIN = 0.001f;  // Scene.T_MINIMUM, M/scene/Scene.java:70

// ---------------------------------------------------------------------------
// RNG: stateless hash keyed by (seed, pixel, sample, dimension).  The
// integer hash is Wellons' "lowbias32".  Dimension layout: 0,1 pixel offset;
// path-tracing bounce b uses 2 + 12*b + {0 light choice, 1-2 light sample,
// 3-4 NEE BSDF sample, 5 Perez sun choice, 6-7 BSDF continuation, 8 Russian
// roulette, 9 clear coat layer choice}; ambient-occlusion sample i uses
// 2 + 2*i and 3 + 2*i.
// ---------------------------------------------------------------------------
DEV uint32_t lowbias32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
struct Rng {
    uint32_t a, b;
    DEV float get(uint32_t dim) const {
        uint32_t h = lowbias32(a ^ lowbias32(b + dim * 0x9e3779b9u));
        return (float)(h >> 8) * (1.0f / 16777216.0f);
    }
};
DEV Rng make_rng(uint64_t seed, uint32_t pixel, uint32_t sample) {
    Rng r;
    r.a = lowbias32(pixel ^ lowbias32((uint32_t)seed));
    r.b = lowbias32(sample ^ lowbias32((uint32_t)(seed >> 32) ^ 0x85ebca6bu) ^ r.a);
    return r;
}
constexpr uint32_t DIM_BOUNCE = 2, DIMS_PER_BOUNCE = 12;

// ---------------------------------------------------------------------------
// Compiled scene
// ---------------------------------------------------------------------------
struct DevScene {
    const float4* nodes;      // 4 float4 per two-wide node
    const float4* tris;       // 3 float4 per triangle (BVH order); tris[3i].w = original id (int bits)
    // per triangle (original order), the shading record of one 128-byte line:
    // (a, nA.x) (b, nA.y) (c, nA.z) (nB, ng.x) (nC, ng.y) (ng.z, -, -, -), two
    // float4 of padding; a, b, c positions, nA..nC vertex normals, ng the
    // geometric normal
    const float4* tshade;
    const DevPrim* prims;
    const int32_t* prim_class;  // per primitive: 4 x the material's sort class
    const int32_t* analytic;  // infinite planes (and spheres and rectangles too few for a tree)
    const DevMaterial* mats;
    const DevLobe* lobes;
    const DevLight* lights;
    const DevPerez* perez;
    // per primitive, six float4 rows: rows 0-2 map a world-space normal to
    // its object-space direction, rows 3-5 an object-space frame vector to
    // world space (read for primitives with `xformed` set)
    const float4* xforms;
    const DevImageLight* images;  // image lights (DevLight::perez indexes them)
    const float* texels;          // RGB texels of image lights and image textures
    const DevTexture* textures;
    // spheres and rectangles: a four-wide BVH of their own over their world
    // boxes (item_nodes), whose leaves index item_prims (primitives in BVH
    // order); walked ahead of the triangles' BVH.  n_items 0: a scene with
    // few of them tests them in the analytic list instead
    const float4* item_nodes;
    const int32_t* item_prims;
    // instances of shared meshes: the top-level tree over their world boxes
    // (root 0; its leaves index `instances`) followed by every shared mesh's
    // own tree (leaves index inst_tris, object-space triangles in the
    // layout of `tris`, a.w the original triangle)
    const float4* inst_nodes;
    const float4* inst_tris;
    const DevInstance* instances;
    int32_t n_instances;
    int32_t n_items;
    // items or instances present: the wavefront's traversal starts from the
    // result k_items_queue left in the hit record
    int32_t prepass;
    int32_t n_analytic;       // entries of analytic, tested linearly
    int32_t n_lights;
    int32_t has_bvh;
    int32_t light_pdf_mode;   // 0 reference (Shape3F pdf without 1/area), 1 area
};

constexpr uint32_t LEAF_FLAG_DEV = 0x80000000u;  // leaf reference bit (see bvh.h)

// Per-thread traversal stack: the first SHORT entries live in LDS (entry k of
// a thread at lds[k * bdim]); deeper entries overflow into an HBM buffer
// (entry k at glob[(k - SHORT) * gstride]).
template <int SHORT>
struct TraversalStack {
    AS_LDS int* lds;
    AS_GLOBAL int* glob;
    int gstride;
    int bdim;
    int sp;
    // `lds_block`: the block's SHORT * bdim ints of LDS; `overflow`: the HBM
    // overflow buffer, one column per thread of the grid
    DEV void init(int* lds_block, int* overflow, int block_dim) {
        lds = (AS_LDS int*)(lds_block + threadIdx.x);
        bdim = block_dim;
        glob = (AS_GLOBAL int*)(overflow + (size_t)blockIdx.x * block_dim + threadIdx.x);
        gstride = gridDim.x * block_dim;
        sp = 0;
    }
    DEV void push(int v) {
        if (sp < SHORT) lds[sp * bdim] = v;
        else glob[(sp - SHORT) * gstride] = v;
        sp++;
    }
    DEV int pop() {
        sp--;
        if (sp < SHORT) return lds[sp * bdim];
        return glob[(sp - SHORT) * gstride];
    }
};

// ---------------------------------------------------------------------------
// Shapes
// ---------------------------------------------------------------------------
DEV float gamma_n(int n) {
    const float e = 5.96046448e-08f;  // ulp(1)/2, M/utility/Floats.java:44
    return (n * e) / (1.0f - n * e);
}

// Triangle3F.intersectionT (M/geometry/shape/Triangle3F.java:505-600):
// pbrt-v3 watertight test with the FP64 recomputation of zero edge
// functions.  The near limit is t > tmin (the reference compares t * det with
// tmin, which hides small triangles).  Returns 0 on a miss.
DEV float tri_t(V p0, V p1, V p2, V o, V d, int kx, int ky, int kz, float sx, float sy, float sz, float tmin,
                float tmax) {
    V a = p0 - o, b = p1 - o, c = p2 - o;
    float ax = comp(a, kx), ay = comp(a, ky), az = comp(a, kz);
    float bx = comp(b, kx), by = comp(b, ky), bz = comp(b, kz);
    float cx = comp(c, kx), cy = comp(c, ky), cz = comp(c, kz);
    ax = ax + sx * az; ay = ay + sy * az;
    bx = bx + sx * bz; by = by + sy * bz;
    cx = cx + sx * cz; cy = cy + sy * cz;
    float e0 = bx * cy - by * cx;
    float e1 = cx * ay - cy * ax;
    float e2 = ax * by - ay * bx;
    if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f) {
        e0 = (float)((double)cy * (double)bx - (double)cx * (double)by);
        e1 = (float)((double)ay * (double)cx - (double)ax * (double)cy);
        e2 = (float)((double)by * (double)ax - (double)bx * (double)ay);
    }
    if ((e0 < 0.0f || e1 < 0.0f || e2 < 0.0f) && (e0 > 0.0f || e1 > 0.0f || e2 > 0.0f)) return 0.0f;
    float det = e0 + e1 + e2;
    if (det == 0.0f) return 0.0f;
    az *= sz; bz *= sz; cz *= sz;
    float ts = e0 * az + e1 * bz + e2 * cz;
    // t = ts / det must lie in (tmin, tmax]: compare in scaled form
    if (det < 0.0f && (ts >= tmin * det || ts < tmax * det)) return 0.0f;
    if (det > 0.0f && (ts <= tmin * det || ts > tmax * det)) return 0.0f;
    float inv = 1.0f / det;
    float t = ts * inv;
    float maxZ = fmaxf(fabsf(az), fmaxf(fabsf(bz), fabsf(cz)));
    float maxX = fmaxf(fabsf(ax), fmaxf(fabsf(bx), fabsf(cx)));
    float maxY = fmaxf(fabsf(ay), fmaxf(fabsf(by), fabsf(cy)));
    float dz = gamma_n(3) * maxZ;
    float dx = gamma_n(5) * (maxX + maxZ);
    float dy = gamma_n(5) * (maxY + maxZ);
    float de = 2.0f * (gamma_n(2) * maxX * maxY + dy * maxX + dx * maxY);
    float maxE = fmaxf(fabsf(e0), fmaxf(fabsf(e1), fabsf(e2)));
    float dt = 3.0f * (gamma_n(3) * maxE * maxZ + de * maxZ + dz * maxE) * fabsf(inv);
    if (t <= dt) return 0.0f;
    return t;
}

// Ray prepared for the watertight test and the slab test.
struct PRay {
    V o, d;
    int kx, ky, kz;
    float sx, sy, sz;
    V inv, oinv;
};
DEV PRay prep_ray(V o, V d) {
    PRay r;
    r.o = o; r.d = d;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    r.kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);  // Vector3F.getMaxDimension
    r.kx = r.kz + 1; if (r.kx == 3) r.kx = 0;
    r.ky = r.kx + 1; if (r.ky == 3) r.ky = 0;
    float dz = comp(d, r.kz);
    r.sx = -comp(d, r.kx) / dz;
    r.sy = -comp(d, r.ky) / dz;
    r.sz = 1.0f / dz;
    auto safe = [](float v) { return fabsf(v) < 1e-20f ? copysignf(1e-20f, v) : v; };
    r.inv = mk(1.0f / safe(d.x), 1.0f / safe(d.y), 1.0f / safe(d.z));
    r.oinv = o * r.inv;
    return r;
}

// EFloat interval arithmetic (M/utility/EFloat.java:34-195).
struct EF { float v, lo, hi; };
DEV float nup(float v) { return nextafterf(v, INFINITY); }
DEV float ndn(float v) { return nextafterf(v, -INFINITY); }
DEV EF ef(float v) { EF e; e.v = v; e.lo = v; e.hi = v; return e; }
DEV EF ef_add(EF a, EF b) { EF e; e.v = a.v + b.v; e.lo = ndn(a.lo + b.lo); e.hi = nup(a.hi + b.hi); return e; }
DEV EF ef_sub(EF a, EF b) { EF e; e.v = a.v - b.v; e.lo = ndn(a.lo - b.hi); e.hi = nup(a.hi - b.lo); return e; }
DEV EF ef_mul(EF a, EF b) {
    float p0 = a.lo * b.lo, p1 = a.hi * b.lo, p2 = a.lo * b.hi, p3 = a.hi * b.hi;
    EF e; e.v = a.v * b.v;
    e.lo = ndn(fminf(fminf(p0, p1), fminf(p2, p3)));
    e.hi = nup(fmaxf(fmaxf(p0, p1), fmaxf(p2, p3)));
    return e;
}
DEV EF ef_div(EF a, EF b) {
    EF e; e.v = a.v / b.v;
    if (b.lo < 0.0f && b.hi > 0.0f) { e.lo = -INFINITY; e.hi = INFINITY; return e; }
    float d0 = a.lo / b.lo, d1 = a.hi / b.lo, d2 = a.lo / b.hi, d3 = a.hi / b.hi;
    e.lo = ndn(fminf(fminf(d0, d1), fminf(d2, d3)));
    e.hi = nup(fmaxf(fmaxf(d0, d1), fmaxf(d2, d3)));
    return e;
}

// Sphere3F.intersectionT for a full sphere (M/geometry/shape/Sphere3F.java).
DEV float sphere_t(const DevPrim& p, V o, V d, float tmin, float tmax) {
    V c = mk(p.data[0], p.data[1], p.data[2]);
    float radius = p.data[3];
    V q = o - c;
    EF ox = ef(q.x), oy = ef(q.y), oz = ef(q.z), dx = ef(d.x), dy = ef(d.y), dz = ef(d.z);
    EF A = ef_add(ef_add(ef_mul(dx, dx), ef_mul(dy, dy)), ef_mul(dz, dz));
    EF B = ef_mul(ef(2.0f), ef_add(ef_add(ef_mul(dx, ox), ef_mul(dy, oy)), ef_mul(dz, oz)));
    EF C = ef_sub(ef_add(ef_add(ef_mul(ox, ox), ef_mul(oy, oy)), ef_mul(oz, oz)), ef_mul(ef(radius), ef(radius)));
    double disc = (double)B.v * (double)B.v - 4.0 * (double)A.v * (double)C.v;
    if (disc < 0.0) return 0.0f;
    double root = sqrt(disc);
    float fr = (float)root;
    EF R; R.v = fr;
    float err = 5.96046448e-08f * fr;
    if (err == 0.0f) { R.lo = fr; R.hi = fr; } else { R.lo = ndn(fr - err); R.hi = nup(fr + err); }
    EF Q = B.v < 0.0f ? ef_mul(ef(-0.5f), ef_sub(B, R)) : ef_mul(ef(-0.5f), ef_add(B, R));
    EF t0 = ef_div(Q, A), t1 = ef_div(C, Q);
    if (t0.v > t1.v) { EF tmp = t0; t0 = t1; t1 = tmp; }
    if (t0.hi > tmax || t1.lo <= tmin) return 0.0f;
    EF th = t0;
    if (th.lo <= 0.0f) {
        th = t1;
        if (th.hi > tmax) return 0.0f;
    }
    return th.v;
}

// Rectangle3F.intersectionT
DEV float rect_t(const DevPrim& p, V o, V d, float tmin, float tmax) {
    V a = ld3(p.data), b = ld3(p.data + 3), c = ld3(p.data + 6);
    V n = ld3(p.normal);
    float dp = dot(n, d);
    if (dp == 0.0f) return 0.0f;
    float t = dot(a - o, n) / dp;
    if (t <= tmin || t >= tmax) return 0.0f;
    V ab = b - a, bc = c - b;
    V ap = (o + d * t) - a;
    float dab = dot(ap, nrm(ab)), dbc = dot(ap, nrm(bc));
    if (dab >= 0.0f && dab <= len(ab) && dbc >= 0.0f && dbc <= len(bc)) return t;
    return 0.0f;
}

// Plane3F.intersectionT
DEV float plane_t(const DevPrim& p, V o, V d, float tmin, float tmax) {
    V org = ld3(p.data), w = ld3(p.data + 9);
    float dp = dot(w, d);
    if (dp == 0.0f) return 0.0f;
    float t = dot(org - o, w) / dp;
    if (t > tmin && t < tmax) return t;
    return 0.0f;
}

// FP64 distance to the winning analytic primitive along the same ray: the
// plane distance for rectangles and planes, and for spheres the root of the
// double-precision quadratic nearest to the float one that won.
DEV float analytic_t64(const DevPrim& p, V o, V d, float t32) {
    double dx = d.x, dy = d.y, dz = d.z;
    if (p.shape == DFR_SHAPE_SPHERE) {
        double ox = (double)o.x - p.data[0], oy = (double)o.y - p.data[1], oz = (double)o.z - p.data[2];
        double r = p.data[3];
        double a = dx * dx + dy * dy + dz * dz;
        double b = 2.0 * (dx * ox + dy * oy + dz * oz);
        double c = ox * ox + oy * oy + oz * oz - r * r;
        double disc = b * b - 4.0 * a * c;
        if (disc < 0.0) return t32;
        double q = b < 0.0 ? -0.5 * (b - sqrt(disc)) : -0.5 * (b + sqrt(disc));
        if (q == 0.0) return t32;
        double t0 = q / a, t1 = c / q;
        double t = fabs(t0 - t32) < fabs(t1 - t32) ? t0 : t1;
        return t > 0.0 ? (float)t : t32;
    }
    const float* n = p.shape == DFR_SHAPE_RECTANGLE ? p.normal : p.data + 9;
    double den = n[0] * dx + n[1] * dy + n[2] * dz;
    if (den == 0.0) return t32;
    double t = (((double)p.data[0] - o.x) * n[0] + ((double)p.data[1] - o.y) * n[1] +
                ((double)p.data[2] - o.z) * n[2]) / den;
    return t > 0.0 ? (float)t : t32;
}

DEV float analytic_t(const DevPrim& p, V o, V d, float tmin, float tmax) {
    switch (p.shape) {
        case DFR_SHAPE_SPHERE: return sphere_t(p, o, d, tmin, tmax);
        case DFR_SHAPE_RECTANGLE: return rect_t(p, o, d, tmin, tmax);
        case DFR_SHAPE_PLANE: return plane_t(p, o, d, tmin, tmax);
        default: return 0.0f;
    }
}

// ---------------------------------------------------------------------------
// BVH traversal: ordered two-wide walk with a short LDS stack that overflows
// into HBM.  Closest hit shrinks tmax; any hit returns at the first hit.
// ---------------------------------------------------------------------------
struct HitRec {
    float t;
    int prim;
    int tri;   // original triangle index or -1
};

DEV bool slab2(const float4& n0, const float4& n1, const float4& n2, const PRay& r, float tmin, float tmax,
               float& t0, float& t1, bool& h0, bool& h1) {
    // child 0
    float lx0 = n0.x * r.inv.x - r.oinv.x, hx0 = n0.y * r.inv.x - r.oinv.x;
    float ly0 = n0.z * r.inv.y - r.oinv.y, hy0 = n0.w * r.inv.y - r.oinv.y;
    float lz0 = n2.x * r.inv.z - r.oinv.z, hz0 = n2.y * r.inv.z - r.oinv.z;
    float near0 = fmaxf(fmaxf(fminf(lx0, hx0), fminf(ly0, hy0)), fmaxf(fminf(lz0, hz0), tmin));
    float far0 = fminf(fminf(fmaxf(lx0, hx0), fmaxf(ly0, hy0)), fminf(fmaxf(lz0, hz0), tmax));
    // child 1
    float lx1 = n1.x * r.inv.x - r.oinv.x, hx1 = n1.y * r.inv.x - r.oinv.x;
    float ly1 = n1.z * r.inv.y - r.oinv.y, hy1 = n1.w * r.inv.y - r.oinv.y;
    float lz1 = n2.z * r.inv.z - r.oinv.z, hz1 = n2.w * r.inv.z - r.oinv.z;
    float near1 = fmaxf(fmaxf(fminf(lx1, hx1), fminf(ly1, hy1)), fmaxf(fminf(lz1, hz1), tmin));
    float far1 = fminf(fminf(fmaxf(lx1, hx1), fmaxf(ly1, hy1)), fminf(fmaxf(lz1, hz1), tmax));
    // widen the far end by 1 + 2*gamma(3) so rounding never culls a grazing hit
    h0 = near0 <= far0 * 1.00000048f;
    h1 = near1 <= far1 * 1.00000048f;
    t0 = near0; t1 = near1;
    return h0 || h1;
}

struct VisitCount {
    int nodes = 0, tris = 0;
};

constexpr uint32_t TRAVERSAL_DONE = 0x7fffffffu;  // no node left (bit 31 clear, never a node index)
// Phase switches of the while-while traversal on 64-wide waves: the wave
// leaves the inner-node loop once at most this many of its lanes still have
// no leaf parked, and leaves the leaf loop once at most this many lanes still
// hold one (the others' work carries over to the next round).  Waiting for
// every lane, as on 32-wide warps, idles most of a 64-wide wave.
constexpr int TRAVERSAL_STRAGGLERS = 8;

// One inner-node step of an ordered traversal: tests the children's boxes
// against (tmin, tmax), continues with the nearest entered child and pushes
// the other entered children (farthest first, so the nearer pop first); with
// none entered, pops.  Returns the next reference or TRAVERSAL_DONE.
template <int SHORT>
DEV uint32_t visit2(const float4* nodes, uint32_t cur, const PRay& r, float tmin, float tmax, TraversalStack<SHORT>& st) {
    const size_t nb = 4 * (size_t)cur;
    const float4 n0 = ld4g(nodes, nb), n1 = ld4g(nodes, nb + 1), n2 = ld4g(nodes, nb + 2), n3 = ld4g(nodes, nb + 3);
    float t0, t1;
    bool h0, h1;
    slab2(n0, n1, n2, r, tmin, tmax, t0, t1, h0, h1);
    const uint32_t c0 = __float_as_uint(n3.x), c1 = __float_as_uint(n3.y);
    if (c1 == 0xffffffffu) h1 = false;
    if (h0 && h1) {
        const bool swap = t1 < t0;
        st.push((int)(swap ? c0 : c1));
        return swap ? c1 : c0;
    }
    if (h0 || h1) return h0 ? c0 : c1;
    return st.sp == 0 ? TRAVERSAL_DONE : (uint32_t)st.pop();
}

// Entry distances tn and widened exit distances tf of the four children of a
// four-wide node -> the next reference (ORDERED: the nearest entered child,
// the other entered children pushed farthest first, through a five-exchange
// sorting network; otherwise, for any-hit queries, the entered children in
// slot order); with none entered, the top of the stack, or TRAVERSAL_DONE
// when the stack is down to `base` entries (a walk nested on top of another
// owns only the entries above its base).
template <int SHORT, bool ORDERED>
DEV uint32_t visit4_next(const float (&tn)[4], const float (&tf)[4], const float4 rf, TraversalStack<SHORT>& st,
                         int base = 0) {
    const uint32_t R[4] = {__float_as_uint(rf.x), __float_as_uint(rf.y), __float_as_uint(rf.z), __float_as_uint(rf.w)};
    float key[4];
    uint32_t ref[4];
    int hits = 0;
    bool entered[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        // far end widened by 1 + 2*gamma(3), as in the two-wide test
        const bool h = tn[c] <= tf[c] * 1.00000048f && R[c] != 0xffffffffu;
        key[c] = h ? tn[c] : INFINITY;
        ref[c] = R[c];
        entered[c] = h;
        hits += h ? 1 : 0;
    }
    // sort the four (key, ref) pairs by key: misses (infinite keys) go last
    auto ce = [&](int a, int b) {
        const bool sw = key[b] < key[a];
        const float ka = key[a], kb = key[b];
        const uint32_t ra = ref[a], rb = ref[b];
        key[a] = sw ? kb : ka; key[b] = sw ? ka : kb;
        ref[a] = sw ? rb : ra; ref[b] = sw ? ra : rb;
    };
    if (!ORDERED) {
        // any-hit: first entered child next, the others pushed as they come
        uint32_t next = TRAVERSAL_DONE;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            if (!entered[c]) continue;
            if (next == TRAVERSAL_DONE) next = ref[c];
            else st.push((int)ref[c]);
        }
        if (next != TRAVERSAL_DONE) return next;
        return st.sp == base ? TRAVERSAL_DONE : (uint32_t)st.pop();
    }
    ce(0, 1); ce(2, 3); ce(0, 2); ce(1, 3); ce(1, 2);
    if (hits == 0) return st.sp == base ? TRAVERSAL_DONE : (uint32_t)st.pop();
    if (hits >= 4) st.push((int)ref[3]);
    if (hits >= 3) st.push((int)ref[2]);
    if (hits >= 2) st.push((int)ref[1]);
    return ref[0];
}

// Near/far plane selection of a ray for visit4_signed: per axis, the byte
// offset (0 or 16) of the entry plane within the axis' (lo, hi) pair of a
// node.  A ray enters a slab through lo where its inverse direction is
// positive (prep_ray never leaves a zero component), so the entry distance is
// the lo plane's and the exit distance the hi plane's, the same values
// min/max would pick: fma is monotonic in its first operand for a fixed
// positive factor.
struct PlaneSel { uint32_t x, y, z; };
DEV PlaneSel plane_sel(const PRay& r) {
    PlaneSel p;
    p.x = r.inv.x < 0.0f ? 16u : 0u;
    p.y = r.inv.y < 0.0f ? 16u : 0u;
    p.z = r.inv.z < 0.0f ? 16u : 0u;
    return p;
}
// a float4 at a byte offset of a buffer whose base is the same for the whole
// wave: the load takes the base from scalar registers and a 32-bit offset
DEV float4 ld4off(const float4* base, uint32_t byte_off, uint32_t imm) {
    const native_float4 v = *(const AS_GLOBAL native_float4*)((const AS_GLOBAL char*)base + byte_off + imm);
    return make_float4(v.x, v.y, v.z, v.w);
}

// One step of a four-wide node (byte offsets below 4 GiB, which compile_scene
// checks): the entry and exit planes of each slab are loaded directly, so no
// min/max orders them.  With a node array shared by the wave, the loads take
// its base from scalar registers.
template <int SHORT, bool ORDERED = true>
DEV uint32_t visit4_signed(const float4* nodes, uint32_t cur, const PRay& r, const PlaneSel& ps, float tmin, float tmax,
                           TraversalStack<SHORT>& st, int base = 0) {
    const uint32_t nb = cur << 7;  // 128 bytes per node
    const uint32_t ox = nb | ps.x, oy = nb | ps.y, oz = nb | ps.z;
    const float4 nx = ld4off(nodes, ox, 0), fx = ld4off(nodes, ox ^ 16u, 0);
    const float4 ny = ld4off(nodes, oy, 32), fy = ld4off(nodes, oy ^ 16u, 32);
    const float4 nz = ld4off(nodes, oz, 64), fz = ld4off(nodes, oz ^ 16u, 64);
    const float4 rf = ld4off(nodes, nb, 96);
    const float N[3][4] = {{nx.x, nx.y, nx.z, nx.w}, {ny.x, ny.y, ny.z, ny.w}, {nz.x, nz.y, nz.z, nz.w}};
    const float F[3][4] = {{fx.x, fx.y, fx.z, fx.w}, {fy.x, fy.y, fy.z, fy.w}, {fz.x, fz.y, fz.z, fz.w}};
    float tn[4], tf[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const float ax = N[0][c] * r.inv.x - r.oinv.x, bx = F[0][c] * r.inv.x - r.oinv.x;
        const float ay = N[1][c] * r.inv.y - r.oinv.y, by = F[1][c] * r.inv.y - r.oinv.y;
        const float az = N[2][c] * r.inv.z - r.oinv.z, bz = F[2][c] * r.inv.z - r.oinv.z;
        tn[c] = fmaxf(fmaxf(ax, ay), fmaxf(az, tmin));
        tf[c] = fminf(fminf(bx, by), fminf(bz, tmax));
    }
    return visit4_next<SHORT, ORDERED>(tn, tf, rf, st, base);
}

template <int SHORT, bool ORDERED = true>
DEV uint32_t visit_node(const float4* nodes, uint32_t cur, const PRay& r, const PlaneSel& ps, float tmin, float tmax,
                        TraversalStack<SHORT>& st, int base = 0) {
    if (BVH_WIDTH == 4) return visit4_signed<SHORT, ORDERED>(nodes, cur, r, ps, tmin, tmax, st, base);
    return visit2<SHORT>(nodes, cur, r, tmin, tmax, st);
}

// Vertices of triangle i in BVH order (48 bytes each)
DEV void load_tri(const float4* tris, uint32_t i, float4& a, float4& b, float4& c) {
    const uint32_t tb = i * 48u;
    a = ld4off(tris, tb, 0); b = ld4off(tris, tb, 16); c = ld4off(tris, tb, 32);
}

// Closest (ANY false) or any hit among the scene's spheres and rectangles:
// an ordered walk of their BVH, one primitive test per item, starting and
// ending with an empty stack.  Updates tbest and prim on a nearer hit.
template <int SHORT, bool ANY>
DEV bool trace_items(const DevScene& s, const PRay& r, const PlaneSel& ps, float tmin, float& tbest, int& prim,
                     TraversalStack<SHORT>& st) {
    bool found = false;
    uint32_t cur = 0;
    for (;;) {
        if (cur & LEAF_FLAG_DEV) {
            const uint32_t first = (cur & 0x7fffffffu) >> 3, count = (cur & 7u) + 1u;
            for (uint32_t k = 0; k < count; k++) {
                const int pi = s.item_prims[first + k];
                const float t = analytic_t(s.prims[pi], r.o, r.d, tmin, tbest);
                if (t != 0.0f && t < tbest) {
                    tbest = t; prim = pi; found = true;
                    if (ANY) { st.sp = 0; return true; }
                }
            }
            if (st.sp == 0) break;
            cur = (uint32_t)st.pop();
        } else {
            cur = visit_node<SHORT, !ANY>(s.item_nodes, cur, r, ps, tmin, tbest, st);
            if (cur == TRAVERSAL_DONE) break;
        }
    }
    return found;
}

// A point or a direction through a 3 x 4 row-major matrix.
DEV V xf34_point(const float* m, V p) {
    return mk(fmaf(m[0], p.x, fmaf(m[1], p.y, fmaf(m[2], p.z, m[3]))),
              fmaf(m[4], p.x, fmaf(m[5], p.y, fmaf(m[6], p.z, m[7]))),
              fmaf(m[8], p.x, fmaf(m[9], p.y, fmaf(m[10], p.z, m[11]))));
}
DEV V xf34_vec(const float* m, V v) {
    return mk(fmaf(m[0], v.x, fmaf(m[1], v.y, m[2] * v.z)), fmaf(m[4], v.x, fmaf(m[5], v.y, m[6] * v.z)),
              fmaf(m[8], v.x, fmaf(m[9], v.y, m[10] * v.z)));
}

// Closest (ANY false) or any hit among the instances of shared meshes: an
// ordered walk of the top-level tree over their world boxes; at an instance
// whose own box the segment meets, the ray is carried into its mesh's object
// space, where its direction keeps its scale so that distances along it are
// the world ray's, and the mesh's tree is walked on the same stack above the
// top level's pending entries (the rest of a top-level leaf waits there as a
// leaf reference stepped to its next instance).  One loop serves both levels,
// so a wave's node steps share one code path whichever tree each lane is in.
// Updates tbest, prim and tri on a nearer hit.
template <int SHORT, bool ANY>
DEV bool trace_instances(const DevScene& s, V o, V d, float tmin, float& tbest, int& prim, int& tri,
                         TraversalStack<SHORT>& st) {
    bool found = false;
    PRay r = prep_ray(o, d);
    PlaneSel ps = plane_sel(r);
    int inst_prim = -1;  // the instance whose mesh the walk is in (-1: the top level)
    int base = 0;        // stack entries below the mesh's
    uint32_t cur = 0;
    for (;;) {
        if (!(cur & LEAF_FLAG_DEV)) {
            cur = visit_node<SHORT, !ANY>(s.inst_nodes, cur, r, ps, tmin, tbest, st, base);
        } else if (inst_prim >= 0) {
            // a leaf of the mesh: its triangles, in object space
            const uint32_t f2 = (cur & 0x7fffffffu) >> 3, n2 = (cur & 7u) + 1u;
            for (uint32_t j = 0; j < n2; j++) {
                float4 a, b, c;
                load_tri(s.inst_tris, f2 + j, a, b, c);
                const float t = tri_t(f4v(a), f4v(b), f4v(c), r.o, r.d, r.kx, r.ky, r.kz, r.sx, r.sy, r.sz, tmin, tbest);
                if (t != 0.0f && t < tbest) {
                    tbest = t; prim = inst_prim; tri = __float_as_int(a.w); found = true;
                    if (ANY) { st.sp = 0; return true; }
                }
            }
            cur = st.sp == base ? TRAVERSAL_DONE : (uint32_t)st.pop();
        } else {
            // a top-level leaf: its first instance; the others wait
            const uint32_t first = (cur & 0x7fffffffu) >> 3;
            if (cur & 7u) st.push((int)(cur + 7u));
            const DevInstance& in = s.instances[first];
            // the instance's own world box against the current segment
            // (widened as the node test widens its exit distance)
            const float ax = in.lo[0] * r.inv.x - r.oinv.x, bx = in.hi[0] * r.inv.x - r.oinv.x;
            const float ay = in.lo[1] * r.inv.y - r.oinv.y, by = in.hi[1] * r.inv.y - r.oinv.y;
            const float az = in.lo[2] * r.inv.z - r.oinv.z, bz = in.hi[2] * r.inv.z - r.oinv.z;
            const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tmin));
            const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tbest));
            if (tn <= tf * 1.00000048f) {
                inst_prim = in.prim;
                base = st.sp;
                r = prep_ray(xf34_point(in.w2o, o), xf34_vec(in.w2o, d));
                ps = plane_sel(r);
                cur = (uint32_t)in.root;
            } else {
                cur = st.sp == 0 ? TRAVERSAL_DONE : (uint32_t)st.pop();
            }
        }
        if (cur == TRAVERSAL_DONE) {
            if (inst_prim < 0) break;
            // the mesh is done: back to the top level's pending entries
            inst_prim = -1;
            base = 0;
            r = prep_ray(o, d);
            ps = plane_sel(r);
            if (st.sp == 0) break;
            cur = (uint32_t)st.pop();
        }
    }
    return found;
}

// The three vertices of triangle record `tri` in world space (an instance's
// records are in its mesh's object space).
DEV void tri_world(const DevScene& s, const DevPrim& p, int tri, V& a, V& b, V& c) {
    const size_t rb = 8 * (size_t)tri;
    a = f4v(ld4g(s.tshade, rb)); b = f4v(ld4g(s.tshade, rb + 1)); c = f4v(ld4g(s.tshade, rb + 2));
    if (p.xformed & PRIM_INSTANCE) { a = xf34_point(p.data, a); b = xf34_point(p.data, b); c = xf34_point(p.data, c); }
}

// With REFINE (closest hit only), the winner's distance is recomputed in
// FP64 for the ray-query outputs; the path tracer keeps the float distance of
// the reference intersection routines, from which its next rays start.
template <bool ANY, int SHORT, bool REFINE = false>
DEV bool trace(const DevScene& s, V o, V d, float tmin, float tmax, HitRec& hit, TraversalStack<SHORT>& st,
               VisitCount& vc) {
    bool found = false;
    hit.t = tmax; hit.prim = -1; hit.tri = -1;
    // infinite planes (and a handful of spheres and rectangles): linear, in primitive order
    for (int i = 0; i < s.n_analytic; i++) {
        int pi = s.analytic[i];
        float t = analytic_t(s.prims[pi], o, d, tmin, hit.t);
        if (t != 0.0f && t < hit.t) {
            hit.t = t; hit.prim = pi; hit.tri = -1; found = true;
            if (ANY) return true;
        }
    }
    PRay r = prep_ray(o, d);
    const PlaneSel ps = plane_sel(r);
    st.sp = 0;
    if (s.n_items > 0 && trace_items<SHORT, ANY>(s, r, ps, tmin, hit.t, hit.prim, st)) {
        hit.tri = -1; found = true;
        if (ANY) return true;
    }
    if (s.n_instances > 0 && trace_instances<SHORT, ANY>(s, o, d, tmin, hit.t, hit.prim, hit.tri, st)) {
        found = true;
        if (ANY) return true;
    }
    if (!s.has_bvh) return found;
    // Speculative while-while traversal: a lane that reaches a leaf parks it
    // and keeps descending inner nodes until all but TRAVERSAL_STRAGGLERS
    // lanes of the wave have a parked leaf (or are done); then the wave tests
    // leaves together, until all but that many lanes have run out of leaves.
    // A lane resumes where it stopped in the next round.
    constexpr uint32_t DONE = TRAVERSAL_DONE;
    uint32_t cur = 0;                       // root node
    uint32_t leaf = 0;                      // parked leaf reference (0: none)
    while (cur != DONE || leaf != 0) {
        while (cur < DONE) {  // inner node
            vc.nodes++;
            cur = visit_node<SHORT, !ANY>(s.nodes, cur, r, ps, tmin, hit.t, st);
            if ((cur & LEAF_FLAG_DEV) && leaf == 0) {
                leaf = cur;
                cur = st.sp == 0 ? DONE : (uint32_t)st.pop();
            }
            if (__popcll(__ballot(leaf == 0)) <= TRAVERSAL_STRAGGLERS) break;
        }
        while (leaf != 0) {
            // one triangle per round (the reference steps by 7, as in k_trace_queue)
            const uint32_t first = (leaf & 0x7fffffffu) >> 3;
            vc.tris++;
            float4 a, b, c;
            load_tri(s.tris, first, a, b, c);
            float t = tri_t(f4v(a), f4v(b), f4v(c), r.o, r.d, r.kx, r.ky, r.kz, r.sx, r.sy, r.sz, tmin, hit.t);
            if (t != 0.0f && t < hit.t) {
                hit.t = t; hit.tri = __float_as_int(a.w); hit.prim = __float_as_int(b.w); found = true;
                if (ANY) return true;
            }
            leaf = (leaf & 7u) ? leaf + 7u : 0u;
            if (leaf == 0 && (cur & LEAF_FLAG_DEV)) {
                leaf = cur;
                cur = st.sp == 0 ? DONE : (uint32_t)st.pop();
            }
            if (__popcll(__ballot(leaf != 0)) <= TRAVERSAL_STRAGGLERS) break;
        }
    }
    if (REFINE && !ANY && found && hit.tri < 0) hit.t = analytic_t64(s.prims[hit.prim], o, d, hit.t);
    if (REFINE && !ANY && found && hit.tri >= 0) {
        // The float watertight t loses digits on small, oblique triangles;
        // the winner's distance to its plane is recomputed in FP64.
        V a, b, c;
        tri_world(s, s.prims[hit.prim], hit.tri, a, b, c);
        double e1x = (double)b.x - a.x, e1y = (double)b.y - a.y, e1z = (double)b.z - a.z;
        double e2x = (double)c.x - a.x, e2y = (double)c.y - a.y, e2z = (double)c.z - a.z;
        double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
        double den = nx * d.x + ny * d.y + nz * d.z;
        if (den != 0.0) {
            double t = (nx * ((double)a.x - o.x) + ny * ((double)a.y - o.y) + nz * ((double)a.z - o.z)) / den;
            if (t > 0.0) hit.t = (float)t;
        }
    }
    return found;
}

template <bool ANY, int SHORT, bool REFINE = false>
DEV bool trace(const DevScene& s, V o, V d, float tmin, float tmax, HitRec& hit, TraversalStack<SHORT>& st) {
    VisitCount vc;
    return trace<ANY, SHORT, REFINE>(s, o, d, tmin, tmax, hit, st, vc);
}

// ---------------------------------------------------------------------------
// Surface reconstruction (world space)
// ---------------------------------------------------------------------------
struct ONB { V u, v, w; };
// M/geometry/Vector3F.java:527-545
DEV V compute_v(V w) {
    V n = nrm(w);
    float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
    if (ax < ay && ax < az) return nrm(mk(0.0f, n.z, -n.y));
    if (ay < az) return nrm(mk(n.z, 0.0f, -n.x));
    return nrm(mk(n.y, -n.x, 0.0f));
}
DEV ONB onb_w(V w) { ONB o; o.w = nrm(w); o.v = compute_v(w); o.u = cross(o.v, o.w); return o; }
DEV V to_local(V v, const ONB& b) { return nrm(mk(dot(v, b.u), dot(v, b.v), dot(v, b.w))); }
DEV V to_world(V v, const ONB& b) {
    return nrm(mk(v.x * b.u.x + v.y * b.v.x + v.z * b.w.x, v.x * b.u.y + v.y * b.v.y + v.z * b.w.y,
                  v.x * b.u.z + v.y * b.v.z + v.z * b.w.z));
}

struct Surf {
    V p;
    ONB g, s;
    int prim;
    int tri;
    // triangles: the hit's barycentric u, v (weights of b and c); spheres:
    // the texture coordinates (Sphere3F.java:340-372)
    float bu, bv;
};

// A row-vector product: (dot(X[0], v), dot(X[1], v), dot(X[2], v)).
DEV V mul_rows(const float4* X, V v) {
    const float4 a = X[0], b = X[1], c = X[2];
    return mk(a.x * v.x + a.y * v.y + a.z * v.z, b.x * v.x + b.y * v.y + b.z * v.z, c.x * v.x + c.y * v.y + c.z * v.z);
}
// The frame of a transformed primitive at world normal `w`: built in object
// space around the normal's object-space direction and carried to world
// space vector by vector, each normalised (AbstractSceneKernel.java:1104-1217).
DEV ONB frame_xformed(const float4* X, V w) {
    const ONB o = onb_w(nrm(mul_rows(X, w)));
    ONB r;
    r.u = nrm(mul_rows(X + 3, o.u));
    r.v = nrm(mul_rows(X + 3, o.v));
    r.w = nrm(mul_rows(X + 3, o.w));
    return r;
}

DEV Surf make_surface(const DevScene& sc, V o, V d, const HitRec& h) {
    Surf s;
    s.prim = h.prim; s.tri = h.tri;
    s.bu = 0.0f; s.bv = 0.0f;
    s.p = o + d * h.t;
    const DevPrim& p = sc.prims[h.prim];
    const float4* X = (p.xformed & PRIM_FRAME_XFORM) ? sc.xforms + 6 * (size_t)h.prim : nullptr;
    if (h.tri >= 0) {
        const size_t rb = 8 * (size_t)h.tri;
        const float4 r0 = ld4g(sc.tshade, rb), r1 = ld4g(sc.tshade, rb + 1), r2 = ld4g(sc.tshade, rb + 2);
        const float4 r3 = ld4g(sc.tshade, rb + 3), r4 = ld4g(sc.tshade, rb + 4), r5 = ld4g(sc.tshade, rb + 5);
        V a = f4v(r0), b = f4v(r1), c = f4v(r2);
        V na = mk(r0.w, r1.w, r2.w), nb = f4v(r3), nc = f4v(r4), ng = mk(r3.w, r4.w, r5.x);
        if (__builtin_expect(p.xformed & PRIM_INSTANCE, 0)) {
            // an instance: its mesh's record is in object space; vertices
            // through the matrix, normals through the normal matrix N (the
            // scene compiler's world-space geometric normal of an unshared
            // mesh, cross(A e1, A e2) = det(A) N cross(e1, e2) flipped back to
            // the object-space side under a reflection, is N ng normalised)
            a = xf34_point(p.data, a); b = xf34_point(p.data, b); c = xf34_point(p.data, c);
            if (X) {
                na = mul_rows(X + 3, na); nb = mul_rows(X + 3, nb); nc = mul_rows(X + 3, nc);
                ng = nrm(mul_rows(X + 3, ng));
            }
        }
        V eab = b - a, eca = a - c;
        V d1 = cross(eab, eca);
        float rdet = __builtin_amdgcn_rcpf(dot(d, d1));
        V d3 = cross(a - o, d);
        float u = dot(d3, eca) * rdet, v = dot(d3, eab) * rdet, w = 1.0f - u - v;
        s.bu = u; s.bv = v;
        V ns = nrm(na * w + nb * u + nc * v);
        if (X) {
            s.g = frame_xformed(X, ng);
            s.s = frame_xformed(X, ns);
        } else {
            s.g = onb_w(ng);
            s.s = onb_w(ns);  // texture coordinates are absent: UV determinant 0
        }
    } else if (p.shape == DFR_SHAPE_SPHERE) {
        V c = ld3(p.data);
        float radius = p.data[3];
        V ph = (o - c) + d * h.t;
        ph = ph * (radius / len(ph));
        const V pw = ph;
        if (X) ph = mul_rows(X, ph);  // the sphere's own (rotated) coordinates
        if (ph.x == 0.0f && ph.y == 0.0f) ph.x = 1.0e-5f * radius;
        float theta = acosf(fminf(fmaxf(ph.z / radius, -1.0f), 1.0f));
        float zr = sqrtf(ph.x * ph.x + ph.y * ph.y);
        float cphi = ph.x / zr, sphi = ph.y / zr;
        V dpdu = mk(-2.0f * PI_F * ph.y, 2.0f * PI_F * ph.x, 0.0f);
        V dpdv = mk(ph.z * cphi, ph.z * sphi, -radius * sinf(theta)) * (0.0f - PI_F);
        V n = nrm(cross(dpdu, dpdv));
        s.g.w = nrm(n); s.g.v = nrm(dpdv); s.g.u = nrm(dpdu);
        float phi = atan2f(ph.y, ph.x);
        if (phi < 0.0f) phi += 2.0f * PI_F;
        s.bu = phi * (1.0f / (2.0f * PI_F));
        s.bv = (theta - PI_F) / (0.0f - PI_F);
        if (X) {
            s.g.u = nrm(mul_rows(X + 3, s.g.u)); s.g.v = nrm(mul_rows(X + 3, s.g.v)); s.g.w = nrm(mul_rows(X + 3, s.g.w));
            s.p = pw + c;
        } else {
            s.p = ph + c;
        }
        s.s = s.g;
    } else if (p.shape == DFR_SHAPE_RECTANGLE) {
        s.g = X ? frame_xformed(X, ld3(p.normal)) : onb_w(ld3(p.normal));
        s.s = s.g;
    } else {
        s.g.u = nrm(ld3(p.data + 3)); s.g.v = nrm(ld3(p.data + 6)); s.g.w = nrm(ld3(p.data + 9));
        s.s = s.g;
    }
    if (dot(s.s.w, s.g.w) < 0.0f) s.s.w = -s.s.w;
    return s;
}

// ---------------------------------------------------------------------------
// Camera.createPrimaryRay (M/scene/Camera.java:243-291) over the 19-float
// camera record.  The thin lens (aperture radius above 0.00001) moves the
// origin to a uniform point of the aperture disk (SampleGeneratorF
// .sampleDiskUniformDistribution, M/geometry/SampleGeneratorF.java:68-76)
// drawn from dimensions DIM_LENS and DIM_LENS + 1; every ray of a pixel
// position then passes through the same point of the focal plane.
// ---------------------------------------------------------------------------
constexpr uint32_t DIM_LENS = 0xfffffff0u;
DEV bool primary_ray(const float* cam, float ix, float iy, float px, float py, const Rng& rng, V& o, V& d) {
    float cx = 2.0f * ((ix + px) / (cam[DFR_CAMERA_RES_X] - 1.0f)) - 1.0f;
    float cy = 2.0f * ((iy + py) / (cam[DFR_CAMERA_RES_Y] - 1.0f)) - 1.0f;
    float wf = 1.0f;
    if ((int)cam[DFR_CAMERA_LENS] == DFR_LENS_FISHEYE) {
        float dd = cx * cx + cy * cy;
        if (dd > 1.0f) return false;
        wf = sqrtf(1.0f - dd);
    }
    V U = ld3(cam + 3), Vv = ld3(cam + 6), W = ld3(cam + 9), eye = ld3(cam + 12);
    V on_plane = (U * (cam[0] * cx) + Vv * (cam[1] * cy)) + (eye + W * wf);
    // Point3F.add(eye, Vector3F.direction(eye, onPlane), focalDistance)
    V on_image = eye + (on_plane - eye) * cam[DFR_CAMERA_FOCAL];
    o = eye;
    const float aperture = cam[DFR_CAMERA_APERTURE];
    if (aperture > 0.00001f) {
        const float r = sqrtf(rng.get(DIM_LENS)), theta = 2.0f * PI_F * rng.get(DIM_LENS + 1);
        const float dx = r * cosf(theta), dy = r * sinf(theta);
        o = eye + (U * (dx * aperture) + Vv * (dy * aperture));
    }
    d = nrm(on_image - o);
    return true;
}

// ---------------------------------------------------------------------------
// BxDFs (M/scene/bxdf/*, M/scene/microfacet/*, M/scene/fresnel/*)
// ---------------------------------------------------------------------------
// The shading-frame trigonometry, the dielectric Fresnel term and the
// Trowbridge-Reitz terms below use the hardware square root and reciprocal
// (about an ulp each) where a correctly rounded form costs ten instructions.
DEV float cos_t(V v) { return v.z; }
DEV float cos_abs(V v) { return fabsf(v.z); }
DEV float sin2_t(V v) { return fmaxf(0.0f, 1.0f - v.z * v.z); }
DEV float sin_t(V v) { return __builtin_amdgcn_sqrtf(sin2_t(v)); }
DEV float cos_phi(V v) { float s = sin_t(v); return s == 0.0f ? 1.0f : fminf(fmaxf(v.x * __builtin_amdgcn_rcpf(s), -1.0f), 1.0f); }
DEV float sin_phi(V v) { float s = sin_t(v); return s == 0.0f ? 0.0f : fminf(fmaxf(v.y * __builtin_amdgcn_rcpf(s), -1.0f), 1.0f); }
DEV bool same_hemi(V a, V b) { return a.z * b.z > 0.0f; }
DEV V face_fwd(V a, V b) { return dot(a, b) < 0.0f ? -a : a; }
DEV V face_fwd_z(V a, V b) { return a.z < 0.0f ? mk(b.x, b.y, -b.z) : b; }
DEV V reflect_pbrt(V wo, V n) { return -wo + n * (2.0f * dot(wo, n)); }
DEV bool refract(V d, V n, float eta, V& out) {
    float ci = dot(d, n);
    float s2i = fmaxf(0.0f, 1.0f - ci * ci);
    float s2t = eta * eta * s2i;
    if (s2t >= 1.0f) return false;
    float ct = sqrtf(1.0f - s2t);
    out = (-d) * eta + n * (eta * ci - ct);
    return true;
}

DEV float fr_dielectric(float cosi, float etai, float etat) {
    float c = fminf(fmaxf(cosi, -1.0f), 1.0f);
    bool entering = c > 0.0f;
    float ci = entering ? c : fabsf(c);
    float ei = entering ? etai : etat, et = entering ? etat : etai;
    float si = __builtin_amdgcn_sqrtf(fmaxf(0.0f, 1.0f - ci * ci));
    float st = ei * __builtin_amdgcn_rcpf(et) * si;
    if (st >= 1.0f) return 1.0f;
    float ct = __builtin_amdgcn_sqrtf(fmaxf(0.0f, 1.0f - st * st));
    float rpa = ((et * ci) - (ei * ct)) * __builtin_amdgcn_rcpf((et * ci) + (ei * ct));
    float rpe = ((ei * ci) - (et * ct)) * __builtin_amdgcn_rcpf((ei * ci) + (et * ct));
    return (rpa * rpa + rpe * rpe) / 2.0f;
}
DEV float fr_conductor(float cosi, float eta, float k) {
    float c = fminf(fmaxf(cosi, -1.0f), 1.0f);
    float c2 = c * c, s2 = 1.0f - c2;
    float eta2 = eta * eta, k2 = k * k;
    float t0 = eta2 - k2 - s2;
    float a2b2 = sqrtf(t0 * t0 + eta2 * k2 * 4.0f);
    float t1 = a2b2 + c2;
    float a = sqrtf((a2b2 + t0) * 0.5f);
    float t2 = a * (c * 2.0f);
    float rs = (t1 - t2) / (t1 + t2);
    float t3 = c2 * a2b2 + s2 * s2;
    float t4 = t2 * s2;
    float rp = rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (rp + rs);
}
DEV float schlick_w(float c) {
    float m = fminf(fmaxf(1.0f - c, 0.0f), 1.0f);
    return (m * m) * (m * m) * m;
}
DEV float lerpf(float a, float b, float t) { return (1.0f - t) * a + t * b; }
DEV float pow5(float x) { return x * x * x * x * x; }

DEV V fresnel(const DevLobe& l, float cosi) {
    switch (l.fresnel) {
        case FR_CONST: return mk(1.0f, 1.0f, 1.0f);
        case FR_DIELECTRIC: { float f = fr_dielectric(cosi, l.frEtaI, l.frEtaT); return mk(f, f, f); }
        case FR_CONDUCTOR: {
            float c = fabsf(cosi);
            return mk(fr_conductor(c, l.condEta[0], l.condK[0]), fr_conductor(c, l.condEta[1], l.condK[1]),
                      fr_conductor(c, l.condEta[2], l.condK[2]));
        }
        default: {
            float die = fr_dielectric(cosi, 1.0f, l.disneyEta);
            float sw = schlick_w(cosi);
            V r;
            r.x = lerpf(die, lerpf(l.r0[0], 1.0f, sw), l.disneyMetallic);
            r.y = lerpf(die, lerpf(l.r0[1], 1.0f, sw), l.disneyMetallic);
            r.z = lerpf(die, lerpf(l.r0[2], 1.0f, sw), l.disneyMetallic);
            return r;
        }
    }
}

// Trowbridge-Reitz (visible-area sampling on)
DEV float tr_D(const DevLobe& l, V h) {
    float s2 = sin2_t(h), c2 = h.z * h.z;
    float t2 = s2 * __builtin_amdgcn_rcpf(c2);
    if (isinf(t2)) return 0.0f;
    float c4 = c2 * c2;
    float cp = cos_phi(h), sp = sin_phi(h);
    float e = (cp * cp * __builtin_amdgcn_rcpf(l.alphaX * l.alphaX) + sp * sp * __builtin_amdgcn_rcpf(l.alphaY * l.alphaY)) * t2;
    return __builtin_amdgcn_rcpf(PI_F * l.alphaX * l.alphaY * c4 * (1.0f + e) * (1.0f + e));
}
DEV float tr_lambda(const DevLobe& l, V w) {
    float ta = fabsf(sin_t(w) * __builtin_amdgcn_rcpf(w.z));
    if (isinf(ta)) return 0.0f;
    float cp = cos_phi(w), sp = sin_phi(w);
    float alpha = __builtin_amdgcn_sqrtf(cp * cp * l.alphaX * l.alphaX + sp * sp * l.alphaY * l.alphaY);
    float a2 = (alpha * ta) * (alpha * ta);
    return (-1.0f + __builtin_amdgcn_sqrtf(1.0f + a2)) * 0.5f;
}
DEV float tr_G1(const DevLobe& l, V w) { return __builtin_amdgcn_rcpf(1.0f + tr_lambda(l, w)); }
DEV float tr_G(const DevLobe& l, V wo, V wi) {
    return l.separable ? tr_G1(l, wo) * tr_G1(l, wi) : __builtin_amdgcn_rcpf(1.0f + tr_lambda(l, wo) + tr_lambda(l, wi));
}
DEV float tr_pdf(const DevLobe& l, V wo, V h) { return tr_D(l, h) * tr_G1(l, wo) * fabsf(dot(wo, h)) * __builtin_amdgcn_rcpf(cos_abs(wo)); }
DEV void tr_slope(float ct, float u, float v, float& sx, float& sy) {
    if (ct > 0.9999f) {
        float r = __builtin_amdgcn_sqrtf(u * __builtin_amdgcn_rcpf(1.0f - u)), phi = 2.0f * PI_F * v;
        sx = r * __cosf(phi); sy = r * __sinf(phi); return;
    }
    float st = __builtin_amdgcn_sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float tt = st * __builtin_amdgcn_rcpf(ct);
    const float itt = __builtin_amdgcn_rcpf(tt);
    float a = 2.0f * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_sqrtf(1.0f + __builtin_amdgcn_rcpf(itt * itt)));
    float b = 2.0f * u * __builtin_amdgcn_rcpf(a) - 1.0f;
    float c = fminf(__builtin_amdgcn_rcpf(b * b - 1.0f), 1.0e10f);
    float dd = tt;
    float e = __builtin_amdgcn_sqrtf(fmaxf(dd * dd * c * c - (b * b - dd * dd) * c, 0.0f));
    float f = dd * c - e, g = dd * c + e;
    float hs = v > 0.5f ? 1.0f : -1.0f;
    float i = v > 0.5f ? 2.0f * (v - 0.5f) : 2.0f * (0.5f - v);
    float j = (i * (i * (i * 0.27385f - 0.73369f) + 0.46341f)) * __builtin_amdgcn_rcpf(i * (i * (i * 0.093073f + 0.309420f) - 1.0f) + 0.597999f);
    float x = (b < 0.0f || g > itt) ? f : g;
    sx = x; sy = hs * j * __builtin_amdgcn_sqrtf(1.0f + x * x);
}
DEV V tr_sample_vis(const DevLobe& l, V wi, float u, float v) {
    V s = nrm(mk(wi.x * l.alphaX, wi.y * l.alphaY, wi.z));
    float sx, sy;
    tr_slope(s.z, u, v, sx, sy);
    float cp = cos_phi(s), sp = sin_phi(s);
    float x = -((cp * sx - sp * sy) * l.alphaX);
    float y = -((sp * sx + cp * sy) * l.alphaY);
    return nrm(mk(x, y, 1.0f));
}
DEV V tr_sample_h(const DevLobe& l, V wo, float u, float v) {
    return wo.z >= 0.0f ? tr_sample_vis(l, wo, u, v) : -tr_sample_vis(l, -wo, u, v);
}

DEV float gtr1(float c, float alpha) {
    float a = alpha * alpha;
    return (a - 1.0f) / (PI_F * logf(a) * (1.0f + (a - 1.0f) * c * c));
}
DEV float smith_ggx(float c, float alpha) {
    float a = alpha * alpha, b = c * c;
    return 1.0f / (c + sqrtf(a + b - a * b));
}

// Lobes of the glossy, substrate, translucent, clear coat and uber
// materials; reached only from kernels instantiated with EXT = true (scenes
// with these materials or with spot or directional lights), so the kernels
// of the other scenes carry none of their code or registers.
DEV V lobe_f_ext(const DevLobe& l, V wo, V wi) {
    switch (l.kind) {
        case LOBE_FRESNEL_BLEND: {  // diffuse under a Schlick-Fresnel glossy coat
            V h = wi + wo;
            if (h.x == 0.0f && h.y == 0.0f && h.z == 0.0f) return mk(0, 0, 0);
            h = nrm(h);
            const float ih = dot(wi, h), ci = cos_abs(wi), co = cos_abs(wo);
            // scalar factors: diffuse 28/(23 pi) (1 - (1 - ci/2)^5) (1 - (1 - co/2)^5),
            // glossy D / (4 |wi.h| max(ci, co)), Schlick weight (1 - wi.h)^5
            const float kd = (28.0f / (23.0f * PI_F)) * (1.0f - pow5(1.0f - 0.5f * ci)) * (1.0f - pow5(1.0f - 0.5f * co));
            const float kg = tr_D(l, h) / (4.0f * fabsf(ih) * fmaxf(ci, co));
            const float w = pow5(1.0f - ih);
            V r;
            r.x = l.R[0] * (1.0f - l.T[0]) * kd + (l.T[0] + (1.0f - l.T[0]) * w) * kg;
            r.y = l.R[1] * (1.0f - l.T[1]) * kd + (l.T[1] + (1.0f - l.T[1]) * w) * kg;
            r.z = l.R[2] * (1.0f - l.T[2]) * kd + (l.T[2] + (1.0f - l.T[2]) * w) * kg;
            return r;
        }
        case LOBE_LAMBERT_T: return ld3(l.T) * INV_PI;
        default: return mk(0, 0, 0);
    }
}
DEV float lobe_pdf_ext(const DevLobe& l, V wo, V wi) {
    switch (l.kind) {
        case LOBE_FRESNEL_BLEND: {  // half cosine-weighted, half microfacet
            if (!same_hemi(wo, wi)) return 0.0f;
            V h = nrm(wo + wi);
            return 0.5f * (cos_abs(wi) * INV_PI + tr_pdf(l, wo, h) / (4.0f * dot(wo, h)));
        }
        case LOBE_LAMBERT_T: return same_hemi(wo, wi) ? 0.0f : cos_abs(wi) * INV_PI;
        default: return 0.0f;
    }
}

template <bool EXT>
DEV V lobe_f(const DevLobe& l, V wo, V wi) {
    if constexpr (EXT) {
        if (l.kind >= LOBE_FRESNEL_BLEND) return lobe_f_ext(l, wo, wi);
    }
    switch (l.kind) {
        case LOBE_LAMBERT: return ld3(l.R) * INV_PI;
        case LOBE_OREN_NAYAR: {
            float ci = cos_abs(wi), co = cos_abs(wo), si = sin_t(wi), so = sin_t(wo);
            float mc = (si > 1.0e-4f && so > 1.0e-4f) ? fmaxf(0.0f, cos_phi(wi) * cos_phi(wo) + sin_phi(wi) * sin_phi(wo)) : 0.0f;
            float sa = ci > co ? so : si;
            float tb = ci > co ? si / ci : so / co;
            return ld3(l.R) * INV_PI * (l.A + l.B * mc * sa * tb);
        }
        case LOBE_TS_R: {
            float co = cos_abs(wo), ci = cos_abs(wi);
            if (co == 0.0f || ci == 0.0f) return mk(0, 0, 0);
            V h = wi + wo;
            if (h.x == 0.0f && h.y == 0.0f && h.z == 0.0f) return mk(0, 0, 0);
            h = nrm(h);
            float D = tr_D(l, h), G = tr_G(l, wo, wi);
            V F = fresnel(l, dot(wi, face_fwd(h, mk(0, 0, 1))));
            return (ld3(l.R) * D * G) * F / (4.0f * ci * co);
        }
        case LOBE_TS_T: {
            if (same_hemi(wo, wi)) return mk(0, 0, 0);
            float ci = cos_t(wi), co = cos_t(wo);
            if (ci == 0.0f || co == 0.0f) return mk(0, 0, 0);
            float eta = co > 0.0f ? l.etaB / l.etaA : l.etaA / l.etaB;
            V h = nrm(wi * eta + wo);
            if (h.z < 0.0f) h = -h;
            float oh = dot(wo, h), ih = dot(wi, h);
            if (oh * ih > 0.0f) return mk(0, 0, 0);
            float F = fr_dielectric(oh, l.etaA, l.etaB);
            float a = oh + eta * ih;
            float b = 1.0f / eta;
            float D = tr_D(l, h), G = tr_G(l, wo, wi);
            float k = fabsf(D * G * eta * eta * fabsf(ih) * fabsf(oh) * b * b / (ci * co * a * a));
            return mk((1.0f - F) * l.T[0] * k, (1.0f - F) * l.T[1] * k, (1.0f - F) * l.T[2] * k);
        }
        case LOBE_DISNEY_DIFFUSE: {
            float fo = schlick_w(cos_abs(wo)), fi = schlick_w(cos_abs(wi));
            float k = INV_PI * (1.0f - fo / 2.0f) * (1.0f - fi / 2.0f);
            return ld3(l.R) * k;
        }
        case LOBE_DISNEY_RETRO: {
            V h = wi + wo;
            if (h.x == 0.0f && h.y == 0.0f && h.z == 0.0f) return mk(0, 0, 0);
            h = nrm(h);
            float cd = dot(wi, h);
            float fo = schlick_w(cos_abs(wo)), fi = schlick_w(cos_abs(wi));
            float a = 2.0f * l.roughness * cd * cd;
            float b = fo + fi + fo * fi * (a - 1.0f);
            return ld3(l.R) * (INV_PI * a * b);
        }
        case LOBE_DISNEY_SHEEN: {
            V h = wi + wo;
            if (h.x == 0.0f && h.y == 0.0f && h.z == 0.0f) return mk(0, 0, 0);
            h = nrm(h);
            return ld3(l.R) * schlick_w(dot(wi, h));
        }
        case LOBE_DISNEY_CLEARCOAT: {
            V h = wi + wo;
            if (h.x == 0.0f && h.y == 0.0f && h.z == 0.0f) return mk(0, 0, 0);
            h = nrm(h);
            float D = gtr1(cos_abs(h), l.gloss);
            float F = lerpf(0.04f, 1.0f, schlick_w(dot(wo, h)));
            float G = smith_ggx(cos_abs(wo), 0.25f) * smith_ggx(cos_abs(wi), 0.25f);
            float v = l.weight * G * F * D / 4.0f;
            return mk(v, v, v);
        }
        default: return mk(0, 0, 0);
    }
}

template <bool EXT>
DEV float lobe_pdf(const DevLobe& l, V wo, V wi) {
    if constexpr (EXT) {
        if (l.kind >= LOBE_FRESNEL_BLEND) return lobe_pdf_ext(l, wo, wi);
    }
    switch (l.kind) {
        case LOBE_LAMBERT: case LOBE_OREN_NAYAR: case LOBE_DISNEY_DIFFUSE: case LOBE_DISNEY_RETRO:
        case LOBE_DISNEY_SHEEN:
            return same_hemi(wo, wi) ? cos_abs(wi) * INV_PI : 0.0f;
        case LOBE_TS_R: {
            if (!same_hemi(wo, wi)) return 0.0f;
            V h = nrm(wo + wi);
            return tr_pdf(l, wo, h) / (4.0f * dot(wo, h));
        }
        case LOBE_TS_T: {
            if (same_hemi(wo, wi)) return 0.0f;
            float eta = cos_t(wo) > 0.0f ? l.etaB / l.etaA : l.etaA / l.etaB;
            V h = nrm(wi * eta + wo);
            float oh = dot(wo, h), ih = dot(wi, h);
            if (oh * ih > 0.0f) return 0.0f;
            float a = oh + eta * ih;
            float b = fabsf((eta * eta * ih) / (a * a));
            return tr_pdf(l, wo, h) * b;
        }
        case LOBE_DISNEY_CLEARCOAT: {
            if (!same_hemi(wo, wi)) return 0.0f;
            V h = wi + wo;
            if (h.x == 0.0f && h.y == 0.0f && h.z == 0.0f) return 0.0f;
            h = nrm(h);
            float ch = cos_abs(h);
            return gtr1(ch, l.gloss) * ch / (4.0f * dot(wo, h));
        }
        default: return 0.0f;
    }
}

DEV void disk_concentric(float u, float v, float& x, float& y) {
    if (u == 0.0f && v == 0.0f) { x = 0.0f; y = 0.0f; return; }
    float a = u * 2.0f - 1.0f, b = v * 2.0f - 1.0f;
    if (a * a > b * b) {
        float phi = (PI_F / 4.0f) * (b * __builtin_amdgcn_rcpf(a));
        x = a * __cosf(phi); y = a * __sinf(phi);
    } else {
        float phi = (PI_F / 2.0f) - (PI_F / 4.0f) * (a * __builtin_amdgcn_rcpf(b));
        x = b * __cosf(phi); y = b * __sinf(phi);
    }
}
DEV V hemi_cos(float u, float v) {
    float x, y;
    disk_concentric(u, v, x, y);
    return mk(x, y, __builtin_amdgcn_sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y)));
}

struct LSample { V f; V wi; float pdf; uint32_t type; };

DEV bool lobe_sample_ext(const DevLobe& l, V wo, float u, float v, LSample& out) {
    out.type = l.type;
    switch (l.kind) {
        case LOBE_FRESNEL_BLEND: {
            V wi;
            if (u < 0.5f) {
                wi = face_fwd_z(wo, hemi_cos(fminf(2.0f * u, 0.99999994f), v));
            } else {
                V h = tr_sample_h(l, wo, fminf(2.0f * (u - 0.5f), 0.99999994f), v);
                wi = reflect_pbrt(wo, h);
                if (!same_hemi(wo, wi)) return false;
            }
            out.wi = wi; out.f = lobe_f_ext(l, wo, wi); out.pdf = lobe_pdf_ext(l, wo, wi);
            return true;
        }
        case LOBE_LAMBERT_T: {  // cosine sample on the side opposite wo
            V s = hemi_cos(u, v);
            V wi = wo.z > 0.0f ? mk(s.x, s.y, -s.z) : s;
            out.wi = wi; out.f = lobe_f_ext(l, wo, wi); out.pdf = lobe_pdf_ext(l, wo, wi);
            return true;
        }
        case LOBE_SPECULAR_T: {
            bool entering = cos_t(wo) > 0.0f;
            float ei = entering ? l.etaA : l.etaB, et = entering ? l.etaB : l.etaA;
            V wi;
            if (!refract(wo, face_fwd(mk(0, 0, 1), wo), ei / et, wi)) return false;
            float F = fr_dielectric(cos_t(wi), l.etaA, l.etaB);
            out.wi = wi; out.f = ld3(l.T) * (1.0f - F) * ((ei * ei) / (et * et)) / cos_abs(wi); out.pdf = 1.0f;
            return true;
        }
        default: return false;
    }
}

template <bool EXT>
DEV bool lobe_sample(const DevLobe& l, V wo, float u, float v, LSample& out) {
    if constexpr (EXT) {
        if (l.kind >= LOBE_FRESNEL_BLEND) return lobe_sample_ext(l, wo, u, v, out);
    }
    out.type = l.type;
    switch (l.kind) {
        case LOBE_LAMBERT: case LOBE_OREN_NAYAR: case LOBE_DISNEY_DIFFUSE: case LOBE_DISNEY_RETRO:
        case LOBE_DISNEY_SHEEN: {
            V wi = face_fwd_z(wo, hemi_cos(u, v));
            out.wi = wi; out.f = lobe_f<false>(l, wo, wi); out.pdf = lobe_pdf<false>(l, wo, wi);
            return true;
        }
        case LOBE_SPECULAR_R: {
            V wi = mk(-wo.x, -wo.y, wo.z);
            out.wi = wi; out.f = fresnel(l, cos_t(wi)) * ld3(l.R) / cos_abs(wi); out.pdf = 1.0f;
            return true;
        }
        case LOBE_FRESNEL_SPECULAR: {
            float refl = fr_dielectric(cos_t(wo), l.etaA, l.etaB);
            if (u < refl) {
                V wi = mk(-wo.x, -wo.y, wo.z);
                out.type = BX_SPEC | BX_REFL;
                out.wi = wi; out.f = ld3(l.R) * refl / cos_abs(wi); out.pdf = refl;
                return true;
            }
            bool entering = cos_t(wo) > 0.0f;
            float ei = entering ? l.etaA : l.etaB, et = entering ? l.etaB : l.etaA;
            V wi;
            if (!refract(wo, face_fwd(mk(0, 0, 1), wo), ei / et, wi)) return false;
            out.type = BX_SPEC | BX_TRANS;
            out.wi = wi;
            out.f = ld3(l.T) * (1.0f - refl) * ((ei * ei) / (et * et)) / cos_abs(wi);
            out.pdf = 1.0f - refl;
            return true;
        }
        case LOBE_TS_R: {
            if (wo.z == 0.0f) return false;
            V h = tr_sample_h(l, wo, u, v);
            if (dot(wo, h) < 0.0f) return false;
            V wi = reflect_pbrt(wo, h);
            if (!same_hemi(wo, wi)) return false;
            out.wi = wi; out.f = lobe_f<false>(l, wo, wi); out.pdf = tr_pdf(l, wo, h) / (4.0f * dot(wo, h));
            return true;
        }
        case LOBE_TS_T: {
            if (wo.z == 0.0f) return false;
            V h = tr_sample_h(l, wo, u, v);
            if (dot(wo, h) < 0.0f) return false;
            float eta = cos_t(wo) > 0.0f ? l.etaA / l.etaB : l.etaB / l.etaA;
            V wi;
            if (!refract(wo, h, eta, wi)) return false;
            out.wi = wi; out.f = lobe_f<false>(l, wo, wi); out.pdf = lobe_pdf<false>(l, wo, wi);
            return true;
        }
        case LOBE_DISNEY_CLEARCOAT: {
            if (wo.z == 0.0f) return false;
            float a2 = l.gloss * l.gloss;
            float ct = sqrtf(fmaxf(0.0f, (1.0f - powf(a2, 1.0f - u)) / (1.0f - a2)));
            float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
            float phi = 2.0f * PI_F * v;
            V hs = mk(st * cosf(phi), st * sinf(phi), ct);
            V h = same_hemi(wo, hs) ? hs : -hs;
            V wi = reflect_pbrt(wo, h);
            if (!same_hemi(wo, wi)) return false;
            out.wi = wi; out.f = lobe_f<false>(l, wo, wi); out.pdf = lobe_pdf<false>(l, wo, wi);
            return true;
        }
        default: return false;
    }
}

// BSDF over a compiled material (M/scene/BSDF.java:272-540)
// Per-hit texture colours of a textured material: c[0], c[1] replace the
// R of lobes 0 and 1, c[2], c[3] their T (bits 0-3 of `set`); lobe i is
// absent at this hit when bit i of `off` is set.
struct TexHit {
    uint32_t set, off;
    V c[4];
    // textured roughness and conductor Fresnel: lobes (bit i: lobe i) whose
    // alphas become (ax, ay), whose conductor eta and k become (eta, k)
    uint32_t rough, cond;
    float ax, ay;
    V eta, k;
    // a textured dielectric eta (GlassMaterial's textureEta): lobes (bit i)
    // whose dielectric eta becomes `geta`, and the BSDF's eta
    uint32_t deta;
    float geta;
};

struct Bsdf {
    const DevLobe* lobes;
    int n;
    float eta;
    ONB s;
    V ng, wo_w, wo;
    TexHit th;
};
// ClearCoatMaterial.java:292-333: the material picks one layer per hit by
// Russian roulette on the dielectric (1 -> 1.5) reflectance of the shading
// normal; its compiled lobes are [specular KS, Lambertian KD].  Total
// internal reflection leaves the specular layer at weight 1.  Returns the
// layer's weight, which scales every value of f; the callers fold it into
// the path throughput, which multiplies both NEE and the continuation.
DEV float clear_coat_layer(Bsdf& b, const Surf& si, V d, float u) {
    const V n = si.s.w;
    const V nco = dot(n, d) > 0.0f ? -n : n;
    const bool entering = dot(n, nco) > 0.0f;
    const float eta = entering ? 1.0f / 1.5f : 1.5f;
    const float ci = dot(d, nco);
    const float s2t = eta * eta * fmaxf(0.0f, 1.0f - ci * ci);
    b.n = 1;
    if (s2t >= 1.0f) return 1.0f;
    const float ct = sqrtf(1.0f - s2t);
    const V refr = (-d) * eta + nco * (eta * ci - ct);
    const float refl = fr_dielectric(entering ? -ci : dot(refr, n), 1.0f, 1.5f);
    const float prr = 0.25f + 0.5f * refl;
    if (u < prr) return refl / prr;
    b.lobes += 1;
    // the diffuse layer is lobe 0 now: so is its textured colour
    b.th.c[0] = b.th.c[1];
    b.th.set >>= 1;
    b.th.off >>= 1;
    return (1.0f - refl) / (1.0f - prr);
}
// `beta` is the path throughput at the hit (after its emission is added).
template <bool EXT>
DEV Bsdf make_bsdf(const DevScene& sc, const DevMaterial& m, const Surf& si, V ray_d, const Rng& rng, uint32_t dim,
                   V& beta, const TexHit& th) {
    Bsdf b;
    b.lobes = sc.lobes + m.first_lobe;
    b.n = m.n_lobes;
    if constexpr (EXT) b.th = th; else { b.th.set = 0u; b.th.off = 0u; b.th.rough = 0u; b.th.cond = 0u; b.th.deta = 0u; }
    b.eta = m.eta;
    if constexpr (EXT) { if (th.deta) b.eta = th.geta; }
    b.s = si.s; b.ng = si.g.w;
    b.wo_w = -ray_d;
    b.wo = to_local(b.wo_w, si.s);
    if constexpr (EXT) {
        if (m.type == DFR_MAT_CLEAR_COAT) beta = beta * clear_coat_layer(b, si, ray_d, rng.get(dim + 9));
    }
    return b;
}
DEV bool matches(uint32_t lt, uint32_t t) { return (lt & t) == lt; }
DEV int count_nonspecular(const Bsdf& b) {
    int c = 0;
    for (int i = 0; i < b.n; i++) if (!((b.th.off >> i) & 1u) && !(b.lobes[i].type & BX_SPEC)) c++;
    return c;
}
// lobe i of the BSDF with the hit's texture values (EXT kernels only)
DEV bool lobe_textured(const Bsdf& b, int i) {
    return ((b.th.set & (5u << i)) | ((b.th.rough | b.th.cond | b.th.deta) & (1u << i))) != 0u;
}
DEV DevLobe textured_lobe(const Bsdf& b, int i) {
    DevLobe c = b.lobes[i];
    if (b.th.set & (1u << i)) { const V r = b.th.c[i]; c.R[0] = r.x; c.R[1] = r.y; c.R[2] = r.z; }
    if (b.th.set & (4u << i)) { const V r = b.th.c[2 + i]; c.T[0] = r.x; c.T[1] = r.y; c.T[2] = r.z; }
    if (b.th.rough & (1u << i)) { c.alphaX = b.th.ax; c.alphaY = b.th.ay; }
    if (b.th.cond & (1u << i)) {
        c.condEta[0] = b.th.eta.x; c.condEta[1] = b.th.eta.y; c.condEta[2] = b.th.eta.z;
        c.condK[0] = b.th.k.x; c.condK[1] = b.th.k.y; c.condK[2] = b.th.k.z;
    }
    if (b.th.deta & (1u << i)) { c.etaB = b.th.geta; c.frEtaT = b.th.geta; }
    return c;
}
template <bool EXT>
DEV V lobe_f_at(const Bsdf& b, int i, V wo, V wi) {
    if constexpr (EXT) {
        if (lobe_textured(b, i)) return lobe_f<EXT>(textured_lobe(b, i), wo, wi);
    }
    return lobe_f<EXT>(b.lobes[i], wo, wi);
}
template <bool EXT>
DEV bool lobe_sample_at(const Bsdf& b, int i, V wo, float u, float v, LSample& out) {
    if constexpr (EXT) {
        if (lobe_textured(b, i)) return lobe_sample<EXT>(textured_lobe(b, i), wo, u, v, out);
    }
    return lobe_sample<EXT>(b.lobes[i], wo, u, v, out);
}
template <bool EXT>
DEV float lobe_pdf_at(const Bsdf& b, int i, V wo, V wi) {
    if constexpr (EXT) {
        if (lobe_textured(b, i)) return lobe_pdf<EXT>(textured_lobe(b, i), wo, wi);
    }
    return lobe_pdf<EXT>(b.lobes[i], wo, wi);
}
DEV bool lobe_on(const Bsdf& b, int i) { return !((b.th.off >> i) & 1u); }
template <bool EXT>
DEV V bsdf_f(const Bsdf& b, uint32_t t, V wi_w) {
    V wi = to_local(wi_w, b.s);
    bool refl = dot(wi_w, b.ng) * dot(b.wo_w, b.ng) > 0.0f;
    V r = mk(0, 0, 0);
    for (int i = 0; i < b.n; i++) {
        const DevLobe& l = b.lobes[i];
        if (lobe_on(b, i) && matches(l.type, t) && ((refl && (l.type & BX_REFL)) || (!refl && (l.type & BX_TRANS))))
            r = r + lobe_f_at<EXT>(b, i, b.wo, wi);
    }
    return r;
}
template <bool EXT>
DEV float bsdf_pdf(const Bsdf& b, uint32_t t, V wi_w) {
    if (b.n == 0) return 0.0f;
    V wi = to_local(wi_w, b.s);
    if (b.wo.z == 0.0f) return 0.0f;
    float p = 0.0f; int m = 0;
    for (int i = 0; i < b.n; i++) {
        const DevLobe& l = b.lobes[i];
        if (!(lobe_on(b, i) && matches(l.type, t))) continue;
        m++;
        p += lobe_pdf_at<EXT>(b, i, b.wo, wi);
    }
    if (m > 1) p /= m;
    return p;
}
template <bool EXT>
DEV bool bsdf_sample(const Bsdf& b, uint32_t t, float u, float v, LSample& out) {
    if (b.wo.z == 0.0f) return false;
    int m = 0;
    for (int i = 0; i < b.n; i++) if (lobe_on(b, i) && matches(b.lobes[i].type, t)) m++;
    if (m == 0) return false;
    int k = min((int)floorf(u * m), m - 1);
    int sel = -1, c = 0;
    for (int i = 0; i < b.n; i++)
        if (lobe_on(b, i) && matches(b.lobes[i].type, t)) { if (c == k) sel = i; c++; }
    const DevLobe& L = b.lobes[sel];
    float ur = fminf(u * m - k, 0.99999994f);
    LSample ls;
    if (!lobe_sample_at<EXT>(b, sel, b.wo, ur, v, ls)) return false;
    V wi_w = to_world(ls.wi, b.s);
    V f = ls.f;
    float pdf = ls.pdf;
    if (!(L.type & BX_SPEC)) {
        bool refl = dot(wi_w, b.ng) * dot(b.wo_w, b.ng) > 0.0f;
        f = mk(0, 0, 0);
        for (int i = 0; i < b.n; i++) {
            const DevLobe& B = b.lobes[i];
            if (!lobe_on(b, i) || !matches(B.type, t)) continue;
            if (i != sel) pdf += lobe_pdf_at<EXT>(b, i, b.wo, ls.wi);
            if ((refl && (B.type & BX_REFL)) || (!refl && (B.type & BX_TRANS))) f = f + lobe_f_at<EXT>(b, i, b.wo, ls.wi);
        }
    }
    if (m > 1) pdf /= m;
    out.f = f; out.wi = wi_w; out.pdf = pdf; out.type = ls.type;
    return true;
}

// ---------------------------------------------------------------------------
// Lights
// ---------------------------------------------------------------------------
struct LightSmp { V L; V p; V wi; float pdf; };

#include "perez_device.inc"

// ---------------------------------------------------------------------------
// Textures (M/scene/texture/*Texture.java getColor) and the texture
// coordinates of a hit (Triangle3F: interpolated vertex coordinates;
// Sphere3F: (phi / 2 pi, (theta - pi) / -pi); Rectangle3F.java:448-475;
// Plane3F.java:280-290).  Only EXT kernel instantiations reach them.
// ---------------------------------------------------------------------------
DEV void hit_uv(const DevScene& sc, const Surf& si, float& u, float& v) {
    const DevPrim& p = sc.prims[si.prim];
    if (si.tri >= 0) {
        const size_t rb = 8 * (size_t)si.tri;
        const float4 r5 = ld4g(sc.tshade, rb + 5), r6 = ld4g(sc.tshade, rb + 6);
        const float w = 1.0f - si.bu - si.bv;
        u = r5.y * w + r5.w * si.bu + r6.y * si.bv;
        v = r5.z * w + r6.x * si.bu + r6.z * si.bv;
    } else if (p.shape == DFR_SHAPE_SPHERE) {
        u = si.bu; v = si.bv;
    } else if (p.shape == DFR_SHAPE_RECTANGLE) {
        // projection on the plane of the normal's largest component, from
        // corners a, d = a + c - b and b
        const V n = mk(fabsf(p.normal[0]), fabsf(p.normal[1]), fabsf(p.normal[2]));
        const bool ix = n.x > n.y && n.x > n.z, iy = n.y > n.z;
        auto pr = [&](float x, float y, float z, float& a, float& b) {
            if (ix) { a = y; b = z; } else if (iy) { a = z; b = x; } else { a = x; b = y; }
        };
        const float* d = p.data;
        const V A = mk(d[0], d[1], d[2]), B = mk(d[3], d[4], d[5]), C = mk(d[6], d[7], d[8]);
        const V D = A + C - B;
        float ax, ay, dx, dy, bx, by, hu, hv;
        pr(A.x, A.y, A.z, ax, ay); pr(D.x, D.y, D.z, dx, dy); pr(B.x, B.y, B.z, bx, by);
        pr(si.p.x, si.p.y, si.p.z, hu, hv);
        const float abx = dx - ax, aby = dy - ay, acx = bx - ax, acy = by - ay;
        const float det = abx * acy - aby * acx, r = 1.0f / det;
        u = hu * (-aby * r) + hv * (abx * r) + (ax * aby - ay * abx) * r;
        v = hu * (acy * r) + hv * (-acx * r) + (acx * ay - acy * ax) * r;
    } else {
        const V q = si.p - ld3(p.data);
        u = dot(nrm(ld3(p.data + 3)), q);
        v = dot(nrm(ld3(p.data + 6)), q);
    }
}

DEV float frac_part(float x) { return x - floorf(x); }

DEV V texture_image(const DevScene& sc, const DevTexture& T, float u, float v) {
    const float ru = (u * T.cos_angle - v * T.sin_angle) * T.scale_u;
    const float rv = (v * T.cos_angle + u * T.sin_angle) * T.scale_v;
    const float x = pmodf_dev(ru * (float)T.width - 0.5f, (float)T.width);
    const float y = pmodf_dev(rv * (float)T.height - 0.5f, (float)T.height);
    const int x0 = (int)floorf(x), x1 = (int)ceilf(x), y0 = (int)floorf(y), y1 = (int)ceilf(y);
    auto tx = [&](int a, int b) {
        const float* t = sc.texels + T.texel + 3 * ((int64_t)pmodi_dev(b, T.height) * T.width + pmodi_dev(a, T.width));
        return mk(t[0], t[1], t[2]);
    };
    if (x0 == x1 && y0 == y1) return tx(x0, y0);
    const float fx = x - (float)x0, fy = y - (float)y0;
    const V a = tx(x0, y0) * (1.0f - fx) + tx(x1, y0) * fx, b = tx(x0, y1) * (1.0f - fx) + tx(x1, y1) * fx;
    return a * (1.0f - fy) + b * fy;
}

// ---------------------------------------------------------------------------
// Noise of the Marble and simplex fBm textures: the reference's Perlin noise
// (M/noise/PerlinNoiseF.java:113-200: improved noise with its own gradient
// choice, u = x for hashes below 8 and 12, 13, v = y for hashes below 4 and
// 12, 13) and turbulence (:231-250), and its 3D simplex noise (M/noise
// /SimplexNoiseF.java:405-507, clamped to [-1, 1]) and fBm (:246-267).
// ---------------------------------------------------------------------------
#include "noise_tables.inc"
static __device__ const uint8_t NOISE_PERM_DEV[512] = {NOISE_PERM_256, NOISE_PERM_256};
static __device__ const float NOISE_GRAD3_DEV[36] = {NOISE_GRAD3_36};

DEV float perlin_grad(int h, float x, float y, float z) {
    const float u = (h < 8 || h == 12 || h == 13) ? x : y;
    const float v = (h < 4 || h == 12 || h == 13) ? y : z;
    return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
}

DEV float perlin_noise(float x, float y, float z) {
    if (!isfinite(x) || !isfinite(y) || !isfinite(z)) return 0.0f;
    const uint8_t* P = NOISE_PERM_DEV;
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const int x0 = (int)fx & 0xff, y0 = (int)fy & 0xff, z0 = (int)fz & 0xff;
    const float x1 = x - fx, y1 = y - fy, z1 = z - fz;
    const float u = x1 * x1 * x1 * (x1 * (x1 * 6.0f - 15.0f) + 10.0f);
    const float v = y1 * y1 * y1 * (y1 * (y1 * 6.0f - 15.0f) + 10.0f);
    const float w = z1 * z1 * z1 * (z1 * (z1 * 6.0f - 15.0f) + 10.0f);
    const int a0 = P[x0] + y0, a1 = P[a0] + z0, a2 = P[a0 + 1] + z0;
    const int b0 = P[x0 + 1] + y0, b1 = P[b0] + z0, b2 = P[b0 + 1] + z0;
    const float g0 = perlin_grad(P[a1] & 15, x1, y1, z1);
    const float g1 = perlin_grad(P[b1] & 15, x1 - 1.0f, y1, z1);
    const float g2 = perlin_grad(P[a2] & 15, x1, y1 - 1.0f, z1);
    const float g3 = perlin_grad(P[b2] & 15, x1 - 1.0f, y1 - 1.0f, z1);
    const float g4 = perlin_grad(P[a1 + 1] & 15, x1, y1, z1 - 1.0f);
    const float g5 = perlin_grad(P[b1 + 1] & 15, x1 - 1.0f, y1, z1 - 1.0f);
    const float g6 = perlin_grad(P[a2 + 1] & 15, x1, y1 - 1.0f, z1 - 1.0f);
    const float g7 = perlin_grad(P[b2 + 1] & 15, x1 - 1.0f, y1 - 1.0f, z1 - 1.0f);
    const float l0 = g0 + u * (g1 - g0), l1 = g2 + u * (g3 - g2), l2 = g4 + u * (g5 - g4), l3 = g6 + u * (g7 - g6);
    const float l4 = l0 + v * (l1 - l0), l5 = l2 + v * (l3 - l2);
    return l4 + w * (l5 - l4);
}

// |noise| summed over octaves at doubling frequencies, each divided by its
// frequency (the first octave at the point itself)
DEV float perlin_turbulence(float x, float y, float z, int octaves) {
    float noise = fabsf(perlin_noise(x, y, z)), weight = 1.0f;
    for (int i = 1; i < octaves; i++) {
        weight *= 2.0f;
        noise += fabsf(perlin_noise(x * weight, y * weight, z * weight)) / weight;
    }
    return noise;
}

DEV int fast_floor(float v) {
    const int i = (int)v;
    return v < (float)i ? i - 1 : i;
}

DEV float simplex_noise(float x, float y, float z) {
    if (!isfinite(x) || !isfinite(y) || !isfinite(z)) return 0.0f;
    const uint8_t* P = NOISE_PERM_DEV;
    const float* G = NOISE_GRAD3_DEV;
    constexpr float F3 = 1.0f / 3.0f, G3 = 1.0f / 6.0f;
    const float s = (x + y + z) * F3;
    const int i = fast_floor(x + s), j = fast_floor(y + s), k = fast_floor(z + s);
    const float t = (float)(i + j + k) * G3;
    const float x0 = x - ((float)i - t), y0 = y - ((float)j - t), z0 = z - ((float)k - t);
    int i1, j1, k1, i2, j2, k2;
    if (x0 >= y0) {
        if (y0 >= z0) { i1 = 1; j1 = 0; k1 = 0; i2 = 1; j2 = 1; k2 = 0; }
        else if (x0 >= z0) { i1 = 1; j1 = 0; k1 = 0; i2 = 1; j2 = 0; k2 = 1; }
        else { i1 = 0; j1 = 0; k1 = 1; i2 = 1; j2 = 0; k2 = 1; }
    } else {
        if (y0 < z0) { i1 = 0; j1 = 0; k1 = 1; i2 = 0; j2 = 1; k2 = 1; }
        else if (x0 < z0) { i1 = 0; j1 = 1; k1 = 0; i2 = 0; j2 = 1; k2 = 1; }
        else { i1 = 0; j1 = 1; k1 = 0; i2 = 1; j2 = 1; k2 = 0; }
    }
    const float x1 = x0 - (float)i1 + G3, y1 = y0 - (float)j1 + G3, z1 = z0 - (float)k1 + G3;
    const float x2 = x0 - (float)i2 + 2.0f * G3, y2 = y0 - (float)j2 + 2.0f * G3, z2 = z0 - (float)k2 + 2.0f * G3;
    const float x3 = x0 - 1.0f + 3.0f * G3, y3 = y0 - 1.0f + 3.0f * G3, z3 = z0 - 1.0f + 3.0f * G3;
    const int ii = i & 0xff, jj = j & 0xff, kk = k & 0xff;
    const int gi0 = P[ii + P[jj + P[kk]]] % 12;
    const int gi1 = P[ii + i1 + P[jj + j1 + P[kk + k1]]] % 12;
    const int gi2 = P[ii + i2 + P[jj + j2 + P[kk + k2]]] % 12;
    const int gi3 = P[ii + 1 + P[jj + 1 + P[kk + 1]]] % 12;
    auto corner = [&](int gi, float a, float b, float c) {
        const float tt = 0.6f - a * a - b * b - c * c;
        return tt < 0.0f ? 0.0f : (tt * tt) * (tt * tt) * (G[gi * 3] * a + G[gi * 3 + 1] * b + G[gi * 3 + 2] * c);
    };
    const float n = 32.0f * (corner(gi0, x0, y0, z0) + corner(gi1, x1, y1, z1) + corner(gi2, x2, y2, z2) +
                             corner(gi3, x3, y3, z3));
    return fminf(fmaxf(n, -1.0f), 1.0f);
}

// fBm of simplex noise mapped to [0, 1] (minimum 0, maximum 1)
DEV float simplex_fbm(float x, float y, float z, float frequency, float gain, int octaves) {
    float amplitude = 1.0f, total = 0.0f, f = frequency, noise = 0.0f;
    for (int i = 0; i < octaves; i++) {
        noise += simplex_noise(x * f, y * f, z * f) * amplitude;
        total += amplitude;
        amplitude *= gain;
        f *= 2.0f;
    }
    noise /= total;
    return noise * (1.0f - 0.0f) / 2.0f + (1.0f + 0.0f) / 2.0f;
}

// A texture that is not a blend, following its chain of sub-textures.
DEV V texture_chain(const DevScene& sc, int t, const Surf& si, V d, float u, float v) {
    for (int guard = 0; guard < 64; guard++) {
        const DevTexture& T = sc.textures[t];
        switch (T.type) {
            case DFR_TEX_CONSTANT: return ld3(T.color);
            case DFR_TEX_CHECKERBOARD: {
                const bool iu = frac_part((u * T.cos_angle - v * T.sin_angle) * T.scale_u) > 0.5f;
                const bool iv = frac_part((v * T.cos_angle + u * T.sin_angle) * T.scale_v) > 0.5f;
                t = (iu ^ iv) ? T.a : T.b;
                break;
            }
            case DFR_TEX_UV: return mk(u, v, 0.0f);
            case DFR_TEX_SURFACE_NORMAL: {
                const V n = dot(si.s.w, d) > 0.0f ? -si.s.w : si.s.w;
                return mk((n.x + 1.0f) / 2.0f, (n.y + 1.0f) / 2.0f, (n.z + 1.0f) / 2.0f);
            }
            case DFR_TEX_DOT_PRODUCT: {
                const float k = fabsf(dot(si.s.w, d));
                return mk(k, k, k);
            }
            case DFR_TEX_BULLSEYE: {
                const float r = fmodf(len(si.p - ld3(T.origin)) * T.scale_u, 1.0f);
                t = r > 0.5f ? T.a : T.b;
                break;
            }
            case DFR_TEX_POLKA_DOT: {
                const float x = frac_part((u * T.cos_angle - v * T.sin_angle) * T.cell) - 0.5f;
                const float y = frac_part((v * T.cos_angle + u * T.sin_angle) * T.cell) - 0.5f;
                t = x * x + y * y < T.radius * T.radius ? T.a : T.b;
                break;
            }
            case DFR_TEX_IMAGE: return texture_image(sc, T, u, v);
            case DFR_TEX_MARBLE: {
                // MarbleTexture.getColor (MarbleTexture.java:131-146): colours A
                // (color), B (origin), C (t); frequency, scale, octaves in
                // scale_u, scale_v, cell
                const float x = si.p.x * T.scale_u, y = si.p.y * T.scale_u, z = si.p.z * T.scale_u;
                const float r = T.scale_v * perlin_turbulence(x, y, z, (int)T.cell);
                const float sv = 2.0f * fabsf(sinf(x + r));
                const float tv = sv < 1.0f ? sv : sv - 1.0f;
                const V lo = sv < 1.0f ? ld3(T.t) : ld3(T.origin), hi = sv < 1.0f ? ld3(T.origin) : ld3(T.color);
                return lo * (1.0f - tv) + hi * tv;
            }
            case DFR_TEX_SIMPLEX_FBM: {
                // SimplexFractionalBrownianMotionTexture.getColor (:127-137):
                // frequency, gain, octaves in scale_u, scale_v, cell
                return ld3(T.color) * simplex_fbm(si.p.x, si.p.y, si.p.z, T.scale_u, T.scale_v, (int)T.cell);
            }
            default: return mk(0.0f, 0.0f, 0.0f);
        }
    }
    return mk(0.0f, 0.0f, 0.0f);
}

DEV V texture_color(const DevScene& sc, int t, const Surf& si, V d) {
    float u, v;
    hit_uv(sc, si, u, v);
    const DevTexture& T = sc.textures[t];
    if (T.type == DFR_TEX_BLEND) {
        const V a = texture_chain(sc, T.a, si, d, u, v), b = texture_chain(sc, T.b, si, d, u, v);
        return mk((1.0f - T.t[0]) * a.x + T.t[0] * b.x, (1.0f - T.t[1]) * a.y + T.t[1] * b.y,
                  (1.0f - T.t[2]) * a.z + T.t[2] * b.z);
    }
    return texture_chain(sc, t, si, d, u, v);
}

// A material's colours at a hit: the textured slots evaluated and clamped at
// zero, the lobes whose colour is black dropped as the material's
// computeBSDF leaves them out (MatteMaterial, PlasticMaterial,
// MirrorMaterial, GlassMaterial); returns false when the material has no
// BSDF here (the ray passes through).
// The values of a material's textured slots at a hit, every slot through one
// call of texture_color (the texture code is inlined once per kernel, not
// once per slot and material).
struct SlotVals {
    V v[8];
};
DEV SlotVals slot_values(const DevScene& sc, const DevMaterial& m, const Surf& si, V d) {
    SlotVals sv;
    V t0 = mk(0, 0, 0), t1 = t0, t2 = t0, t3 = t0, t4 = t0, t5 = t0, t6 = t0, t7 = t0;
#pragma unroll 1
    for (int k = 0; k < 8; k++) {
        if (!m.tex[k]) continue;
        const V c = texture_color(sc, m.tex[k] - 1, si, d);
        switch (k) {
            case 0: t0 = c; break; case 1: t1 = c; break; case 2: t2 = c; break; case 3: t3 = c; break;
            case 4: t4 = c; break; case 5: t5 = c; break; case 6: t6 = c; break; default: t7 = c; break;
        }
    }
    sv.v[0] = t0; sv.v[1] = t1; sv.v[2] = t2; sv.v[3] = t3; sv.v[4] = t4; sv.v[5] = t5; sv.v[6] = t6; sv.v[7] = t7;
    return sv;
}
// a colour slot: the texture's colour clamped at zero, else `fallback`
DEV V slot_color(const DevMaterial& m, const SlotVals& sv, int slot, V fallback) {
    if (!m.tex[slot]) return fallback;
    const V c = sv.v[slot];
    return mk(fmaxf(c.x, 0.0f), fmaxf(c.y, 0.0f), fmaxf(c.z, 0.0f));
}
// a colour slot read unclamped
DEV V slot_raw(const DevMaterial& m, const SlotVals& sv, int slot, V fallback) {
    return m.tex[slot] ? sv.v[slot] : fallback;
}
// A textured roughness: the texture's colour average (Texture.getFloat),
// remapped as the material asks (MicrofacetDistribution
// .convertRoughnessToAlpha), floored at 0.001 as the lobe's alphas are;
// `fallback` (an alpha) when the slot has no texture.
DEV float slot_alpha(const DevMaterial& m, const SlotVals& sv, int slot, float fallback) {
    if (!m.tex[slot]) return fallback;
    const V c = sv.v[slot];
    float r = (c.x + c.y) + c.z;
    r = r / 3.0f;
    if (m.remap) {
        const float y = (float)log((double)fmaxf(r, 1.0e-3f));
        r = 1.62142f + 0.819955f * y + 0.1734f * y * y + 0.0171201f * y * y * y + 0.000640711f * y * y * y * y;
    }
    return fmaxf(r, 0.001f);
}
template <bool EXT>
DEV bool shade_material(const DevScene& sc, const DevMaterial& m, const Surf& si, V d, TexHit& th) {
    th.set = 0u; th.off = 0u; th.rough = 0u; th.cond = 0u; th.deta = 0u;
    if constexpr (EXT) {
        if (m.tex[0] | m.tex[1] | m.tex[2] | m.tex[3] | m.tex[4] | m.tex[5] | m.tex[6] | m.tex[7]) {
            const DevLobe* L = sc.lobes + m.first_lobe;
            const SlotVals sv = slot_values(sc, m, si, d);
            switch (m.type) {
                case DFR_MAT_MATTE:
                    th.c[0] = slot_color(m, sv, DFR_SLOT_KD, ld3(L[0].R)); th.set = 1u;
                    return true;
                case DFR_MAT_PLASTIC:
                    th.c[0] = slot_color(m, sv, DFR_SLOT_KD, ld3(L[0].R));
                    th.c[1] = slot_color(m, sv, DFR_SLOT_KS, ld3(L[1].R));
                    th.set = 3u;
                    th.off = (black(th.c[0]) ? 1u : 0u) | (black(th.c[1]) ? 2u : 0u);
                    if (m.tex[DFR_SLOT_ROUGHNESS_U]) {  // the glossy lobe's one roughness
                        th.ax = th.ay = slot_alpha(m, sv, DFR_SLOT_ROUGHNESS_U, 0.0f);
                        th.rough = 2u;
                    }
                    return th.off != 3u;
                case DFR_MAT_SUBSTRATE: {  // one FresnelBlend lobe: Rd = kd in R, Rs = ks in T
                    th.c[0] = slot_color(m, sv, DFR_SLOT_KD, ld3(L[0].R));
                    th.c[2] = slot_color(m, sv, DFR_SLOT_KS, ld3(L[0].T));
                    th.set = 1u | 4u;
                    th.ax = slot_alpha(m, sv, DFR_SLOT_ROUGHNESS_U, L[0].alphaX);
                    th.ay = slot_alpha(m, sv, DFR_SLOT_ROUGHNESS_V, L[0].alphaY);
                    th.rough = 1u;
                    return !(black(th.c[0]) && black(th.c[2]));
                }
                case DFR_MAT_GLOSSY: {  // TorranceSparrowBRDF(KR, ConductorFresnel(KR, 1, 1)): eta = k = 1 / KR
                    // GlossyMaterial reads KR unclamped
                    const V kr = slot_raw(m, sv, DFR_SLOT_KR, ld3(L[0].R));
                    th.c[0] = kr; th.set = 1u;
                    th.eta = mk(1.0f / kr.x, 1.0f / kr.y, 1.0f / kr.z); th.k = th.eta;
                    th.ax = th.ay = slot_alpha(m, sv, DFR_SLOT_ROUGHNESS_U, L[0].alphaX);
                    th.cond = 1u; th.rough = 1u;
                    return true;
                }
                case DFR_MAT_CLEAR_COAT:  // the specular layer (ks) and the diffuse one (kd), unclamped
                    th.c[0] = slot_raw(m, sv, DFR_SLOT_KS, ld3(L[0].R));
                    th.c[1] = slot_raw(m, sv, DFR_SLOT_KD, ld3(L[1].R));
                    th.set = 3u;
                    return true;
                case DFR_MAT_METAL:  // one conductor Torrance-Sparrow lobe (MetalMaterial.computeBSDF)
                    th.ax = slot_alpha(m, sv, DFR_SLOT_ROUGHNESS_U, L[0].alphaX);
                    th.ay = slot_alpha(m, sv, DFR_SLOT_ROUGHNESS_V, L[0].alphaY);
                    th.eta = slot_raw(m, sv, DFR_SLOT_ETA, ld3(L[0].condEta));
                    th.k = slot_raw(m, sv, DFR_SLOT_K, ld3(L[0].condK));
                    th.rough = 1u; th.cond = 1u;
                    return true;
                case DFR_MAT_MIRROR:
                    th.c[0] = slot_color(m, sv, DFR_SLOT_KR, ld3(L[0].R)); th.set = 1u;
                    return !black(th.c[0]);
                case DFR_MAT_GLASS:
                    if (m.tex[DFR_SLOT_ETA]) {  // textureEta: the colour average, every lobe
                        const V e = sv.v[DFR_SLOT_ETA];
                        th.geta = (e.x + e.y + e.z) / 3.0f;
                        th.deta = m.n_lobes == 1 ? 1u : 3u;
                    }
                    if (m.n_lobes == 1) {  // smooth: one Fresnel-specular lobe
                        th.c[0] = slot_color(m, sv, DFR_SLOT_KR, ld3(L[0].R));
                        th.c[2] = slot_color(m, sv, DFR_SLOT_KT, ld3(L[0].T));
                        th.set = 1u | 4u;
                        return !(black(th.c[0]) && black(th.c[2]));
                    }
                    th.c[0] = slot_color(m, sv, DFR_SLOT_KR, ld3(L[0].R));
                    th.c[3] = slot_color(m, sv, DFR_SLOT_KT, ld3(L[1].T));
                    th.set = 1u | 8u;
                    th.off = (black(th.c[0]) ? 1u : 0u) | (black(th.c[3]) ? 2u : 0u);
                    return th.off != 3u;
                default: break;
            }
        }
    }
    return m.has_bsdf != 0;
}

// SpotLight and DirectionalLight (delta distributions); reached only from
// kernels instantiated with EXT = true.
DEV bool sample_delta_light_ext(const DevLight& L, const Surf& si, LightSmp& out) {
    if (L.type == DFR_LIGHT_SPOT) {  // SpotLight.sampleRadianceIncoming, doComputeFalloff
        V pos = ld3(L.position);
        V wi = nrm(pos - si.p);
        const float ct = dot(-wi, ld3(L.dir));
        float fall = 1.0f;
        if (ct < L.area) fall = 0.0f;
        else if (ct < L.param) { const float t = (ct - L.area) / (L.param - L.area); fall = (t * t) * (t * t); }
        V dd = si.p - pos;
        out.L = ld3(L.radiance) * fall / dot(dd, dd);
        out.p = pos;
        out.wi = wi;
        out.pdf = 1.0f;
        return true;
    }
    {  // DirectionalLight.sampleRadianceIncoming
        V d = ld3(L.dir);
        out.L = ld3(L.radiance);
        out.p = si.p + d * (2.0f * L.param);
        out.wi = d;
        out.pdf = 1.0f;
        return true;
    }
}

// DiffuseAreaLight over a Sphere3F (M/geometry/shape/Sphere3F.java:156-226):
// the reference samples its unit sphere in object space, where a world point
// x of a sphere with centre c and radius r is (x - c) / r.  Inside the sphere
// a uniform point on it, pdf d^2 / |cos| in object units (light_pdf_mode 1:
// d^2 / (|cos| 4 pi r^2) in world units); outside a direction of the cone the
// sphere subtends, pdf 1 / (2 pi (1 - cos theta_max)).
DEV bool sample_sphere_light(const DevScene& sc, const DevLight& L, const Surf& si, float u, float v, LightSmp& out) {
    const DevPrim& sp = sc.prims[L.prim];
    const V c = ld3(sp.data);
    const float r = sp.data[3];
    const V po = (si.p - c) * (1.0f / r);
    V pt, n;
    float pdf;
    if (dot(po, po) <= 1.0f) {
        const float ct = 1.0f - 2.0f * u, st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct)), phi = v * 2.0f * PI_F;
        n = nrm(mk(st * cosf(phi), st * sinf(phi), ct));
        pt = n;
        const V inc = pt - po;
        if (dot(inc, inc) == 0.0f) return false;
        pdf = dot(inc, inc) / fabsf(dot(n, -nrm(inc)));
        if (sc.light_pdf_mode == 1) pdf = pdf / (4.0f * PI_F);
        if (isinf(pdf) || isnan(pdf)) return false;
    } else {
        const float dist = len(po), dr = 1.0f / dist;
        const ONB b = onb_w(-po * dr);
        const float s2 = dr * dr, sr = 1.0f / dr;
        const float cmax = sqrtf(fmaxf(0.0f, 1.0f - s2));
        const bool small = s2 < 0.00068523f;
        const float ctheta = small ? sqrtf(1.0f - s2 * u) : (cmax - 1.0f) * u + 1.0f;
        const float st2 = small ? s2 * u : 1.0f - ctheta * ctheta;
        const float ca = st2 * sr + ctheta * sqrtf(fmaxf(0.0f, 1.0f - st2 * sr * sr));
        const float sa = sqrtf(fmaxf(0.0f, 1.0f - ca * ca));
        const float phi = v * 2.0f * PI_F;
        const V dir = (-b.u) * (sa * cosf(phi)) + (-b.v) * (sa * sinf(phi)) + (-b.w) * ca;
        pt = dir;
        n = nrm(dir);
        pdf = 1.0f / (2.0f * PI_F * (1.0f - cmax));
    }
    const V pw = c + pt * r;
    const V wi = nrm(pw - si.p);
    if (pdf > 0.0f && (L.two_sided || dot(n, -wi) > 0.0f)) {
        out.L = ld3(L.radiance); out.p = pw; out.wi = wi; out.pdf = pdf;
        return true;
    }
    return false;
}

// Sphere3F.evaluateProbabilityDensityFunction (Sphere3F.java:460-479):
// d^2 / |cos| * 4 pi in object units (the reference multiplies by the area);
// light_pdf_mode 1: the cone pdf outside the sphere, d^2 / (|cos| 4 pi r^2)
// inside.
DEV float sphere_light_pdf(const DevScene& sc, const DevLight& L, const Surf& si, V wi) {
    const DevPrim& sp = sc.prims[L.prim];
    const V c = ld3(sp.data);
    const float r = sp.data[3];
    const V d = nrm(wi);
    const V po = (si.p - c) * (1.0f / r);
    if (sc.light_pdf_mode == 1 && dot(po, po) > 1.0f) {
        const float dist = len(po);
        const float cmax = sqrtf(fmaxf(0.0f, 1.0f - 1.0f / (dist * dist)));
        if (dot(d, -po * (1.0f / dist)) < cmax) return 0.0f;
        return 1.0f / (2.0f * PI_F * (1.0f - cmax));
    }
    const float t = sphere_t(sp, si.p, d, T_MIN, FLT_MAX);
    if (t == 0.0f) return 0.0f;
    const V n = nrm((si.p + d * t) - c);
    const float ca = fabsf(dot(n, -d));
    const float to = t / r;
    float pdf = to * to / ca * (4.0f * PI_F);
    if (sc.light_pdf_mode == 1) pdf = (t * t) / (ca * (4.0f * PI_F * r * r));
    if (isinf(pdf) || isnan(pdf)) return 0.0f;
    return pdf;
}

// DiffuseAreaLight over Triangle3F, PointLight, PerezLight sampling (and the
// spot and directional lights with EXT).
template <bool EXT>
DEV bool sample_light(const DevScene& sc, const DevLight& L, const Surf& si, float u, float v, float usun,
                      LightSmp& out) {
    if (L.type == DFR_LIGHT_POINT) {
        V pos = ld3(L.position);
        V dd = si.p - pos;
        out.L = ld3(L.radiance) / dot(dd, dd);
        out.p = pos;
        out.wi = nrm(pos - si.p);
        out.pdf = 1.0f;
        return true;
    }
    if constexpr (EXT) {
        if (L.type == DFR_LIGHT_SPOT || L.type == DFR_LIGHT_DIRECTIONAL) return sample_delta_light_ext(L, si, out);
        if (L.type == DFR_LIGHT_IMAGE) return image_sample(sc, sc.images[L.perez], si, u, v, out);
    }
    if (L.type == DFR_LIGHT_PEREZ) return perez_sample(sc.perez[L.perez], si, u, v, usun, out);
    if (L.tri < 0) return sample_sphere_light(sc, L, si, u, v, out);
    const size_t rb = 8 * (size_t)L.tri;
    const float4 r0 = ld4g(sc.tshade, rb), r1 = ld4g(sc.tshade, rb + 1), r2 = ld4g(sc.tshade, rb + 2);
    const float4 r3 = ld4g(sc.tshade, rb + 3), r4 = ld4g(sc.tshade, rb + 4);
    V a = f4v(r0), b = f4v(r1), c = f4v(r2);
    float su = sqrtf(u);
    float b0 = 1.0f - su, b1 = v * su, b2 = 1.0f - b0 - b1;
    V p = a * b0 + b * b1 + c * b2;
    V n = nrm(mk(r0.w, r1.w, r2.w) * b0 + f4v(r3) * b1 + f4v(r4) * b2);
    V inc = p - si.p;
    if (dot(inc, inc) == 0.0f) return false;
    V incn = nrm(inc);
    float pdf = dot(inc, inc) / fabsf(dot(n, -incn));
    if (sc.light_pdf_mode == 1) pdf *= 1.0f / L.area;
    if (isinf(pdf) || isnan(pdf)) return false;
    V wi = nrm(p - si.p);
    if (pdf > 0.0f && (L.two_sided || dot(n, -wi) > 0.0f)) {
        out.L = ld3(L.radiance); out.p = p; out.wi = wi; out.pdf = pdf;
        return true;
    }
    return false;
}

// Shape3F.evaluateProbabilityDensityFunction for a triangle or sphere light;
// the infinite lights' solid-angle pdfs (image lights in EXT kernels only)
template <bool EXT>
DEV float light_pdf(const DevScene& sc, const DevLight& L, const Surf& si, V wi) {
    if (L.type == DFR_LIGHT_PEREZ) return perez_pdf(sc.perez[L.perez], wi);
    if constexpr (EXT) {
        if (L.type == DFR_LIGHT_IMAGE) return image_pdf(sc.images[L.perez], wi);
    }
    if (L.type != DFR_LIGHT_DIFFUSE_AREA) return 0.0f;
    if (L.tri < 0) return sphere_light_pdf(sc, L, si, wi);
    const size_t rb = 8 * (size_t)L.tri;
    V a = f4v(ld4g(sc.tshade, rb)), b = f4v(ld4g(sc.tshade, rb + 1)), c = f4v(ld4g(sc.tshade, rb + 2));
    V d = nrm(wi);
    PRay r = prep_ray(si.p, d);
    float t = tri_t(a, b, c, r.o, r.d, r.kx, r.ky, r.kz, r.sx, r.sy, r.sz, T_MIN, FLT_MAX);
    if (t == 0.0f) return 0.0f;
    HitRec h; h.t = t; h.prim = L.prim; h.tri = L.tri;
    Surf s2 = make_surface(sc, si.p, d, h);
    V dp = s2.p - si.p;
    float pdf = dot(dp, dp) / (fabsf(dot(s2.s.w, -wi)) * L.area);
    if (isinf(pdf) || isnan(pdf)) return 0.0f;
    return pdf;
}

// Intersection.evaluateRadianceEmitted (M/scene/Intersection.java:89-101);
// a textured emission (Material.emittance reads textureEmission) only in
// the EXT instantiations, where texture code lives
template <bool EXT = false>
DEV V emitted(const DevScene& sc, const Surf& si, V dir) {
    const DevPrim& p = sc.prims[si.prim];
    if (p.area_light >= 0) {
        const DevLight& L = sc.lights[p.area_light];
        return (L.two_sided || dot(si.s.w, dir) > 0.0f) ? ld3(L.radiance) : mk(0, 0, 0);
    }
    const DevMaterial& m = sc.mats[p.material];
    if constexpr (EXT) {
        if (m.tex_emission) return texture_color(sc, m.tex_emission - 1, si, -dir);
    }
    return ld3(m.emission);
}

// Light.evaluateRadianceEmitted(ray): what a ray that leaves the scene sees
// (image lights in EXT instantiations only)
template <bool EXT>
DEV V emitted_ray(const DevScene& sc, const DevLight& L, V d) {
    if (L.type == DFR_LIGHT_PEREZ) return perez_sky(sc.perez[L.perez], d);
    if constexpr (EXT) {
        if (L.type == DFR_LIGHT_IMAGE) return image_radiance(sc, sc.images[L.perez], d);
    }
    return mk(0, 0, 0);
}

// lights sampled by a delta distribution: no MIS, no BSDF-sampled request
template <bool EXT>
DEV bool delta_light(int type) {
    return type == DFR_LIGHT_POINT || (EXT && (type == DFR_LIGHT_SPOT || type == DFR_LIGHT_DIRECTIONAL));
}

DEV float power_heuristic(float a, float b) { return a * a * __builtin_amdgcn_rcpf(a * a + b * b); }

}  // namespace dfr
