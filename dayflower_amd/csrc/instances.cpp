// instances.cpp — shared meshes as instances.
//
// Several mesh primitives that name the same triangle range are instances of
// one mesh, each under its own object-to-world matrix.  The range's
// triangles stay in object space, once, under a four-wide BVH of their own;
// a top-level four-wide BVH over the instances' world boxes leads a ray to
// an instance, where the ray is carried into the mesh's object space and
// walks the mesh's tree.  Moving an instance refits the top-level tree only.
// The reference stores a mesh once per distinct TriangleMesh3F in its
// compiled arrays whatever number of primitives name it (Shape3FCache.java
// :176-177, 298-300), carries
// the ray into object space per primitive (AbstractSceneKernel.java
// :1669-1772), and bounds the primitives in world space for its CPU tree
// (M/scene/Scene.java:1333-1357).
#include "api_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>

namespace dfr {

int refit_nodes(std::vector<Node4>& nodes, const std::function<void(uint32_t, uint32_t, float*)>& leaf, std::string& err);

int find_instances(const SceneHost& in, InstanceSet& is, std::string& err) {
    const int64_t nt = (int64_t)in.indices.size() / 3;
    is = InstanceSet{};
    is.prim_group.assign(in.prims.size(), -1);
    is.tri_group.assign(nt, -1);
    std::vector<int32_t> mesh;
    for (size_t p = 0; p < in.prims.size(); p++)
        if (in.prims[p].shape == DFR_SHAPE_TRIANGLE_MESH && in.prims[p].triangle_count > 0) mesh.push_back((int32_t)p);
    auto key = [&](int32_t p) { return std::make_pair(in.prims[p].first_triangle, in.prims[p].triangle_count); };
    std::stable_sort(mesh.begin(), mesh.end(), [&](int32_t a, int32_t b) { return key(a) < key(b); });
    int64_t end = 0;  // the farthest end of the distinct ranges so far
    for (size_t i = 0; i < mesh.size(); i++) {
        const dfr_primitive& pr = in.prims[mesh[i]];
        if (i > 0 && key(mesh[i]) == key(mesh[i - 1])) {
            int32_t g = is.prim_group[mesh[i - 1]];
            if (g < 0) {
                g = (int32_t)is.groups.size();
                InstanceSet::Group G;
                G.first = pr.first_triangle;
                G.count = pr.triangle_count;
                is.groups.push_back(G);
                is.prim_group[mesh[i - 1]] = g;
            }
            is.prim_group[mesh[i]] = g;
            continue;
        }
        if (pr.first_triangle < end) {
            err = "mesh primitives' triangle ranges overlap without being equal";
            return DFR_E_ARGUMENT;
        }
        end = pr.first_triangle + pr.triangle_count;
    }
    for (size_t g = 0; g < is.groups.size(); g++)
        for (int64_t t = is.groups[g].first; t < is.groups[g].first + is.groups[g].count; t++)
            is.tri_group[t] = (int32_t)g;
    // area lights sample world-space triangle records
    for (size_t p = 0; p < in.prims.size(); p++)
        if (is.prim_group[p] >= 0 && in.prims[p].area_light >= 0) {
            err = "an area light on a mesh shared by several primitives";
            return DFR_E_UNSUPPORTED;
        }
    for (const dfr_light& L : in.lights) {
        if (L.type != DFR_LIGHT_DIFFUSE_AREA) continue;
        const bool on_shared = (L.primitive >= 0 && L.primitive < (int32_t)in.prims.size() && is.prim_group[L.primitive] >= 0) ||
                               (L.triangle >= 0 && L.triangle < nt && is.tri_group[L.triangle] >= 0);
        if (on_shared) { err = "an area light on a mesh shared by several primitives"; return DFR_E_UNSUPPORTED; }
    }
    return DFR_OK;
}

namespace {

uint32_t ref_of(const Node4& n, int c) {
    uint32_t r;
    std::memcpy(&r, &n.f[24 + c], 4);
    return r;
}
void set_ref(Node4& n, int c, uint32_t r) { std::memcpy(&n.f[24 + c], &r, 4); }

float4 f4(float x, float y, float z, float w) {
    float4 v;
    v.x = x; v.y = y; v.z = z; v.w = w;
    return v;
}

// World box of instance primitive p: the eight corners of its mesh's
// object-space box under the primitive's matrix, padded (a ray reaches the
// mesh through its own rounded world-to-object map, so the box must not
// cull a hit that the object-space walk reports).
void instance_box(const std::vector<Xform>& xf, int32_t p, const float* ob, float* wb) {
    const Xform* x = xf.empty() || xf[p].identity ? nullptr : &xf[p];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 8; k++) {
        const double c[3] = {ob[(k & 1) ? 3 : 0], ob[(k & 2) ? 4 : 1], ob[(k & 4) ? 5 : 2]};
        double w[3];
        for (int r = 0; r < 3; r++)
            w[r] = x ? x->a[r][0] * c[0] + x->a[r][1] * c[1] + x->a[r][2] * c[2] + x->t[r] : c[r];
        for (int r = 0; r < 3; r++) { lo[r] = std::min(lo[r], w[r]); hi[r] = std::max(hi[r], w[r]); }
    }
    double mag = 0.0;
    for (int r = 0; r < 3; r++) mag = std::max(mag, std::max(std::fabs(lo[r]), std::fabs(hi[r])));
    const double pad = 1e-5 * mag + 1e-30;
    for (int r = 0; r < 3; r++) {
        wb[r] = std::nextafter((float)(lo[r] - pad), -INFINITY);
        wb[3 + r] = std::nextafter((float)(hi[r] + pad), INFINITY);
    }
}

}  // namespace

void build_instances(const std::vector<float>& opos, const std::vector<Xform>& xf, InstanceSet& is) {
    is.bottom.clear();
    is.tris.clear();
    is.top.clear();
    is.order.clear();
    const int threads = (int)std::max(1u, std::thread::hardware_concurrency());
    int bottom_stack = 0;
    for (InstanceSet::Group& G : is.groups) {
        std::vector<int32_t> idx(3 * G.count);
        for (int64_t k = 0; k < G.count; k++)
            for (int j = 0; j < 3; j++) idx[3 * k + j] = (int32_t)(3 * (G.first + k) + j);
        BvhResult b = build_bvh(opos.data(), idx.data(), G.count, threads, 4);
        const uint32_t node_base = (uint32_t)is.bottom.size(), tri_base = (uint32_t)(is.tris.size() / 3);
        G.root = node_base;
        for (Node4 n : b.nodes4) {
            for (int c = 0; c < 4; c++) {
                const uint32_t r = ref_of(n, c);
                if (r == EMPTY_REF) continue;
                if (r & LEAF_FLAG) set_ref(n, c, make_leaf(((r & 0x7fffffffu) >> 3) + tri_base, (r & 7u) + 1u));
                else set_ref(n, c, r + node_base);
            }
            is.bottom.push_back(n);
        }
        // the mesh's bounds: the union of its root's children
        float bx[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        const Node4& root = b.nodes4[0];
        for (int c = 0; c < 4; c++) {
            if (ref_of(root, c) == EMPTY_REF) continue;
            for (int a = 0; a < 3; a++) {
                bx[a] = std::min(bx[a], root.f[8 * a + c]);
                bx[3 + a] = std::max(bx[3 + a], root.f[8 * a + 4 + c]);
            }
        }
        std::memcpy(G.box, bx, sizeof(bx));
        for (int64_t i = 0; i < G.count; i++) {
            const int64_t t = G.first + b.tri_order[i];
            const float* P = &opos[9 * t];
            float tid;
            const int32_t ti = (int32_t)t;
            std::memcpy(&tid, &ti, 4);
            is.tris.push_back(f4(P[0], P[1], P[2], tid));
            is.tris.push_back(f4(P[3], P[4], P[5], 0.0f));
            is.tris.push_back(f4(P[6], P[7], P[8], 0.0f));
        }
        bottom_stack = std::max(bottom_stack, b.max_stack4);
    }
    // the top-level tree over the instances' world boxes, each box given to
    // the builder as a degenerate triangle (lo, hi, lo)
    std::vector<int32_t> prims;
    for (size_t p = 0; p < is.prim_group.size(); p++)
        if (is.prim_group[p] >= 0) prims.push_back((int32_t)p);
    const int64_t ni = (int64_t)prims.size();
    std::vector<float> ipos(9 * ni);
    std::vector<int32_t> iidx(3 * ni);
    for (int64_t k = 0; k < ni; k++) {
        float B[6];
        instance_box(xf, prims[k], is.groups[is.prim_group[prims[k]]].box, B);
        const float v[9] = {B[0], B[1], B[2], B[3], B[4], B[5], B[0], B[1], B[2]};
        std::memcpy(&ipos[9 * k], v, sizeof(v));
        for (int j = 0; j < 3; j++) iidx[3 * k + j] = (int32_t)(3 * k + j);
    }
    BvhResult tb = build_bvh(ipos.data(), iidx.data(), ni, 1, 4);
    is.top = std::move(tb.nodes4);
    is.order.resize(ni);
    for (int64_t k = 0; k < ni; k++) is.order[k] = prims[tb.tri_order[k]];
    // a top-level walk holds its pending nodes while a mesh's walk runs on top
    is.max_stack = tb.max_stack4 + bottom_stack + 1;
}

void refit_instances(const std::vector<Xform>& xf, InstanceSet& is) {
    const size_t ni = is.order.size();
    std::vector<float> boxes(6 * ni);
    for (size_t k = 0; k < ni; k++)
        instance_box(xf, is.order[k], is.groups[is.prim_group[is.order[k]]].box, &boxes[6 * k]);
    std::string err;
    refit_nodes(is.top, [&](uint32_t first, uint32_t count, float* cb) {
        for (uint32_t i = first; i < first + count; i++)
            for (int a = 0; a < 3; a++) {
                cb[a] = std::min(cb[a], boxes[6 * i + a]);
                cb[3 + a] = std::max(cb[3 + a], boxes[6 * i + 3 + a]);
            }
    }, err);
}

std::vector<DevInstance> instance_records(const std::vector<Xform>& xf, const InstanceSet& is) {
    std::vector<DevInstance> out(is.order.size());
    const uint32_t top = (uint32_t)is.top.size();
    for (size_t k = 0; k < is.order.size(); k++) {
        const int32_t p = is.order[k];
        DevInstance& d = out[k];
        std::memset(&d, 0, sizeof(d));
        // the inverse of x -> A x + t is y -> A^-1 (y - t); A^-1 is the
        // transpose of the normal matrix N = A^-T
        double inv[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, t[3] = {0, 0, 0};
        if (!xf.empty() && !xf[p].identity) {
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) inv[r][c] = xf[p].n[c][r];
            for (int r = 0; r < 3; r++) t[r] = xf[p].t[r];
        }
        for (int r = 0; r < 3; r++) {
            d.w2o[4 * r + 0] = (float)inv[r][0];
            d.w2o[4 * r + 1] = (float)inv[r][1];
            d.w2o[4 * r + 2] = (float)inv[r][2];
            d.w2o[4 * r + 3] = (float)-(inv[r][0] * t[0] + inv[r][1] * t[1] + inv[r][2] * t[2]);
        }
        float box[6];
        instance_box(xf, p, is.groups[is.prim_group[p]].box, box);
        std::memcpy(d.lo, box, 12);
        std::memcpy(d.hi, box + 3, 12);
        d.root = (int32_t)(top + is.groups[is.prim_group[p]].root);
        d.prim = p;
    }
    return out;
}

std::vector<Node4> instance_nodes(const InstanceSet& is) {
    std::vector<Node4> all = is.top;
    const uint32_t top = (uint32_t)is.top.size();
    all.reserve(is.top.size() + is.bottom.size());
    for (Node4 n : is.bottom) {
        for (int c = 0; c < 4; c++) {
            const uint32_t r = ref_of(n, c);
            if (r != EMPTY_REF && !(r & LEAF_FLAG)) set_ref(n, c, r + top);
        }
        all.push_back(n);
    }
    return all;
}

}  // namespace dfr
