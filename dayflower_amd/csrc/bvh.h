// bvh.h — host-side BVH construction for the scene compiler.
//
// The reference flattens a per-mesh BVH built by a swept SAH
// (M/geometry/boundingvolume/hierarchy/BVHNode3Fs.java:63-...) into a
// stackless threaded int[] (M/scene/compiler/CompiledShape3FCache.java
// :2611-2734).  Here one BVH spans every triangle of the scene; it is built
// by binned SAH over triangle centroids in parallel and emitted in a
// two-wide "child pair" layout: each node holds the boxes of both children,
// so one 64-byte fetch decides which children a ray enters.
#pragma once
#include <cstdint>
#include <vector>

namespace dfr {

// Leaf reference: bit 31 set, bits 3..30 first triangle (BVH order), bits
// 0..2 triangle count - 1.  At most 8 triangles per leaf and 2^28 triangles.
constexpr uint32_t LEAF_FLAG = 0x80000000u;
constexpr int MAX_LEAF = 8;
constexpr uint32_t EMPTY_REF = 0xffffffffu;  // unused child of a one-child root

inline uint32_t make_leaf(uint32_t first, uint32_t count) { return LEAF_FLAG | (first << 3) | (count - 1); }

struct WideNode {
    // [0] c0.lo.x c0.hi.x c0.lo.y c0.hi.y
    // [1] c1.lo.x c1.hi.x c1.lo.y c1.hi.y
    // [2] c0.lo.z c0.hi.z c1.lo.z c1.hi.z
    // [3] ref0 ref1 (uint32 bits) pad pad
    float f[16];
};
static_assert(sizeof(WideNode) == 64, "node is one 64-byte record");

// Four-wide node (128 bytes, one L2 line): child boxes as structure of
// arrays, then the child references.  An unused child slot has EMPTY_REF.
struct Node4 {
    // f[0..3] lo.x, f[4..7] hi.x, f[8..11] lo.y, f[12..15] hi.y,
    // f[16..19] lo.z, f[20..23] hi.z, f[24..27] refs (uint32 bits), f[28..31] 0
    float f[32];
};
static_assert(sizeof(Node4) == 128, "node is one 128-byte record");

struct BvhResult {
    std::vector<WideNode> nodes;        // two-wide layout; nodes[0] is the root
    std::vector<Node4> nodes4;          // four-wide layout (when requested); nodes4[0] is the root
    std::vector<uint32_t> tri_order;    // BVH position -> original triangle
    int max_depth = 0;                  // of the binary tree
    int max_stack4 = 0;                 // deepest traversal stack of the four-wide tree
    int64_t leaf_count = 0;
    double sah_cost = 0.0;
};

// positions: 3 floats per vertex; indices: 3 per triangle.  width 2 emits
// `nodes`, width 4 collapses the binary tree into `nodes4`.
BvhResult build_bvh(const float* positions, const int32_t* indices, int64_t n_triangles, int threads, int width = 2);

}  // namespace dfr
