// bvh_probe.cpp — test helper: builds the scene compiler's BVH on the host and
// hands its arrays to the tests (compiled with g++ next to the tests; not
// part of the product library).
#include <cstring>

#include "../../dayflower_amd/csrc/bvh.h"

extern "C" {

// Builds the tree and copies it out.  Returns the node count, or -1 when
// `cap_nodes` is too small.  nodes: 32 floats (width 4) or 16 floats (width 2)
// per node; order: BVH position -> original triangle; info: max_depth,
// max_stack4, leaf_count.
long long probe_build_bvh(const float* positions, const int* indices, long long n, int width, float* nodes,
                          long long cap_nodes, unsigned* order, int* info) {
    dfr::BvhResult r = dfr::build_bvh(positions, indices, n, 4, width);
    const long long count = width == 4 ? (long long)r.nodes4.size() : (long long)r.nodes.size();
    if (count > cap_nodes) return -1;
    if (width == 4) std::memcpy(nodes, r.nodes4.data(), r.nodes4.size() * sizeof(dfr::Node4));
    else std::memcpy(nodes, r.nodes.data(), r.nodes.size() * sizeof(dfr::WideNode));
    std::memcpy(order, r.tri_order.data(), r.tri_order.size() * sizeof(unsigned));
    info[0] = r.max_depth;
    info[1] = r.max_stack4;
    info[2] = (int)r.leaf_count;
    return count;
}

}  // extern "C"
