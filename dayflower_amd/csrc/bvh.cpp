// bvh.cpp — parallel binned-SAH BVH builder (host).  See bvh.h.
#include "bvh.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <future>
#include <stdexcept>

namespace dfr {
namespace {

struct Box {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const Box& b) {
        for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); }
    }
    void grow(const float* p) {
        for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
    }
    float area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct BinNode {
    Box box;
    int64_t first = 0, count = 0;
    int64_t left = -1;  // index of the left child; right child = left + 1
};

constexpr int BINS = 32;
constexpr float COST_TRAVERSE = 1.0f;
constexpr float COST_INTERSECT = 1.0f;
constexpr int64_t PARALLEL_THRESHOLD = 1 << 16;
constexpr int MAX_DEPTH = 60;

struct Builder {
    const float* pos;
    const int32_t* idx;
    std::vector<Box> tbox;
    std::vector<float> cen;       // 3 per triangle
    std::vector<uint32_t> order;
    std::vector<BinNode> nodes;
    std::atomic<int64_t> next_node{0};
    std::atomic<int> depth_max{0};
    std::atomic<int64_t> leaves{0};
    int threads;

    int64_t alloc_pair() {
        int64_t i = next_node.fetch_add(2);
        if (i + 2 > (int64_t)nodes.size()) throw std::runtime_error("bvh node overflow");
        return i;
    }

    void make_leaf(BinNode& n, int depth) {
        n.left = -1;
        leaves.fetch_add(1);
        int d = depth_max.load();
        while (depth > d && !depth_max.compare_exchange_weak(d, depth)) {}
    }

    // Partition [first, first+count) at `split` along `axis` by bin of the
    // centroid; returns the number on the left.
    int64_t partition_bins(int64_t first, int64_t count, int axis, float cmin, float scale, int split) {
        uint32_t* b = order.data() + first;
        uint32_t* e = b + count;
        uint32_t* m = std::partition(b, e, [&](uint32_t t) {
            int bin = std::min(BINS - 1, (int)((cen[3 * t + axis] - cmin) * scale));
            return bin < split;
        });
        return m - b;
    }

    void build(int64_t ni, int depth) {
        BinNode& n = nodes[ni];
        const int64_t first = n.first, count = n.count;
        // node and centroid bounds
        Box cb;
        Box nb;
        for (int64_t i = first; i < first + count; i++) {
            uint32_t t = order[i];
            nb.grow(tbox[t]);
            cb.grow(&cen[3 * t]);
        }
        n.box = nb;
        if (count <= 2) { make_leaf(n, depth); return; }
        // binned SAH over the three axes
        float best_cost = FLT_MAX;
        int best_axis = -1, best_split = -1;
        for (int axis = 0; axis < 3; axis++) {
            float ext = cb.hi[axis] - cb.lo[axis];
            if (!(ext > 0.0f)) continue;
            float scale = BINS / ext * (1.0f - 1e-6f);
            Box bb[BINS];
            int64_t bc[BINS] = {0};
            for (int64_t i = first; i < first + count; i++) {
                uint32_t t = order[i];
                int bin = std::min(BINS - 1, (int)((cen[3 * t + axis] - cb.lo[axis]) * scale));
                bb[bin].grow(tbox[t]);
                bc[bin]++;
            }
            float rarea[BINS];
            int64_t rcount[BINS];
            Box acc;
            int64_t c = 0;
            for (int i = BINS - 1; i > 0; i--) {
                acc.grow(bb[i]); c += bc[i];
                rarea[i] = acc.area(); rcount[i] = c;
            }
            acc = Box(); c = 0;
            for (int i = 0; i < BINS - 1; i++) {
                acc.grow(bb[i]); c += bc[i];
                if (c == 0 || rcount[i + 1] == 0) continue;
                float cost = acc.area() * (float)c + rarea[i + 1] * (float)rcount[i + 1];
                if (cost < best_cost) { best_cost = cost; best_axis = axis; best_split = i + 1; }
            }
        }
        float area = nb.area();
        float leaf_cost = COST_INTERSECT * (float)count;
        float split_cost = area > 0.0f ? COST_TRAVERSE + COST_INTERSECT * best_cost / area : FLT_MAX;
        if (best_axis >= 0 && count <= MAX_LEAF && leaf_cost <= split_cost) { make_leaf(n, depth); return; }
        int64_t nl;
        if (best_axis < 0 || depth >= MAX_DEPTH) {
            // degenerate centroids or too deep: median split on the widest axis
            int axis = 0;
            for (int k = 1; k < 3; k++)
                if (cb.hi[k] - cb.lo[k] > cb.hi[axis] - cb.lo[axis]) axis = k;
            nl = count / 2;
            std::nth_element(order.begin() + first, order.begin() + first + nl, order.begin() + first + count,
                             [&](uint32_t a, uint32_t b) { return cen[3 * a + axis] < cen[3 * b + axis]; });
            if (count <= MAX_LEAF && depth >= MAX_DEPTH) { make_leaf(n, depth); return; }
        } else {
            float scale = BINS / (cb.hi[best_axis] - cb.lo[best_axis]) * (1.0f - 1e-6f);
            nl = partition_bins(first, count, best_axis, cb.lo[best_axis], scale, best_split);
            if (nl == 0 || nl == count) nl = count / 2;
        }
        int64_t l = alloc_pair();
        n.left = l;
        nodes[l].first = first; nodes[l].count = nl;
        nodes[l + 1].first = first + nl; nodes[l + 1].count = count - nl;
        if (count >= PARALLEL_THRESHOLD && threads > 1) {
            auto f = std::async(std::launch::async, [this, l, depth] { build(l, depth + 1); });
            build(l + 1, depth + 1);
            f.get();
        } else {
            build(l, depth + 1);
            build(l + 1, depth + 1);
        }
    }
};

}  // namespace

// Collapse of the binary tree into four-wide nodes: each four-wide node takes
// a binary node's two children and keeps opening the largest-area internal
// child until it has four children or only leaves.  Breadth-first numbering.
static void emit_four_wide(const std::vector<BinNode>& bn, BvhResult& r) {
    auto leaf_ref = [&](const BinNode& x) { return make_leaf((uint32_t)x.first, (uint32_t)x.count); };
    std::vector<std::vector<int64_t>> kids;  // binary children of each four-wide node
    std::vector<int64_t> queue;              // binary node of each four-wide node
    auto open = [&](int64_t root) {
        std::vector<int64_t> c;
        if (bn[root].left < 0) { c.push_back(root); return c; }
        c.push_back(bn[root].left);
        c.push_back(bn[root].left + 1);
        while (c.size() < 4) {
            int best = -1;
            float best_area = -1.0f;
            for (int k = 0; k < (int)c.size(); k++)
                if (bn[c[k]].left >= 0 && bn[c[k]].box.area() > best_area) { best = k; best_area = bn[c[k]].box.area(); }
            if (best < 0) break;
            int64_t x = c[best];
            c[best] = bn[x].left;
            c.push_back(bn[x].left + 1);
        }
        return c;
    };
    queue.push_back(0);
    for (size_t q = 0; q < queue.size(); q++) {
        kids.push_back(open(queue[q]));
        for (int64_t ch : kids.back())
            if (bn[ch].left >= 0) queue.push_back(ch);
    }
    const int64_t n4 = (int64_t)queue.size();
    r.nodes4.resize(n4);
    // four-wide index of every queued binary node
    std::vector<int64_t> index_of(bn.size(), -1);
    for (int64_t q = 0; q < n4; q++) index_of[queue[q]] = q;
    std::vector<int> stack_need(n4, 0);
    for (int64_t q = n4 - 1; q >= 0; q--) {
        Node4& w = r.nodes4[q];
        std::memset(&w, 0, sizeof(w));
        const std::vector<int64_t>& c = kids[q];
        int deepest = 0;
        for (int k = 0; k < 4; k++) {
            uint32_t ref = EMPTY_REF;
            Box box;  // inverted: never entered (the reference is EMPTY_REF as well)
            if (k < (int)c.size()) {
                const BinNode& ch = bn[c[k]];
                box = ch.box;
                ref = ch.left >= 0 ? (uint32_t)index_of[c[k]] : leaf_ref(ch);
                if (ch.left >= 0) deepest = std::max(deepest, stack_need[index_of[c[k]]]);
            }
            w.f[0 + k] = box.lo[0]; w.f[4 + k] = box.hi[0];
            w.f[8 + k] = box.lo[1]; w.f[12 + k] = box.hi[1];
            w.f[16 + k] = box.lo[2]; w.f[20 + k] = box.hi[2];
            std::memcpy(&w.f[24 + k], &ref, 4);
        }
        // entering this node pushes all but one child
        stack_need[q] = (int)c.size() - 1 + deepest;
    }
    r.max_stack4 = stack_need.empty() ? 0 : stack_need[0];
}

BvhResult build_bvh(const float* positions, const int32_t* indices, int64_t n, int threads, int width) {
    BvhResult r;
    if (n <= 0) return r;
    if (n >= (int64_t(1) << 28)) throw std::runtime_error("too many triangles for the leaf encoding");
    Builder b;
    b.pos = positions; b.idx = indices; b.threads = threads;
    b.tbox.resize(n);
    b.cen.resize(3 * n);
    b.order.resize(n);
    for (int64_t t = 0; t < n; t++) {
        Box bx;
        for (int k = 0; k < 3; k++) bx.grow(positions + 3 * (int64_t)indices[3 * t + k]);
        b.tbox[t] = bx;
        for (int k = 0; k < 3; k++) b.cen[3 * t + k] = 0.5f * (bx.lo[k] + bx.hi[k]);
        b.order[t] = (uint32_t)t;
    }
    b.nodes.resize(2 * n + 2);
    b.next_node = 1;
    b.nodes[0].first = 0; b.nodes[0].count = n;
    b.build(0, 0);
    const int64_t nb = b.next_node.load();

    double sah = 0.0;
    if (width == 4) {
        emit_four_wide(b.nodes, r);
    } else {
        // Emit the two-wide layout: one wide node per internal binary node.
        std::vector<int64_t> wide_of(nb, -1);
        int64_t nw = 0;
        std::vector<int64_t> stack{0};
        // breadth-first numbering keeps siblings close in memory
        std::vector<int64_t> queue{0};
        if (b.nodes[0].left < 0) {
            wide_of[0] = nw++;
        } else {
            for (size_t q = 0; q < queue.size(); q++) {
                int64_t i = queue[q];
                wide_of[i] = nw++;
                for (int c = 0; c < 2; c++) {
                    int64_t ch = b.nodes[i].left + c;
                    if (b.nodes[ch].left >= 0) queue.push_back(ch);
                }
            }
        }
        r.nodes.resize(nw);
        auto leaf_ref = [&](const BinNode& x) { return make_leaf((uint32_t)x.first, (uint32_t)x.count); };
        auto put = [&](WideNode& w, int c, const BinNode& ch, uint32_t ref) {
            w.f[c * 4 + 0] = ch.box.lo[0]; w.f[c * 4 + 1] = ch.box.hi[0];
            w.f[c * 4 + 2] = ch.box.lo[1]; w.f[c * 4 + 3] = ch.box.hi[1];
            w.f[8 + c * 2 + 0] = ch.box.lo[2]; w.f[8 + c * 2 + 1] = ch.box.hi[2];
            std::memcpy(&w.f[12 + c], &ref, 4);
        };
        float root_area = b.nodes[0].box.area();
        if (b.nodes[0].left < 0) {
            WideNode w; std::memset(&w, 0, sizeof(w));
            put(w, 0, b.nodes[0], leaf_ref(b.nodes[0]));
            BinNode empty;  // inverted box never hit
            put(w, 1, empty, EMPTY_REF);
            r.nodes[0] = w;
        } else {
            for (int64_t i = 0; i < nb; i++) {
                if (wide_of[i] < 0) continue;
                const BinNode& n = b.nodes[i];
                WideNode w; std::memset(&w, 0, sizeof(w));
                for (int c = 0; c < 2; c++) {
                    const BinNode& ch = b.nodes[n.left + c];
                    uint32_t ref = ch.left >= 0 ? (uint32_t)wide_of[n.left + c] : leaf_ref(ch);
                    put(w, c, ch, ref);
                    if (root_area > 0.0f) {
                        float a = ch.box.area() / root_area;
                        sah += ch.left >= 0 ? COST_TRAVERSE * a : COST_INTERSECT * a * (double)ch.count;
                    }
                }
                r.nodes[wide_of[i]] = w;
            }
        }
    }
    r.tri_order = std::move(b.order);
    r.max_depth = b.depth_max.load();
    r.leaf_count = b.leaves.load();
    r.sah_cost = sah;
    return r;
}

}  // namespace dfr
