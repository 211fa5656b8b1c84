// bvh_build.cpp — binned-SAH BVH over float64 primitive bounds, collapsed to 4-wide nodes.
//
// Replaces the reference's midpoint-split binary BVH (pkg/geometry/bvh.go:64-160, leaf <= 8,
// unordered traversal) with a structure built for the device traversal: SAH splits, leaf <= 4,
// 128-byte 4-wide nodes with conservatively rounded float32 boxes. Closest-hit results do
// not depend on the tree shape, so the reference's hit semantics are kept.
#include "bvh_build.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>

namespace prt {

namespace {

struct Box {
    double lo[3], hi[3];
    void empty() {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::numeric_limits<double>::infinity();
            hi[a] = -std::numeric_limits<double>::infinity();
        }
    }
    void grow(const Box& b) {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    void grow(const double* p) {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    double area() const {
        double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct Node2 {
    Box box;
    int32_t left = -1, right = -1;  // inner node children
    uint32_t first = 0, count = 0;  // leaf range in the index array (count > 0)
};

constexpr int kBins = 32;

struct Builder {
    const std::vector<PrimBounds>& prims;
    std::vector<double> cent;  // 3 per prim
    std::vector<uint32_t> idx;
    std::vector<Node2> nodes;
    int max_leaf;

    Builder(const std::vector<PrimBounds>& p, int ml) : prims(p), max_leaf(ml) {
        size_t n = p.size();
        cent.resize(3 * n);
        idx.resize(n);
        for (size_t i = 0; i < n; i++) {
            for (int a = 0; a < 3; a++) cent[3 * i + a] = 0.5 * (p[i].lo[a] + p[i].hi[a]);
            idx[i] = (uint32_t)i;
        }
    }

    Box prim_box(uint32_t i) const {
        Box b;
        for (int a = 0; a < 3; a++) {
            b.lo[a] = prims[i].lo[a];
            b.hi[a] = prims[i].hi[a];
        }
        return b;
    }

    int32_t make_leaf(const Box& b, uint32_t first, uint32_t count) {
        Node2 n;
        n.box = b;
        n.first = first;
        n.count = count;
        nodes.push_back(n);
        return (int32_t)nodes.size() - 1;
    }

    int32_t build(uint32_t begin, uint32_t end) {
        Box b, cb;
        b.empty();
        cb.empty();
        for (uint32_t k = begin; k < end; k++) {
            uint32_t i = idx[k];
            b.grow(prim_box(i));
            cb.grow(&cent[3 * i]);
        }
        uint32_t n = end - begin;
        if (n <= 1) return make_leaf(b, begin, n);

        // binned SAH over the centroid bounds
        double best_cost = std::numeric_limits<double>::infinity();
        int best_axis = -1, best_split = -1;
        for (int a = 0; a < 3; a++) {
            double ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0)) continue;
            Box bb[kBins];
            uint32_t bc[kBins] = {0};
            for (int k = 0; k < kBins; k++) bb[k].empty();
            double scale = kBins / ext;
            for (uint32_t k = begin; k < end; k++) {
                uint32_t i = idx[k];
                int bi = (int)((cent[3 * i + a] - cb.lo[a]) * scale);
                bi = std::min(std::max(bi, 0), kBins - 1);
                bc[bi]++;
                bb[bi].grow(prim_box(i));
            }
            double right_area[kBins];
            uint32_t right_count[kBins];
            Box acc;
            acc.empty();
            uint32_t cnt = 0;
            for (int k = kBins - 1; k > 0; k--) {
                acc.grow(bb[k]);
                cnt += bc[k];
                right_area[k] = acc.area();
                right_count[k] = cnt;
            }
            acc.empty();
            cnt = 0;
            for (int k = 0; k < kBins - 1; k++) {
                acc.grow(bb[k]);
                cnt += bc[k];
                if (cnt == 0 || right_count[k + 1] == 0) continue;
                double cost = acc.area() * cnt + right_area[k + 1] * right_count[k + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = a;
                    best_split = k;
                }
            }
        }
        double parent_area = b.area();
        double leaf_cost = (double)n;
        double split_cost = parent_area > 0 ? 1.0 + best_cost / parent_area : leaf_cost;
        if (n <= (uint32_t)max_leaf && (best_axis < 0 || leaf_cost <= split_cost)) {
            return make_leaf(b, begin, n);
        }
        uint32_t mid;
        if (best_axis < 0) {
            mid = begin + n / 2;  // coincident centroids: split by index
        } else {
            int a = best_axis;
            double ext = cb.hi[a] - cb.lo[a];
            double scale = kBins / ext;
            double lo = cb.lo[a];
            auto it = std::partition(idx.begin() + begin, idx.begin() + end, [&](uint32_t i) {
                int bi = (int)((cent[3 * i + a] - lo) * scale);
                bi = std::min(std::max(bi, 0), kBins - 1);
                return bi <= best_split;
            });
            mid = (uint32_t)(it - idx.begin());
            if (mid == begin || mid == end) mid = begin + n / 2;
        }
        int32_t id = (int32_t)nodes.size();
        nodes.emplace_back();
        nodes[id].box = b;
        int32_t l = build(begin, mid);
        int32_t r = build(mid, end);
        nodes[id].left = l;
        nodes[id].right = r;
        return id;
    }
};

float round_down(double x) {
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -std::numeric_limits<float>::infinity());
    return f;
}
float round_up(double x) {
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, std::numeric_limits<float>::infinity());
    return f;
}

struct Collapser {
    const Builder& b;
    BVHBuildResult& out;
    Collapser(const Builder& bb, BVHBuildResult& o) : b(bb), out(o) {}

    uint32_t leaf_word(const Node2& n) {
        uint32_t first = (uint32_t)out.order.size();
        for (uint32_t k = 0; k < n.count; k++) out.order.push_back(b.idx[n.first + k]);
        return make_leaf(first, n.count);
    }

    // Emits the 4-wide tree in breadth-first order (the top levels are the first nodes, so a
    // kernel can keep them in LDS), starting from binary node `root2` (inner or a single leaf).
    void emit_bfs(int32_t root2) {
        struct Item {
            int32_t n2;
            int depth;
            int32_t id;
        };
        std::vector<Item> queue;
        queue.reserve(b.nodes.size() / 2 + 4);
        out.nodes.emplace_back();
        queue.push_back({root2, 1, 0});
        for (size_t qi = 0; qi < queue.size(); qi++) {
            const Item it = queue[qi];
            out.depth = std::max(out.depth, it.depth);
            int32_t kids[4];
            int nk = 0;
            const Node2& root = b.nodes[it.n2];
            if (root.count > 0 || (root.left < 0 && root.right < 0)) {
                kids[nk++] = it.n2;
            } else {
                kids[nk++] = root.left;
                kids[nk++] = root.right;
                while (nk < 4) {
                    int best = -1;
                    double best_area = -1;
                    for (int k = 0; k < nk; k++) {
                        const Node2& c = b.nodes[kids[k]];
                        if (c.count > 0) continue;
                        double a = c.box.area();
                        if (a > best_area) {
                            best_area = a;
                            best = k;
                        }
                    }
                    if (best < 0) break;
                    int32_t open = kids[best];
                    kids[best] = b.nodes[open].left;
                    kids[nk++] = b.nodes[open].right;
                }
            }
            Node4 node;
            for (int k = 0; k < 4; k++) {
                node.lox[k] = node.loy[k] = node.loz[k] = std::numeric_limits<float>::infinity();
                node.hix[k] = node.hiy[k] = node.hiz[k] = -std::numeric_limits<float>::infinity();
                node.child[k] = kEmptyChild;
                node.pad[k] = 0;
            }
            for (int k = 0; k < nk; k++) {
                const Node2& c = b.nodes[kids[k]];
                uint32_t word;
                if (c.count > 0) {
                    word = leaf_word(c);
                } else if (c.left < 0) {
                    continue;  // empty leaf (no primitives)
                } else {
                    word = (uint32_t)out.nodes.size();
                    out.nodes.emplace_back();
                    queue.push_back({kids[k], it.depth + 1, (int32_t)word});
                }
                node.lox[k] = round_down(c.box.lo[0]);
                node.loy[k] = round_down(c.box.lo[1]);
                node.loz[k] = round_down(c.box.lo[2]);
                node.hix[k] = round_up(c.box.hi[0]);
                node.hiy[k] = round_up(c.box.hi[1]);
                node.hiz[k] = round_up(c.box.hi[2]);
                node.child[k] = word;
            }
            out.nodes[it.id] = node;
        }
    }
};

}  // namespace

BVHBuildResult build_bvh4(const std::vector<PrimBounds>& prims, int max_leaf) {
    BVHBuildResult out;
    if (max_leaf < 1) max_leaf = 1;
    if (max_leaf > 7) max_leaf = 7;
    if (prims.empty()) {
        Node4 node;
        for (int k = 0; k < 4; k++) {
            node.lox[k] = node.loy[k] = node.loz[k] = std::numeric_limits<float>::infinity();
            node.hix[k] = node.hiy[k] = node.hiz[k] = -std::numeric_limits<float>::infinity();
            node.child[k] = kEmptyChild;
            node.pad[k] = 0;
        }
        out.nodes.push_back(node);
        out.depth = 1;
        out.max_stack = 4;
        return out;
    }
    Builder b(prims, max_leaf);
    b.nodes.reserve(prims.size() * 2 / max_leaf + 16);
    int32_t root = b.build(0, (uint32_t)prims.size());
    out.nodes.reserve(b.nodes.size() / 2 + 4);
    out.order.reserve(prims.size());
    Collapser c(b, out);
    c.emit_bfs(root);
    out.max_stack = 3 * out.depth + 4;
    return out;
}

}  // namespace prt
