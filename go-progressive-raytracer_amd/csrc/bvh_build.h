// bvh_build.h — host-side binned-SAH BVH builder (float64 bounds) emitting 4-wide nodes.
#pragma once
#include <stdint.h>

#include <vector>

#include "prt_common.h"

namespace prt {

struct PrimBounds {
    double lo[3], hi[3];
};

struct BVHBuildResult {
    std::vector<Node4> nodes;       // nodes[0] is the root
    std::vector<uint32_t> order;    // leaf order: order[k] = input primitive index
    int32_t depth = 0;              // levels of 4-wide nodes
    int32_t max_stack = 0;          // traversal stack bound (3 pushes per level + 1)
};

// Binned SAH over primitive bounds (32 bins per axis, leaf <= max_leaf primitives),
// then collapse of the binary tree into 4-wide nodes by repeatedly opening the child
// with the largest surface area. Boxes are rounded outward to float32.
BVHBuildResult build_bvh4(const std::vector<PrimBounds>& prims, int max_leaf = 4);

}  // namespace prt
