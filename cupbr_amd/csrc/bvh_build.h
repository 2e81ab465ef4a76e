// Host BVH builder: binned-SAH binary tree down to single primitives (task-parallel over
// subtrees), collapsed into a BVH4 of 128-byte nodes in depth-first order by a dynamic program
// that minimises the SAH cost of the BVH4 and chooses its leaves (up to max_leaf primitives).
//
// The reference builds a median split over whole objects with one device
// allocation per node (lib/src/DataStructure/BoundingVolumeHierarchy.cpp:46-136);
// this builder works on single primitives (mesh triangles included) so
// traversal cost is logarithmic in the triangle count. Which primitive wins a
// tie is decided at traversal time by the reference rank, not by the tree.
#pragma once

#include <cstdint>
#include <vector>

#include "scene_types.h"

namespace cupbr {

struct BuildPrim {
    float mn[3], mx[3];
};

struct Bvh4 {
    std::vector<Node4> nodes;       // nodes[0] is the root
    std::vector<uint32_t> order;    // leaf order: order[k] = input primitive index
    uint32_t depth = 0;             // BVH4 levels
    uint32_t leaves = 0;
    double sah_cost = 0.0;
};

// SAH bins per axis and split of the binary build
#ifndef CUPBR_BVH_BINS
#define CUPBR_BVH_BINS 32
#endif
#ifndef CUPBR_BVH_MAX_LEAF
#define CUPBR_BVH_MAX_LEAF 8
#endif
// SAH cost of a BVH4 node visit (its 4 box tests) relative to one primitive test. 2.0 measured
// best for PT+NEE at 1080p on 1M triangles with the binary16 node step (Msamples/s: 1.0 -> 1023,
// 1.25 -> 1015, 1.5 -> 1016, 1.75 -> 1021, 2.0 -> 1026-1028, 2.5 -> 1015, 3.0 -> 999); with the
// earlier FP32 node step 1.5 was best (0.7 -> 627, 1.0 -> 633, 1.5 -> 637, 2.0 -> 631, 3.0 -> 610).
#ifndef CUPBR_BVH_TRAVERSAL_COST
#define CUPBR_BVH_TRAVERSAL_COST 2.0f
#endif
struct BvhBuildOptions {
    uint32_t max_leaf = CUPBR_BVH_MAX_LEAF;  // <= kMaxLeafPrims
    float traversal_cost = CUPBR_BVH_TRAVERSAL_COST;  // BVH4 node visit
    float intersect_cost = 1.0f;                      // one primitive test
    int threads = 0;            // 0: hardware concurrency
};

Bvh4 build_bvh4(const std::vector<BuildPrim>& prims, const BvhBuildOptions& opt);

// The NodeH form of every node. A decoded plane c + h 2^k lies at least 16 x 2^-24 x E outside the
// FP32 plane it stands for (E: the node's extent on that axis), which covers the rounding of the
// traversal's decode t = fma(h, ix 2^k, (c - o) ix): its error is at most 2^-23 (E |ix|) plus a
// few ulp of t, and the slab test's relative slack takes the latter. NaN bounds become the
// infinities that never cull; infinite ones stay infinite.
std::vector<NodeH> half_nodes(const std::vector<Node4>& nodes);

}  // namespace cupbr
