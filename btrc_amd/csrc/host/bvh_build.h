// Host construction of the 4-wide BVH traversed by the gfx950 trace kernels.
// Replaces the OptiX acceleration-structure builds (utils/optix/context.cpp:
// 158-295): a binned-SAH binary tree over the flattened world-space triangle
// list, collapsed greedily to 4-wide nodes by largest surface area.
#pragma once

#include <cstdint>
#include <vector>

#include "../core/traverse.h"

namespace btrc {

struct BvhBuildOptions
{
    int bins = 32;
    int max_leaf_size = 4;
    float traversal_cost = 1.0f;
    float intersection_cost = 1.0f;
};

struct Bvh4
{
    std::vector<BvhNode4> nodes;         // root is node 0
    std::vector<uint32_t> prim_order;    // leaf order -> flattened triangle index
    int depth = 0;                       // levels of 4-wide nodes
    int stack_bound = 0;                 // upper bound on traversal stack entries
    bool root_is_leaf = false;           // true when the whole scene is one leaf
};

// vertices: 9 floats (v0, v1, v2) per triangle
Bvh4 build_bvh4(const float *vertices, uint32_t triangle_count, const BvhBuildOptions &opt);

// boxes: 6 floats (lo xyz, hi xyz) per primitive; an empty box (lo > hi)
// is allowed and never hit
Bvh4 build_bvh4_boxes(const float *boxes, uint32_t count, const BvhBuildOptions &opt);

} // namespace btrc
