// BVH construction on the GPU: the parallel locally-ordered clustering
// builder (PLOC, Meister and Bittner) as the reference's vendored lib/bvh
// states it (LocallyOrderedClusteringBuilder<Bvh<float>, uint64_t>,
// lib/bvh/include/bvh/locally_ordered_clustering_builder.hpp, over
// MortonCodeBasedBuilder, morton_code_based_builder.hpp, and MortonEncoder,
// morton.hpp), its SAH leaf collapse (LeafCollapser, leaf_collapser.hpp) and
// the bottom-up refit (HierarchyRefitter, hierarchy_refitter.hpp), then the
// greedy collapse to the 4-wide nodes the trace kernels read. The reference
// builds its triangle structures with optixAccelBuild
// (utils/optix/context.cpp:158-295) and its volume tree with these lib/bvh
// builders on the CPU (core/volume/bvh.cpp:40-63).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../core/traverse.h"

namespace btrc {

// lib/bvh's Bvh<float>::Node (bvh.hpp:25-28): bounds (min x, max x, min y,
// max y, min z, max z); a leaf has primitive_count > 0 primitives starting
// at `first` in the primitive index list, an inner node has its children at
// `first` and `first + 1`
struct PlocNode
{
    float bounds[6];
    uint32_t count;
    uint32_t first;
};
static_assert(sizeof(PlocNode) == 32, "PlocNode mirrors lib/bvh's 32-byte node");

constexpr int PLOC_SEARCH_RADIUS = 14;        // LocallyOrderedClusteringBuilder::search_radius
constexpr float PLOC_TRAVERSAL_COST = 1.0f;   // SahBasedAlgorithm::traversal_cost

// Scratch and results of one build, all in device memory. Buffers grow to
// the largest build seen and are reused.
struct PlocBuffers
{
    PlocNode *nodes = nullptr, *nodes_b = nullptr;  // 2n - 1 each
    uint32_t *prims = nullptr, *prims_b = nullptr;  // n each
    uint64_t *codes = nullptr, *codes_b = nullptr;  // n each
    uint32_t *aux_a = nullptr, *aux_b = nullptr, *aux_c = nullptr; // 2n - 1 each
    float *box_partial = nullptr;                   // reduction partials
    void *temp = nullptr;                           // library scratch (sort, scan)
    size_t temp_bytes = 0;
    size_t capacity = 0;                            // primitives the buffers hold
    BvhNode4 *nodes4 = nullptr;                     // 4-wide nodes
    size_t nodes4_capacity = 0;
    uint32_t *frontier = nullptr, *frontier_b = nullptr; // 3 words per item
    int32_t *counters = nullptr;                    // small device counters

    void reserve(size_t n);
    void release();
    ~PlocBuffers() { release(); }
};

struct PlocResult
{
    uint32_t node_count = 0;     // binary nodes in PlocBuffers::nodes
    // bottom-up schedule: node index ranges [lo, hi) whose nodes depend only
    // on higher ranges, in the order they became final (the last is the root)
    std::vector<std::pair<uint32_t, uint32_t>> ranges;
};

// Binary tree over n boxes (6 floats each: lo xyz, hi xyz) in device memory.
// max_leaf < 0: the clustering's tree as built; 0: LeafCollapser as lib/bvh
// states it; k > 0: LeafCollapser that keeps leaves at k primitives or fewer.
// Results in buf.nodes / buf.prims.
PlocResult ploc_build(PlocBuffers &buf, const float *boxes, uint32_t n, int max_leaf, hipStream_t s);

// Boxes of world-space triangles (9 floats each) into `boxes` (6 floats each).
void triangle_boxes(const float *vertices, uint32_t n, float *boxes, hipStream_t s);

// HierarchyRefitter::refit over the tree of the last ploc_build: leaf boxes
// from new primitive boxes, inner boxes as the union of their children.
void ploc_refit(PlocBuffers &buf, const PlocResult &r, const float *boxes, hipStream_t s);

struct Bvh4Device
{
    uint32_t node_count = 0;
    int depth = 0;
    int stack_bound = 0;
    std::vector<uint32_t> level_offsets; // nodes of depth d: [offsets[d], offsets[d + 1])
};

// 4-wide nodes (into buf.nodes4, child slots as traversal codes) from the
// binary tree: each 4-wide node takes a binary node's children and opens
// the largest-area inner member until it holds four. Leaves must hold at
// most four primitives.
Bvh4Device ploc_to_bvh4(PlocBuffers &buf, const PlocResult &r, hipStream_t s);

// Refit of a one-level 4-wide tree: every slot box becomes the union of
// what it holds (a leaf's triangles from the leaf-ordered records, an inner
// slot's child node), level by level from the deepest; `order` lists the
// nodes grouped by level, `levels` the group boundaries (root level first).
void refit_bvh4(BvhNode4 *nodes, const uint32_t *order, const std::vector<uint32_t> &levels, const float4 *tris,
                hipStream_t s);

// The leaf-ordered triangle records of a one-level structure:
// (v0, flattened index), (v1, instance), (v2, primitive) per leaf slot.
void gather_triangles(const float *vertices, const uint32_t *tri_inst, const uint32_t *tri_prim,
                      const uint32_t *prims, uint32_t n, float4 *tris, hipStream_t s);

} // namespace btrc
