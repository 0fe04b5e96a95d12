// C entry points over lib/bvh (the reference's vendored CPU BVH library),
// compiled from its headers where they lie in the reference tree by
// oracle/build_ref.sh. Test infrastructure only.
//
// Builder: BinnedSahBuilder<Bvh, 16> as the lib/bvh benchmark's default
// (lib/bvh/test/benchmark.cpp:331-337); traversal: SingleRayTraverser with
// FastNodeIntersector and ClosestPrimitiveIntersector / AnyPrimitiveIntersector
// (benchmark.cpp:133-150). Statistics per ray as SingleRayTraverser::Statistics
// (single_ray_traverser.hpp:132-135).
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include <bvh/bvh.hpp>
#include <bvh/binned_sah_builder.hpp>
#include <bvh/sweep_sah_builder.hpp>
#include <bvh/single_ray_traverser.hpp>
#include <bvh/locally_ordered_clustering_builder.hpp>
#include <bvh/leaf_collapser.hpp>
#include <bvh/hierarchy_refitter.hpp>
#include <bvh/primitive_intersectors.hpp>
#include <bvh/triangle.hpp>

using Scalar = float;
using Vector3 = bvh::Vector3<Scalar>;
using Triangle = bvh::Triangle<Scalar>;
using Ray = bvh::Ray<Scalar>;
using Bvh = bvh::Bvh<Scalar>;

struct RefScene
{
    std::vector<Triangle> tris;
    Bvh bvh;
};

extern "C" {

void *bvhref_create(const float *v, int64_t n, int32_t builder)
{
    auto *s = new RefScene;
    s->tris.reserve(n);
    for(int64_t i = 0; i < n; ++i)
    {
        const float *p = v + 9 * i;
        s->tris.emplace_back(Vector3(p[0], p[1], p[2]), Vector3(p[3], p[4], p[5]), Vector3(p[6], p[7], p[8]));
    }
    auto [bboxes, centers] = bvh::compute_bounding_boxes_and_centers(s->tris.data(), s->tris.size());
    auto global_bbox = bvh::compute_bounding_boxes_union(bboxes.get(), s->tris.size());
    if(builder == 1)
    {
        bvh::SweepSahBuilder<Bvh> b(s->bvh);
        b.build(global_bbox, bboxes.get(), centers.get(), s->tris.size());
    }
    else
    {
        bvh::BinnedSahBuilder<Bvh, 16> b(s->bvh);
        b.build(global_bbox, bboxes.get(), centers.get(), s->tris.size());
    }
    return s;
}

void bvhref_destroy(void *p) { delete static_cast<RefScene *>(p); }

// LocallyOrderedClusteringBuilder<Bvh, uint64_t> over n boxes (6 floats:
// lo xyz, hi xyz) with centres 0.5 (lo + hi) and the boxes' union as the
// global box, the way core/volume/bvh.cpp:40-60 calls it; then LeafCollapser
// when collapse != 0, then HierarchyRefitter onto refit_boxes when given
// (leaf box: the union of its primitives' boxes in order). nodes: room for
// 2n - 1 nodes of 8 words (bounds, primitive_count, first); prims: n.
// Returns the node count.
int64_t bvhref_ploc(const float *boxes, int64_t n, int32_t collapse, const float *refit_boxes, uint32_t *nodes,
                    uint32_t *prims)
{
    using BBox = bvh::BoundingBox<Scalar>;
    std::vector<BBox> bb(n);
    std::vector<Vector3> centers(n);
    for(int64_t i = 0; i < n; ++i)
    {
        const float *b = boxes + 6 * i;
        bb[i] = BBox(Vector3(b[0], b[1], b[2]), Vector3(b[3], b[4], b[5]));
        centers[i] = Vector3((b[0] + b[3]) * 0.5f, (b[1] + b[4]) * 0.5f, (b[2] + b[5]) * 0.5f);
    }
    const BBox global = bvh::compute_bounding_boxes_union(bb.data(), (size_t)n);
    Bvh tree;
    bvh::LocallyOrderedClusteringBuilder<Bvh, uint64_t> builder(tree);
    builder.build(global, bb.data(), centers.data(), (size_t)n);
    if(collapse)
    {
        bvh::LeafCollapser<Bvh> collapser(tree);
        collapser.collapse();
    }
    if(refit_boxes)
    {
        bvh::HierarchyRefitter<Bvh> refitter(tree);
        refitter.refit([&](Bvh::Node &leaf) {
            auto box = [&](size_t k) {
                const float *b = refit_boxes + 6 * tree.primitive_indices[leaf.first_child_or_primitive + k];
                return BBox(Vector3(b[0], b[1], b[2]), Vector3(b[3], b[4], b[5]));
            };
            BBox u = box(0);
            for(size_t k = 1; k < leaf.primitive_count; ++k)
                u.extend(box(k));
            leaf.bounding_box_proxy() = u;
        });
    }
    static_assert(sizeof(Bvh::Node) == 32, "lib/bvh node layout");
    std::memcpy(nodes, tree.nodes.get(), tree.node_count * sizeof(Bvh::Node));
    for(int64_t i = 0; i < n; ++i)
        prims[i] = (uint32_t)tree.primitive_indices[i];
    return (int64_t)tree.node_count;
}

// rays: 8 floats each (o.xyz, tmin, d.xyz, tmax)
void bvhref_trace(void *p, const float *rays, int64_t n, int32_t any, uint32_t *prim, float *t, float *u, float *v,
                  uint32_t *steps, uint32_t *isects, int32_t threads)
{
    const RefScene &s = *static_cast<RefScene *>(p);
    bvh::SingleRayTraverser<Bvh> traverser(s.bvh);
    bvh::ClosestPrimitiveIntersector<Bvh, Triangle> closest(s.bvh, s.tris.data());
    bvh::AnyPrimitiveIntersector<Bvh, Triangle> anyhit(s.bvh, s.tris.data());
#pragma omp parallel for schedule(dynamic, 1024) num_threads(threads)
    for(int64_t i = 0; i < n; ++i)
    {
        const float *r = rays + 8 * i;
        Ray ray(Vector3(r[0], r[1], r[2]), Vector3(r[4], r[5], r[6]), r[3], r[7]);
        bvh::SingleRayTraverser<Bvh>::Statistics st;
        if(any)
        {
            auto hit = traverser.traverse(ray, anyhit, st);
            prim[i] = hit ? 0u : 0xffffffffu;
            t[i] = hit ? hit->t : r[7];
            u[i] = v[i] = 0.0f;
        }
        else
        {
            auto hit = traverser.traverse(ray, closest, st);
            prim[i] = hit ? (uint32_t)hit->primitive_index : 0xffffffffu;
            t[i] = hit ? hit->intersection.t : r[7];
            u[i] = hit ? hit->intersection.u : 0.0f;
            v[i] = hit ? hit->intersection.v : 0.0f;
        }
        steps[i] = (uint32_t)st.traversal_steps;
        isects[i] = (uint32_t)st.intersections;
    }
}

} // extern "C"
