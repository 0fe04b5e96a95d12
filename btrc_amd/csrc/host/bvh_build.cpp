#include "bvh_build.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <thread>

namespace btrc {

namespace {

struct Box
{
    float lo[3] = { INFINITY, INFINITY, INFINITY };
    float hi[3] = { -INFINITY, -INFINITY, -INFINITY };

    void grow(const float *p)
    {
        for(int a = 0; a < 3; ++a)
        {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    void grow(const Box &b)
    {
        for(int a = 0; a < 3; ++a)
        {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    bool empty() const { return lo[0] > hi[0]; }
    float area() const
    {
        if(empty())
            return 0.0f;
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct BNode
{
    Box box;
    int32_t left = -1, right = -1; // children (binary); -1 for leaves
    uint32_t first = 0, count = 0; // leaf range in prim_order
};

struct Builder
{
    const BvhBuildOptions &opt;
    std::vector<Box> prim_box;
    std::vector<float> centroid; // 3 per prim
    std::vector<uint32_t> order;
    std::vector<BNode> nodes;

    explicit Builder(const BvhBuildOptions &o) : opt(o) {}

    int32_t make_leaf(const Box &box, uint32_t first, uint32_t count)
    {
        BNode n;
        n.box = box;
        n.first = first;
        n.count = count;
        nodes.push_back(n);
        return (int32_t)nodes.size() - 1;
    }

    // iterative build with an explicit work list; returns the root index
    int32_t build(uint32_t n)
    {
        struct Task { int32_t node; uint32_t begin, end; };
        nodes.reserve(2 * (size_t)n / std::max(1, opt.max_leaf_size) + 16);
        Box root_box;
        for(uint32_t i = 0; i < n; ++i)
            root_box.grow(prim_box[i]);
        nodes.push_back(BNode{});
        nodes[0].box = root_box;
        std::vector<Task> stack;
        stack.push_back({ 0, 0, n });

        const int B = opt.bins;
        std::vector<Box> bin_box(B);
        std::vector<uint32_t> bin_cnt(B);
        std::vector<float> right_area(B);
        std::vector<uint32_t> right_cnt(B);

        while(!stack.empty())
        {
            const Task task = stack.back();
            stack.pop_back();
            const uint32_t count = task.end - task.begin;
            Box node_box = nodes[task.node].box;

            Box cbox;
            for(uint32_t i = task.begin; i < task.end; ++i)
                cbox.grow(&centroid[3 * order[i]]);

            int best_axis = -1;
            int best_split = -1;
            float best_cost = INFINITY;
            for(int axis = 0; axis < 3 && count > 1; ++axis)
            {
                const float extent = cbox.hi[axis] - cbox.lo[axis];
                if(!(extent > 0.0f))
                    continue;
                const float scale = (float)B / extent;
                for(int b = 0; b < B; ++b)
                {
                    bin_box[b] = Box{};
                    bin_cnt[b] = 0;
                }
                for(uint32_t i = task.begin; i < task.end; ++i)
                {
                    const uint32_t p = order[i];
                    int b = (int)((centroid[3 * p + axis] - cbox.lo[axis]) * scale);
                    b = std::min(std::max(b, 0), B - 1);
                    bin_box[b].grow(prim_box[p]);
                    bin_cnt[b]++;
                }
                Box acc;
                uint32_t cnt = 0;
                for(int b = B - 1; b > 0; --b)
                {
                    acc.grow(bin_box[b]);
                    cnt += bin_cnt[b];
                    right_area[b] = acc.area();
                    right_cnt[b] = cnt;
                }
                acc = Box{};
                cnt = 0;
                for(int b = 0; b < B - 1; ++b)
                {
                    acc.grow(bin_box[b]);
                    cnt += bin_cnt[b];
                    if(cnt == 0 || right_cnt[b + 1] == 0)
                        continue;
                    const float cost = acc.area() * (float)cnt + right_area[b + 1] * (float)right_cnt[b + 1];
                    if(cost < best_cost)
                    {
                        best_cost = cost;
                        best_axis = axis;
                        best_split = b + 1;
                    }
                }
            }

            const float parent_area = node_box.area();
            const float leaf_cost = opt.intersection_cost * (float)count;
            const float split_cost = parent_area > 0.0f
                ? opt.traversal_cost + opt.intersection_cost * best_cost / parent_area
                : INFINITY;

            uint32_t mid;
            if(best_axis < 0)
            {
                if(count <= (uint32_t)opt.max_leaf_size)
                {
                    nodes[task.node].first = task.begin;
                    nodes[task.node].count = count;
                    continue;
                }
                mid = task.begin + count / 2; // coincident centroids: split by index
            }
            else
            {
                if(count <= (uint32_t)opt.max_leaf_size && leaf_cost <= split_cost)
                {
                    nodes[task.node].first = task.begin;
                    nodes[task.node].count = count;
                    continue;
                }
                const float extent = cbox.hi[best_axis] - cbox.lo[best_axis];
                const float scale = (float)B / extent;
                auto it = std::partition(order.begin() + task.begin, order.begin() + task.end, [&](uint32_t p)
                {
                    int b = (int)((centroid[3 * p + best_axis] - cbox.lo[best_axis]) * scale);
                    b = std::min(std::max(b, 0), B - 1);
                    return b < best_split;
                });
                mid = (uint32_t)(it - order.begin());
                if(mid == task.begin || mid == task.end)
                    mid = task.begin + count / 2;
            }

            Box lb, rb;
            for(uint32_t i = task.begin; i < mid; ++i)
                lb.grow(prim_box[order[i]]);
            for(uint32_t i = mid; i < task.end; ++i)
                rb.grow(prim_box[order[i]]);
            BNode l, r;
            l.box = lb;
            r.box = rb;
            nodes.push_back(l);
            const int32_t li = (int32_t)nodes.size() - 1;
            nodes.push_back(r);
            const int32_t ri = (int32_t)nodes.size() - 1;
            nodes[task.node].left = li;
            nodes[task.node].right = ri;
            stack.push_back({ ri, mid, task.end });
            stack.push_back({ li, task.begin, mid });
        }
        return 0;
    }
};

void set_slot(BvhNode4 &n, int j, const Box &b)
{
    n.lo_x[j] = b.lo[0]; n.hi_x[j] = b.hi[0];
    n.lo_y[j] = b.lo[1]; n.hi_y[j] = b.hi[1];
    n.lo_z[j] = b.lo[2]; n.hi_z[j] = b.hi[2];
}

} // namespace

namespace {

Bvh4 build_from_boxes(std::vector<Box> &&boxes, const BvhBuildOptions &opt)
{
    const uint32_t triangle_count = (uint32_t)boxes.size();
    Bvh4 out;
    if(triangle_count == 0)
    {
        BvhNode4 n;
        for(int j = 0; j < 4; ++j)
        {
            set_slot(n, j, Box{});
            n.child[j] = -1;
            n.count[j] = 0;
        }
        out.nodes.push_back(n);
        out.depth = 1;
        out.stack_bound = 1;
        return out;
    }
    if(triangle_count >= (1u << 30))
        throw std::runtime_error("too many primitives for the BVH leaf encoding");

    Builder b(opt);
    b.prim_box = std::move(boxes);
    b.centroid.resize(3 * (size_t)triangle_count);
    b.order.resize(triangle_count);
    for(uint32_t i = 0; i < triangle_count; ++i)
    {
        const Box &box = b.prim_box[i];
        for(int a = 0; a < 3; ++a)
            b.centroid[3 * (size_t)i + a] = box.empty() ? 0.0f : 0.5f * (box.lo[a] + box.hi[a]);
        b.order[i] = i;
    }
    b.build(triangle_count);
    out.prim_order = b.order;

    // collapse to 4-wide nodes
    const std::vector<BNode> &bn = b.nodes;
    struct Item { int32_t bnode; int32_t slot_parent; int slot; int level; int pushed; };
    // A root that is itself a leaf becomes a 4-wide node with one leaf slot.
    std::vector<Item> work;
    out.nodes.push_back(BvhNode4{});
    work.push_back({ 0, 0, -1, 1, 0 });
    int max_depth = 1;
    int max_stack = 1;
    while(!work.empty())
    {
        const Item it = work.back();
        work.pop_back();
        const int32_t self = it.slot < 0 ? 0 : (int32_t)out.nodes.size();
        if(it.slot >= 0)
        {
            out.nodes.push_back(BvhNode4{});
            out.nodes[it.slot_parent].child[it.slot] = self;
        }
        std::vector<int32_t> set;
        if(bn[it.bnode].left < 0)
            set.push_back(it.bnode);
        else
        {
            set.push_back(bn[it.bnode].left);
            set.push_back(bn[it.bnode].right);
            while(set.size() < 4)
            {
                int best = -1;
                float best_area = -1.0f;
                for(int k = 0; k < (int)set.size(); ++k)
                {
                    if(bn[set[k]].left >= 0 && bn[set[k]].box.area() > best_area)
                    {
                        best_area = bn[set[k]].box.area();
                        best = k;
                    }
                }
                if(best < 0)
                    break;
                const int32_t x = set[best];
                set[best] = bn[x].left;
                set.push_back(bn[x].right);
            }
        }
        int inner = 0;
        for(int j = 0; j < 4; ++j)
        {
            BvhNode4 &n = out.nodes[self];
            if(j < (int)set.size())
            {
                const BNode &c = bn[set[j]];
                set_slot(n, j, c.box);
                if(c.left < 0)
                {
                    n.child[j] = ~(int32_t)c.first;
                    n.count[j] = (int32_t)c.count;
                }
                else
                {
                    n.child[j] = 0; // patched when the child node is emitted
                    n.count[j] = 0;
                    ++inner;
                }
            }
            else
            {
                set_slot(n, j, Box{});
                n.child[j] = ~0;
                n.count[j] = 0;
            }
        }
        // a visit pushes every hit child but the nearest, leaves included
        const int pushed_here = it.pushed + std::max(0, (int)set.size() - 1);
        max_stack = std::max(max_stack, pushed_here + 1);
        max_depth = std::max(max_depth, it.level);
        for(int j = (int)set.size() - 1; j >= 0; --j)
        {
            if(bn[set[j]].left >= 0)
                work.push_back({ set[j], self, j, it.level + 1, pushed_here });
        }
    }
    out.depth = max_depth;
    out.stack_bound = max_stack;
    out.root_is_leaf = bn[0].left < 0;
    return out;
}

} // namespace

Bvh4 build_bvh4(const float *vertices, uint32_t triangle_count, const BvhBuildOptions &opt)
{
    std::vector<Box> boxes(triangle_count);
    for(uint32_t i = 0; i < triangle_count; ++i)
    {
        boxes[i].grow(vertices + 9 * (size_t)i + 0);
        boxes[i].grow(vertices + 9 * (size_t)i + 3);
        boxes[i].grow(vertices + 9 * (size_t)i + 6);
    }
    return build_from_boxes(std::move(boxes), opt);
}

Bvh4 build_bvh4_boxes(const float *boxes, uint32_t count, const BvhBuildOptions &opt)
{
    std::vector<Box> bx(count);
    for(uint32_t i = 0; i < count; ++i)
    {
        const float *b = boxes + 6 * (size_t)i;
        if(b[0] <= b[3] && b[1] <= b[4] && b[2] <= b[5])
            for(int a = 0; a < 3; ++a)
            {
                bx[i].lo[a] = b[a];
                bx[i].hi[a] = b[3 + a];
            }
    }
    return build_from_boxes(std::move(bx), opt);
}

} // namespace btrc
