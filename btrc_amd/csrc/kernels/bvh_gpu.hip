// BVH construction on the GPU (see bvh_gpu.h for what each step restates).
//
// One thread per primitive or per node throughout; the clustering loop
// reads back one count per iteration to place the next. Sorting and scans
// use rocPRIM's device-wide radix sort (stable, as lib/bvh's LSD radix sort
// is) and scans. Bottom-up passes (leaf collapse, refit) need no atomics:
// every clustering iteration finalizes a contiguous index range whose nodes
// have their children in ranges finalized earlier, so one launch per range,
// in that order, is a valid bottom-up schedule.
#include "bvh_gpu.h"
#include "trace.h"

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <stdexcept>
#include <string>

#pragma clang fp contract(off)

namespace btrc {

namespace {

#define PLOC_CHECK(expr)                                                                            \
    do                                                                                              \
    {                                                                                               \
        hipError_t e_ = (expr);                                                                     \
        if(e_ != hipSuccess)                                                                        \
            throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_));            \
    } while(0)

constexpr int BLOCK = 256;
constexpr int REDUCE_BLOCKS = 1024;
constexpr uint64_t GRID_DIM = uint64_t(1) << 21; // MortonEncoder::max_grid_dim for 64-bit codes

inline dim3 grid_for(uint64_t n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }

// std::min / std::max argument order, so that signed zeros land as lib/bvh's do
__device__ __forceinline__ float mn(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float mx(float a, float b) { return a < b ? b : a; }

struct B6
{
    float v[6]; // node bounds layout: min x, max x, min y, max y, min z, max z
};

__device__ __forceinline__ B6 node_box(const PlocNode *nodes, uint32_t i)
{
    const float4 a = *reinterpret_cast<const float4 *>(&nodes[i].bounds[0]);
    const float2 b = *reinterpret_cast<const float2 *>(&nodes[i].bounds[4]);
    return B6{ { a.x, a.y, a.z, a.w, b.x, b.y } };
}

// BoundingBox::extend (bounding_box.hpp:23-27): a grown by b
__device__ __forceinline__ B6 extend(const B6 &a, const B6 &b)
{
    return B6{ { mn(a.v[0], b.v[0]), mx(a.v[1], b.v[1]), mn(a.v[2], b.v[2]), mx(a.v[3], b.v[3]), mn(a.v[4], b.v[4]),
                 mx(a.v[5], b.v[5]) } };
}

// BoundingBox::half_area, bounding_box.hpp:43-46
__device__ __forceinline__ float half_area(const B6 &b)
{
    const float d0 = b.v[1] - b.v[0], d1 = b.v[3] - b.v[2], d2 = b.v[5] - b.v[4];
    return (d0 + d1) * d2 + d0 * d1;
}

// a primitive box given as lo xyz, hi xyz
__device__ __forceinline__ B6 prim_box(const float *boxes, uint32_t i)
{
    const float *p = boxes + 6 * (size_t)i;
    return B6{ { p[0], p[3], p[1], p[4], p[2], p[5] } };
}

__device__ __forceinline__ void store_node(PlocNode *nodes, uint32_t i, const B6 &b, uint32_t count, uint32_t first)
{
    float4 *d = reinterpret_cast<float4 *>(&nodes[i]);
    d[0] = make_float4(b.v[0], b.v[1], b.v[2], b.v[3]);
    d[1] = make_float4(b.v[4], b.v[5], __uint_as_float(count), __uint_as_float(first));
}

__device__ __forceinline__ void copy_node(const PlocNode *src, uint32_t i, PlocNode *dst, uint32_t k)
{
    const float4 *s = reinterpret_cast<const float4 *>(&src[i]);
    float4 *d = reinterpret_cast<float4 *>(&dst[k]);
    d[0] = s[0];
    d[1] = s[1];
}

// ------------------------------------------------------------------ boxes

__global__ void k_triangle_boxes(const float *v, uint32_t n, float *boxes)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n)
        return;
    const float *p = v + 9 * (size_t)i;
    float *b = boxes + 6 * (size_t)i;
    for(int a = 0; a < 3; ++a)
    {
        b[a] = mn(mn(p[a], p[3 + a]), p[6 + a]);
        b[3 + a] = mx(mx(p[a], p[3 + a]), p[6 + a]);
    }
}

// union of all primitive boxes: per-block partials, then one block
__global__ void k_box_partials(const float *boxes, uint32_t n, float *partial)
{
    __shared__ float red[6][BLOCK];
    float lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
    for(uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const float *b = boxes + 6 * (size_t)i;
        for(int a = 0; a < 3; ++a)
        {
            lo[a] = fminf(lo[a], b[a]);
            hi[a] = fmaxf(hi[a], b[3 + a]);
        }
    }
    for(int a = 0; a < 3; ++a)
    {
        red[a][threadIdx.x] = lo[a];
        red[3 + a][threadIdx.x] = hi[a];
    }
    __syncthreads();
    for(int s = BLOCK / 2; s > 0; s >>= 1)
    {
        if((int)threadIdx.x < s)
            for(int a = 0; a < 3; ++a)
            {
                red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + s]);
                red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if(threadIdx.x < 6)
        partial[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_box_final(float *partial, int count)
{
    if(threadIdx.x != 0 || blockIdx.x != 0)
        return;
    float r[6] = { INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY };
    for(int k = 0; k < count; ++k)
        for(int a = 0; a < 3; ++a)
        {
            r[a] = fminf(r[a], partial[6 * k + a]);
            r[3 + a] = fmaxf(r[3 + a], partial[6 * k + 3 + a]);
        }
    for(int a = 0; a < 6; ++a)
        partial[a] = r[a];
}

// ------------------------------------------------------------------ Morton codes

// morton_split (morton.hpp:14-25) for 64-bit codes: bits spaced by two zeros
__device__ __forceinline__ uint64_t morton_split(uint64_t x)
{
    uint64_t mask = ~uint64_t(0) >> 32;
    x &= mask;
    for(int i = 5, n = 32; i > 0; --i, n >>= 1)
    {
        mask = (mask | (mask << n)) & ~(mask << (n / 2));
        x = (x | (x << n)) & mask;
    }
    return x;
}

// MortonEncoder::encode (morton.hpp:35-60) of each box centre 0.5 (lo + hi)
// on the 2^21 grid over the union box; the identity permutation beside it
__global__ void k_morton(const float *boxes, uint32_t n, const float *gbox, uint64_t *codes, uint32_t *idx)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n)
        return;
    const float *b = boxes + 6 * (size_t)i;
    uint64_t q[3];
    for(int a = 0; a < 3; ++a)
    {
        const float w = (float)GRID_DIM * (1.0f / (gbox[3 + a] - gbox[a]));
        const float off = -gbox[a] * w;
        const float c = 0.5f * (b[a] + b[3 + a]);
        const float g = c * w + off;
        // a flat axis (0 * inf) gives NaN: grid coordinate 0
        const uint64_t v = g > 0.0f ? (uint64_t)g : 0;
        q[a] = v < GRID_DIM - 1 ? v : GRID_DIM - 1;
    }
    codes[i] = morton_split(q[0]) | (morton_split(q[1]) << 1) | (morton_split(q[2]) << 2);
    idx[i] = i;
}

// leaves at [n - 1, 2n - 1) in sorted order (locally_ordered_clustering_builder.hpp:198-204)
__global__ void k_leaves(const float *boxes, const uint32_t *order, uint32_t n, PlocNode *nodes)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
        store_node(nodes, n - 1 + i, prim_box(boxes, order[i]), 1u, i);
}

// ------------------------------------------------------------------ clustering

// nearest neighbour within the search radius by half the area of the
// union box; ties go to the lowest index (cluster(), :64-101)
__global__ void k_nearest(const PlocNode *in, uint32_t begin, uint32_t end, uint32_t *nn)
{
    const uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= end)
        return;
    const uint32_t lo = i > begin + PLOC_SEARCH_RADIUS ? i - PLOC_SEARCH_RADIUS : begin;
    const uint32_t hi = min(i + PLOC_SEARCH_RADIUS + 1, end);
    const B6 bi = node_box(in, i);
    float best = 3.40282347e+38f;
    uint32_t best_j = 0xffffffffu;
    for(uint32_t j = lo; j < hi; ++j)
    {
        if(j == i)
            continue;
        const float d = half_area(extend(bi, node_box(in, j)));
        if(d < best)
        {
            best = d;
            best_j = j;
        }
    }
    nn[i] = best_j;
}

// 1 where i leads a mutual pair (i < j, nn[j] == i)
__global__ void k_mark(const uint32_t *nn, uint32_t begin, uint32_t end, uint32_t *mark)
{
    const uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= end)
        return;
    const uint32_t j = nn[i];
    mark[i] = (j != 0xffffffffu && i < j && nn[j] == i) ? 1u : 0u;
}

// cluster(), :124-156: mutual pairs become a parent in the next active
// range and their two nodes children at its end; the rest move down
__global__ void k_merge(const PlocNode *in, PlocNode *out, const uint32_t *nn, const uint32_t *mi, uint32_t begin,
                        uint32_t end, uint32_t unmerged_begin, uint32_t children_begin, int32_t *error)
{
    const uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= end)
        return;
    const uint32_t j = nn[i];
    if(j != 0xffffffffu && nn[j] == i)
    {
        if(i < j)
        {
            const uint32_t pos = unmerged_begin + j - begin - mi[j];
            const uint32_t first_child = children_begin + (mi[i] - 1) * 2;
            if(pos < unmerged_begin || pos >= children_begin || first_child < children_begin || first_child + 1 >= end)
            {
                atomicMax(error, 1);
                return;
            }
            store_node(out, pos, extend(node_box(in, j), node_box(in, i)), 0u, first_child);
            copy_node(in, i, out, first_child);
            copy_node(in, j, out, first_child + 1);
        }
    }
    else
    {
        const uint32_t pos = unmerged_begin + i - begin - mi[i];
        if(pos < unmerged_begin || pos >= children_begin)
        {
            atomicMax(error, 1);
            return;
        }
        copy_node(in, i, out, pos);
    }
}

__global__ void k_copy_nodes(const PlocNode *in, PlocNode *out, uint32_t lo, uint32_t hi)
{
    const uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if(i < hi)
        copy_node(in, i, out, i);
}

// ------------------------------------------------------------------ leaf collapse

__global__ void k_collapse_init(const PlocNode *nodes, uint32_t count, uint32_t *pc, uint32_t *nc)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= count)
        return;
    pc[i] = nodes[i].count;
    nc[i] = 1;
}

// LeafCollapser::collapse's bottom-up step (leaf_collapser.hpp:45-75) for
// the inner nodes of one range; cap > 0 also keeps leaves at cap or fewer
__global__ void k_collapse_range(const PlocNode *nodes, uint32_t lo, uint32_t hi, uint32_t *pc, uint32_t *nc,
                                 uint32_t cap)
{
    const uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= hi || nodes[i].count != 0)
        return;
    const uint32_t l = nodes[i].first, r = l + 1;
    const uint32_t lc = pc[l], rc = pc[r], total = lc + rc;
    if(lc == 0 || rc == 0 || (cap != 0 && total > cap))
        return;
    const float collapse_cost = half_area(node_box(nodes, i)) * ((float)total - PLOC_TRAVERSAL_COST);
    const float base_cost = half_area(node_box(nodes, l)) * (float)lc + half_area(node_box(nodes, r)) * (float)rc;
    if(collapse_cost <= base_cost)
    {
        pc[i] = total;
        pc[l] = 0;
        pc[r] = 0;
        nc[l] = 0;
        nc[r] = 0;
    }
}

__global__ void k_parents(const PlocNode *nodes, uint32_t count, uint32_t *parents)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= count)
        return;
    if(i == 0)
        parents[0] = 0;
    if(nodes[i].count == 0)
    {
        parents[nodes[i].first] = i;
        parents[nodes[i].first + 1] = i;
    }
}

// LeafCollapser::collapse's compaction (leaf_collapser.hpp:96-146): kept
// nodes move to their prefix-sum index; a collapsed node takes the
// primitives of its subtree in depth-first order (left first, through the
// parent links; children pairs start at odd indices, bvh.hpp:81-91)
__global__ void k_collapse_compact(const PlocNode *old_nodes, uint32_t count, const uint32_t *ncs, const uint32_t *pcs,
                                   const uint32_t *parents, const uint32_t *old_prims, PlocNode *new_nodes,
                                   uint32_t *new_prims, int32_t *error)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= count)
        return;
    if(i == 0)
    {
        const B6 b = node_box(old_nodes, 0);
        store_node(new_nodes, 0, b, 0u, ncs[old_nodes[0].first - 1]);
        return;
    }
    if(ncs[i] == ncs[i - 1])
        return;
    const uint32_t ni = ncs[i - 1];
    const B6 b = node_box(old_nodes, i);
    uint32_t first_primitive = pcs[i - 1];
    if(first_primitive != pcs[i])
    {
        store_node(new_nodes, ni, b, pcs[i] - first_primitive, first_primitive);
        uint32_t j = i;
        for(uint32_t steps = 0;; ++steps)
        {
            if(steps > 2 * count)
            {
                atomicMax(error, 1); // not a tree of the clustering's layout
                break;
            }
            const uint32_t c = old_nodes[j].count, f = old_nodes[j].first;
            if(c != 0)
            {
                if(first_primitive + c > pcs[i])
                {
                    atomicMax(error, 1);
                    break;
                }
                for(uint32_t k = 0; k < c; ++k)
                    new_prims[first_primitive + k] = old_prims[f + k];
                first_primitive += c;
                while((j & 1u) == 0 && j != i)
                    j = parents[j];
                if(j == i)
                    break;
                j = j + 1;
            }
            else
                j = f;
        }
    }
    else
        store_node(new_nodes, ni, b, 0u, ncs[old_nodes[i].first - 1]);
}

// old range boundaries to compacted ones: lo -> ncs[lo - 1], hi -> ncs[hi - 1]
__global__ void k_map_ranges(uint32_t *r, uint32_t count, const uint32_t *ncs)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k < count)
        r[k] = r[k] == 0 ? 0 : ncs[r[k] - 1];
}

// ------------------------------------------------------------------ refit

// HierarchyRefitter::refit (hierarchy_refitter.hpp:17-33) over one range:
// a leaf is the union of its primitives' boxes in order, an inner node the
// union of its children
__global__ void k_refit_range(PlocNode *nodes, uint32_t lo, uint32_t hi, const uint32_t *prims, const float *boxes)
{
    const uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= hi)
        return;
    const uint32_t c = nodes[i].count, f = nodes[i].first;
    B6 b;
    if(c != 0)
    {
        b = prim_box(boxes, prims[f]);
        for(uint32_t k = 1; k < c; ++k)
            b = extend(b, prim_box(boxes, prims[f + k]));
    }
    else
        b = extend(node_box(nodes, f), node_box(nodes, f + 1));
    store_node(nodes, i, b, c, f);
}

// ------------------------------------------------------------------ 4-wide nodes

struct SlotSet
{
    uint32_t node[4];
    int m;
};

// a binary node's children, opening the largest-area inner member until four
__device__ SlotSet slot_set(const PlocNode *nodes, uint32_t b)
{
    SlotSet s;
    if(nodes[b].count != 0)
    {
        s.node[0] = b;
        s.m = 1;
        return s;
    }
    s.node[0] = nodes[b].first;
    s.node[1] = nodes[b].first + 1;
    s.m = 2;
    while(s.m < 4)
    {
        int best = -1;
        float best_area = -1.0f;
        for(int k = 0; k < s.m; ++k)
        {
            if(nodes[s.node[k]].count == 0)
            {
                const float a = half_area(node_box(nodes, s.node[k]));
                if(a > best_area)
                {
                    best_area = a;
                    best = k;
                }
            }
        }
        if(best < 0)
            break;
        const uint32_t x = s.node[best];
        s.node[best] = nodes[x].first;
        s.node[s.m++] = nodes[x].first + 1;
    }
    return s;
}

// frontier item: binary node, 4-wide node index, stack entries pushed above it
__global__ void k_bvh4_count(const PlocNode *nodes, const uint32_t *frontier, uint32_t f, uint32_t *cnt,
                             int32_t *counters)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= f)
        return;
    const SlotSet s = slot_set(nodes, frontier[3 * t]);
    uint32_t inner = 0;
    for(int k = 0; k < s.m; ++k)
        inner += nodes[s.node[k]].count == 0 ? 1u : 0u;
    cnt[t] = inner;
    const int pushed_here = (int)frontier[3 * t + 2] + (s.m > 1 ? s.m - 1 : 0);
    atomicMax(&counters[0], pushed_here + 1);
}

__global__ void k_bvh4_write(const PlocNode *nodes, const uint32_t *frontier, uint32_t f, const uint32_t *off,
                             uint32_t base, BvhNode4 *out, uint32_t *next, int32_t *counters)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= f)
        return;
    const SlotSet s = slot_set(nodes, frontier[3 * t]);
    const uint32_t q = frontier[3 * t + 1];
    const uint32_t pushed_here = frontier[3 * t + 2] + (s.m > 1 ? (uint32_t)(s.m - 1) : 0u);
    BvhNode4 n;
    uint32_t k_inner = 0;
    for(int j = 0; j < 4; ++j)
    {
        if(j < s.m)
        {
            const B6 b = node_box(nodes, s.node[j]);
            n.lo_x[j] = b.v[0]; n.hi_x[j] = b.v[1];
            n.lo_y[j] = b.v[2]; n.hi_y[j] = b.v[3];
            n.lo_z[j] = b.v[4]; n.hi_z[j] = b.v[5];
            const uint32_t c = nodes[s.node[j]].count, first = nodes[s.node[j]].first;
            if(c != 0)
            {
                if(c > 4 || first > MAX_LEAF_FIRST)
                    atomicMax(&counters[1], 1); // leaf the traversal cannot encode
                n.child[j] = (int32_t)(CODE_LEAF | ((min(c, 4u) - 1) << 28) | (first & 0x0fffffffu));
                n.count[j] = (int32_t)c;
            }
            else
            {
                const uint32_t slot = off[t] + k_inner++;
                n.child[j] = (int32_t)(base + slot);
                n.count[j] = 0;
                next[3 * slot + 0] = s.node[j];
                next[3 * slot + 1] = base + slot;
                next[3 * slot + 2] = pushed_here;
            }
        }
        else
        {
            n.lo_x[j] = n.lo_y[j] = n.lo_z[j] = INFINITY;
            n.hi_x[j] = n.hi_y[j] = n.hi_z[j] = -INFINITY;
            n.child[j] = (int32_t)CODE_DONE;
            n.count[j] = 0;
        }
    }
    out[q] = n;
}

// one level of refit_bvh4
__global__ void k_refit4(BvhNode4 *nodes, const uint32_t *order, uint32_t lo, uint32_t hi, const float4 *tris)
{
    const uint32_t k = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if(k >= hi)
        return;
    BvhNode4 &n = nodes[order[k]];
    for(int j = 0; j < 4; ++j)
    {
        const uint32_t code = (uint32_t)n.child[j];
        if(code == CODE_DONE)
            continue;
        float b[6] = { INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY };
        if(code & CODE_LEAF)
        {
            const uint32_t first = code & 0x0fffffffu, count = ((code >> 28) & 3u) + 1;
            for(uint32_t t = 0; t < 3 * count; ++t)
            {
                const float4 v = tris[3 * (size_t)first + t];
                b[0] = fminf(b[0], v.x); b[1] = fminf(b[1], v.y); b[2] = fminf(b[2], v.z);
                b[3] = fmaxf(b[3], v.x); b[4] = fmaxf(b[4], v.y); b[5] = fmaxf(b[5], v.z);
            }
        }
        else
        {
            const BvhNode4 &c = nodes[code];
            for(int q = 0; q < 4; ++q)
            {
                b[0] = fminf(b[0], c.lo_x[q]); b[1] = fminf(b[1], c.lo_y[q]); b[2] = fminf(b[2], c.lo_z[q]);
                b[3] = fmaxf(b[3], c.hi_x[q]); b[4] = fmaxf(b[4], c.hi_y[q]); b[5] = fmaxf(b[5], c.hi_z[q]);
            }
        }
        n.lo_x[j] = b[0]; n.lo_y[j] = b[1]; n.lo_z[j] = b[2];
        n.hi_x[j] = b[3]; n.hi_y[j] = b[4]; n.hi_z[j] = b[5];
    }
}

__global__ void k_gather_tris(const float *v, const uint32_t *tri_inst, const uint32_t *tri_prim, const uint32_t *prims,
                              uint32_t n, float4 *tris)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k >= n)
        return;
    const uint32_t gid = prims[k];
    const float *w = v + 9 * (size_t)gid;
    tris[3 * (size_t)k + 0] = make_float4(w[0], w[1], w[2], __uint_as_float(gid));
    tris[3 * (size_t)k + 1] = make_float4(w[3], w[4], w[5], __uint_as_float(tri_inst[gid]));
    tris[3 * (size_t)k + 2] = make_float4(w[6], w[7], w[8], __uint_as_float(tri_prim[gid]));
}

// the first frontier: binary root, 4-wide node 0, nothing pushed above it
__global__ void k_root_frontier(uint32_t *frontier)
{
    frontier[0] = 0;
    frontier[1] = 0;
    frontier[2] = 0;
}

template<class T>
void dev_alloc(T *&p, size_t n)
{
    if(p)
        (void)hipFree(p);
    p = nullptr;
    if(n && hipMalloc(&p, n * sizeof(T)) != hipSuccess)
    {
        p = nullptr;
        throw std::bad_alloc();
    }
}

size_t scan_bytes(size_t n)
{
    size_t bytes = 0;
    PLOC_CHECK(rocprim::inclusive_scan(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr, n,
                                       rocprim::plus<uint32_t>()));
    size_t b2 = 0;
    PLOC_CHECK(rocprim::exclusive_scan(nullptr, b2, (const uint32_t *)nullptr, (uint32_t *)nullptr, 0u, n,
                                       rocprim::plus<uint32_t>()));
    return std::max(bytes, b2);
}

size_t sort_bytes(size_t n)
{
    size_t bytes = 0;
    PLOC_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                         (const uint32_t *)nullptr, (uint32_t *)nullptr, n, 0, 63));
    return bytes;
}

void inclusive_scan(PlocBuffers &buf, const uint32_t *in, uint32_t *out, size_t n, hipStream_t s)
{
    size_t bytes = buf.temp_bytes;
    PLOC_CHECK(rocprim::inclusive_scan(buf.temp, bytes, in, out, n, rocprim::plus<uint32_t>(), s));
}

void exclusive_scan(PlocBuffers &buf, const uint32_t *in, uint32_t *out, size_t n, hipStream_t s)
{
    size_t bytes = buf.temp_bytes;
    PLOC_CHECK(rocprim::exclusive_scan(buf.temp, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), s));
}

uint32_t read_u32(const uint32_t *p, hipStream_t s)
{
    uint32_t v = 0;
    PLOC_CHECK(hipMemcpyAsync(&v, p, sizeof(v), hipMemcpyDeviceToHost, s));
    PLOC_CHECK(hipStreamSynchronize(s));
    return v;
}

} // namespace

void PlocBuffers::reserve(size_t n)
{
    if(n <= capacity)
        return;
    release();
    const size_t m = 2 * n - 1;
    dev_alloc(nodes, m);
    dev_alloc(nodes_b, m);
    dev_alloc(prims, n);
    dev_alloc(prims_b, n);
    dev_alloc(codes, n);
    dev_alloc(codes_b, n);
    dev_alloc(aux_a, m);
    dev_alloc(aux_b, m);
    dev_alloc(aux_c, m);
    dev_alloc(box_partial, 6 * (size_t)REDUCE_BLOCKS);
    temp_bytes = std::max(sort_bytes(n), scan_bytes(m));
    uint8_t *t = nullptr;
    dev_alloc(t, temp_bytes);
    temp = t;
    dev_alloc(counters, 8);
    dev_alloc(frontier, 3 * n + 3);
    dev_alloc(frontier_b, 3 * n + 3);
    capacity = n;
}

void PlocBuffers::release()
{
    for(void *p : { (void *)nodes, (void *)nodes_b, (void *)prims, (void *)prims_b, (void *)codes, (void *)codes_b,
                    (void *)aux_a, (void *)aux_b, (void *)aux_c, (void *)box_partial, temp, (void *)counters,
                    (void *)frontier, (void *)frontier_b, (void *)nodes4 })
        if(p)
            (void)hipFree(p);
    nodes = nodes_b = nullptr;
    prims = prims_b = nullptr;
    codes = codes_b = nullptr;
    aux_a = aux_b = aux_c = nullptr;
    box_partial = nullptr;
    temp = nullptr;
    temp_bytes = 0;
    counters = nullptr;
    frontier = frontier_b = nullptr;
    nodes4 = nullptr;
    nodes4_capacity = 0;
    capacity = 0;
}

void triangle_boxes(const float *vertices, uint32_t n, float *boxes, hipStream_t s)
{
    if(n)
        hipLaunchKernelGGL(k_triangle_boxes, grid_for(n), dim3(BLOCK), 0, s, vertices, n, boxes);
}

PlocResult ploc_build(PlocBuffers &buf, const float *boxes, uint32_t n, int max_leaf, hipStream_t s)
{
    if(n == 0)
        throw std::runtime_error("ploc_build: no primitives");
    if(n > (1u << 30))
        throw std::runtime_error("ploc_build: too many primitives");
    buf.reserve(n);
    PlocResult r;
    const uint32_t node_count = 2 * n - 1;

    // union box, Morton codes, stable sort by code
    const int blocks = (int)std::min<uint64_t>(REDUCE_BLOCKS, (n + BLOCK - 1) / BLOCK);
    hipLaunchKernelGGL(k_box_partials, dim3(blocks), dim3(BLOCK), 0, s, boxes, n, buf.box_partial);
    hipLaunchKernelGGL(k_box_final, dim3(1), dim3(1), 0, s, buf.box_partial, blocks);
    hipLaunchKernelGGL(k_morton, grid_for(n), dim3(BLOCK), 0, s, boxes, n, buf.box_partial, buf.codes, buf.prims_b);
    size_t bytes = buf.temp_bytes;
    PLOC_CHECK(rocprim::radix_sort_pairs(buf.temp, bytes, buf.codes, buf.codes_b, buf.prims_b, buf.prims, n, 0, 63, s));
    hipLaunchKernelGGL(k_leaves, grid_for(n), dim3(BLOCK), 0, s, boxes, buf.prims, n, buf.nodes);
    PLOC_CHECK(hipGetLastError());

    // clustering (LocallyOrderedClusteringBuilder::build, :186-237)
    PLOC_CHECK(hipMemsetAsync(buf.counters, 0, 8 * sizeof(int32_t), s));
    PlocNode *in = buf.nodes, *out = buf.nodes_b;
    uint32_t begin = node_count - n, end = node_count, previous_end = end;
    while(end - begin > 1)
    {
        const uint32_t active = end - begin;
        hipLaunchKernelGGL(k_nearest, grid_for(active), dim3(BLOCK), 0, s, in, begin, end, buf.aux_a);
        hipLaunchKernelGGL(k_mark, grid_for(active), dim3(BLOCK), 0, s, buf.aux_a, begin, end, buf.aux_b);
        inclusive_scan(buf, buf.aux_b + begin, buf.aux_c + begin, active, s);
        const uint32_t merged = read_u32(buf.aux_c + end - 1, s);
        if(merged == 0)
            throw std::runtime_error("ploc_build: clustering made no progress");
        const uint32_t unmerged = active - merged;
        const uint32_t children_begin = end - 2 * merged;
        const uint32_t unmerged_begin = end - (2 * merged + unmerged);
        hipLaunchKernelGGL(k_merge, grid_for(active), dim3(BLOCK), 0, s, in, out, buf.aux_a, buf.aux_c, begin, end,
                           unmerged_begin, children_begin, buf.counters + 3);
        if(previous_end > end)
            hipLaunchKernelGGL(k_copy_nodes, grid_for(previous_end - end), dim3(BLOCK), 0, s, in, out, end,
                               previous_end);
        PLOC_CHECK(hipGetLastError());
        r.ranges.push_back({ children_begin, end });
        std::swap(in, out);
        previous_end = end;
        begin = unmerged_begin;
        end = children_begin;
    }
    if(begin != 0 || end != 1)
        throw std::runtime_error("ploc_build: clustering ended off the root");
    if(read_u32(reinterpret_cast<const uint32_t *>(buf.counters + 3), s) != 0)
        throw std::runtime_error("ploc_build: a cluster was placed outside its range");
    r.ranges.push_back({ 0u, 1u });
    r.node_count = node_count;
    if(in != buf.nodes)
        std::swap(buf.nodes, buf.nodes_b);

    if(max_leaf < 0 || node_count == 1)
        return r;

    // LeafCollapser::collapse (leaf_collapser.hpp:31-150)
    uint32_t *pc = buf.aux_a, *nc = buf.aux_b, *parents = buf.aux_c;
    hipLaunchKernelGGL(k_collapse_init, grid_for(node_count), dim3(BLOCK), 0, s, buf.nodes, node_count, pc, nc);
    for(const auto &rg : r.ranges)
        hipLaunchKernelGGL(k_collapse_range, grid_for(rg.second - rg.first), dim3(BLOCK), 0, s, buf.nodes, rg.first,
                           rg.second, pc, nc, (uint32_t)max_leaf);
    hipLaunchKernelGGL(k_parents, grid_for(node_count), dim3(BLOCK), 0, s, buf.nodes, node_count, parents);
    // prefix sums in place: node_counts into nc, primitive_counts into pc
    // (the scan reads its input before it writes the same index)
    uint32_t *ncs = buf.aux_b, *pcs = buf.aux_a;
    {
        // scan into the code buffers' space to keep in-place hazards out
        uint32_t *tmp = reinterpret_cast<uint32_t *>(buf.codes);
        inclusive_scan(buf, nc, tmp, node_count, s);
        PLOC_CHECK(hipMemcpyAsync(ncs, tmp, node_count * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        inclusive_scan(buf, pc, tmp, node_count, s);
        PLOC_CHECK(hipMemcpyAsync(pcs, tmp, node_count * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    }
    const uint32_t root_prims = read_u32(pcs, s);
    if(root_prims > 0)
    {
        // the root became a leaf over every primitive
        PlocNode root;
        PLOC_CHECK(hipMemcpyAsync(&root, buf.nodes, sizeof(root), hipMemcpyDeviceToHost, s));
        PLOC_CHECK(hipStreamSynchronize(s));
        root.count = root_prims;
        root.first = 0;
        PLOC_CHECK(hipMemcpyAsync(buf.nodes, &root, sizeof(root), hipMemcpyHostToDevice, s));
        PLOC_CHECK(hipStreamSynchronize(s));
        r.node_count = 1;
        r.ranges.assign(1, { 0u, 1u });
        return r;
    }
    const uint32_t new_count = read_u32(ncs + node_count - 1, s);
    PLOC_CHECK(hipMemsetAsync(buf.counters, 0, 8 * sizeof(int32_t), s));
    hipLaunchKernelGGL(k_collapse_compact, grid_for(node_count), dim3(BLOCK), 0, s, buf.nodes, node_count, ncs, pcs,
                       parents, buf.prims, buf.nodes_b, buf.prims_b, buf.counters + 2);
    PLOC_CHECK(hipGetLastError());
    if(read_u32(reinterpret_cast<const uint32_t *>(buf.counters + 2), s) != 0)
        throw std::runtime_error("ploc_build: leaf collapse walked off the tree");
    std::swap(buf.nodes, buf.nodes_b);
    std::swap(buf.prims, buf.prims_b);
    // the schedule carries over through the same prefix sums
    std::vector<uint32_t> bounds(2 * r.ranges.size());
    for(size_t k = 0; k < r.ranges.size(); ++k)
    {
        bounds[2 * k] = r.ranges[k].first;
        bounds[2 * k + 1] = r.ranges[k].second;
    }
    uint32_t *db = buf.frontier;
    PLOC_CHECK(hipMemcpyAsync(db, bounds.data(), bounds.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_map_ranges, grid_for(bounds.size()), dim3(BLOCK), 0, s, db, (uint32_t)bounds.size(), ncs);
    PLOC_CHECK(hipMemcpyAsync(bounds.data(), db, bounds.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PLOC_CHECK(hipStreamSynchronize(s));
    r.ranges.clear();
    for(size_t k = 0; k + 1 < bounds.size(); k += 2)
        if(bounds[k + 1] > bounds[k])
            r.ranges.push_back({ bounds[k], bounds[k + 1] });
    r.node_count = new_count;
    return r;
}

void ploc_refit(PlocBuffers &buf, const PlocResult &r, const float *boxes, hipStream_t s)
{
    for(const auto &rg : r.ranges)
        hipLaunchKernelGGL(k_refit_range, grid_for(rg.second - rg.first), dim3(BLOCK), 0, s, buf.nodes, rg.first,
                           rg.second, buf.prims, boxes);
    PLOC_CHECK(hipGetLastError());
}

Bvh4Device ploc_to_bvh4(PlocBuffers &buf, const PlocResult &r, hipStream_t s)
{
    // at most one 4-wide node per binary inner node, and one for a leaf root
    const size_t cap = std::max<size_t>(1, r.node_count / 2 + 1);
    if(cap > buf.nodes4_capacity)
    {
        dev_alloc(buf.nodes4, cap);
        buf.nodes4_capacity = cap;
    }
    PLOC_CHECK(hipMemsetAsync(buf.counters, 0, 8 * sizeof(int32_t), s));
    hipLaunchKernelGGL(k_root_frontier, dim3(1), dim3(1), 0, s, buf.frontier);
    uint32_t f = 1, base = 1;
    Bvh4Device out;
    out.level_offsets.push_back(0);
    uint32_t *cur = buf.frontier, *nxt = buf.frontier_b;
    while(f > 0)
    {
        hipLaunchKernelGGL(k_bvh4_count, grid_for(f), dim3(BLOCK), 0, s, buf.nodes, cur, f, buf.aux_a, buf.counters);
        exclusive_scan(buf, buf.aux_a, buf.aux_b, f, s);
        const uint32_t last_cnt = read_u32(buf.aux_a + f - 1, s);
        const uint32_t last_off = read_u32(buf.aux_b + f - 1, s);
        const uint32_t total = last_off + last_cnt;
        if((size_t)base + total > cap)
            throw std::runtime_error("ploc_to_bvh4: node estimate exceeded");
        hipLaunchKernelGGL(k_bvh4_write, grid_for(f), dim3(BLOCK), 0, s, buf.nodes, cur, f, buf.aux_b, base, buf.nodes4,
                           nxt, buf.counters);
        PLOC_CHECK(hipGetLastError());
        ++out.depth;
        out.level_offsets.push_back(base);
        base += total;
        f = total;
        std::swap(cur, nxt);
    }
    int32_t counters[2] = { 0, 0 };
    PLOC_CHECK(hipMemcpyAsync(counters, buf.counters, sizeof(counters), hipMemcpyDeviceToHost, s));
    PLOC_CHECK(hipStreamSynchronize(s));
    if(counters[1] != 0)
        throw std::runtime_error("ploc_to_bvh4: a leaf exceeds the traversal's encoding (4 triangles)");
    out.node_count = base;
    out.stack_bound = counters[0];
    return out;
}

void refit_bvh4(BvhNode4 *nodes, const uint32_t *order, const std::vector<uint32_t> &levels, const float4 *tris,
                hipStream_t s)
{
    for(size_t d = levels.size() - 1; d-- > 0;)
    {
        const uint32_t lo = levels[d], hi = levels[d + 1];
        if(hi > lo)
            hipLaunchKernelGGL(k_refit4, grid_for(hi - lo), dim3(BLOCK), 0, s, nodes, order, lo, hi, tris);
    }
    PLOC_CHECK(hipGetLastError());
}

void gather_triangles(const float *vertices, const uint32_t *tri_inst, const uint32_t *tri_prim,
                      const uint32_t *prims, uint32_t n, float4 *tris, hipStream_t s)
{
    if(n)
        hipLaunchKernelGGL(k_gather_tris, grid_for(n), dim3(BLOCK), 0, s, vertices, tri_inst, tri_prim, prims, n, tris);
}

} // namespace btrc
