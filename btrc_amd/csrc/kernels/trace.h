// Software BVH traversal for gfx950 (CDNA4 has no ray-tracing instructions).
//
// Persistent "while-while" traversal (Aila and Laine, "Understanding the
// Efficiency of Ray Traversal on GPUs", HPG 2009) laid out for 64-wide
// wavefronts:
//  * one wavefront per workgroup, one lane per ray; once REFILL_MIN lanes
//    have ended their rays they take new ones from the wavefront's granule
//    (one atomic per wavefront for a new granule, below), so no lane waits
//    for the slowest ray of its batch;
//  * a lane tests a leaf's triangles as soon as it meets the leaf (the
//    default), or, with leaf postponing, keeps stepping through inner nodes
//    until every lane of the wavefront holds a leaf, so that leaf tests run
//    together;
//  * the traversal stack keeps its top SHORT_STACK entries in LDS, laid out
//    [entry][lane] (a wavefront's pushes hit consecutive words, no bank
//    conflicts), and deeper entries in a per-lane global overflow area sized
//    from the tree's exact worst case;
//  * children are visited nearest first; each stack entry keeps its entry
//    distance so popped subtrees beyond the current hit are skipped.
// Boxes use Ize's conservative slab test and a 1 + 2^-20 culling margin, so
// rounding never culls a triangle: the closest hit is the BVH-independent
// minimum of (t, flattened triangle index).
#pragma once

#include <hip/hip_runtime.h>

#include "../core/traverse.h"

namespace btrc {

constexpr int TRACE_BLOCK = 64;   // one wavefront per workgroup
constexpr int SHORT_STACK = 16;   // LDS stack entries per lane
constexpr int REFILL_MIN = 16;    // idle lanes that trigger a refill
constexpr int TRACE_WAVES_PER_CU = 20; // persistent grid: wavefronts per compute unit (default)
constexpr int TRACE_WAVES_PER_CU_MAX = 32;

// Device-scope atomics on one address serialise at about 11 ns each on
// MI355X (a wavefront-per-atomic microbenchmark: 65,536 atomics on one word
// take 746 us, spread over 64 words 22 us). Work counters are therefore
// striped: COUNTER_STRIPES counters, each on its own 128-byte line
// (COUNTER_STRIDE 32-bit words apart).
constexpr int COUNTER_STRIPES = 16;
constexpr int COUNTER_STRIDE = 32;
constexpr int FETCH_WORDS = COUNTER_STRIPES * COUNTER_STRIDE; // a traversal's fetch counters
constexpr int GRANULE = 64; // rays a traversal wavefront takes from a chunk counter at a time

// traversal schedules: leaves tested when reached, leaves postponed until
// every lane holds one, or the two-level (instance) structure
enum : int
{
    SCHED_EAGER = 0,
    SCHED_POSTPONE = 1,
    SCHED_TWO_LEVEL = 2,          // | SCHED_POSTPONE for postponed leaves
    SCHED_TWO_LEVEL_POSTPONE = 3,
};

// stack entry / node codes
constexpr uint32_t CODE_LEAF = 0x80000000u;   // | (count - 1) << 28 | first triangle
constexpr uint32_t CODE_INST = 0x40000000u;   // with CODE_LEAF: | instance (top-level leaf)
constexpr uint32_t CODE_DONE = 0xffffffffu;
constexpr uint32_t CODE_EXIT = 0xfffffffeu;   // stack marker: back to the top level
constexpr uint32_t CODE_EXIT_NEXT = 0xfffffffdu; // node: the stack top is CODE_EXIT (leaf postponed)
constexpr uint32_t MAX_LEAF_FIRST = 0x0ffffffeu;

// an instance of the two-level structure: the top level's leaves point here
struct TraceInstance
{
    float w2l[12];       // world_to_local rows
    uint32_t blas_root;  // node index of its geometry's tree
    uint32_t prim_base;  // flattened index of its first triangle
    uint32_t pad[2];
};

struct TraceScene
{
    const BvhNode4 *nodes;
    // 3 float4 per triangle in leaf order. One level (world space):
    //   (v0.xyz, flattened index), (v1.xyz, instance), (v2.xyz, prim in geometry);
    // two levels (local space): (v0.xyz, prim in geometry), (v1.xyz, 0), (v2.xyz, 0)
    const float4 *tris;
    const TraceInstance *inst; // two levels only
    float extent;              // largest world coordinate magnitude of the triangles
};

// per-lane global stack entries below the LDS part: [entry - SHORT_STACK][lane]
struct TraceOverflow
{
    uint32_t *code;
    float *t;
    int32_t lanes;   // stride between entries = lanes of the launch grid
};

struct HitRecord
{
    float t, b1, b2;
    uint32_t gid;   // flattened triangle index, 0xffffffff on a miss
    uint32_t inst;
    uint32_t prim;
};

__device__ __forceinline__ uint32_t leaf_code(int32_t first, int32_t count)
{
    return CODE_LEAF | ((uint32_t)(count - 1) << 28) | (uint32_t)first;
}



// The two-level variant (TWO): node 0 roots the instance tree, whose leaves
// are instance codes; entering one pushes CODE_EXIT, moves the ray into the
// instance's local space (same parametric t) and continues at its
// geometry's tree; popping CODE_EXIT restores the world ray. A lane holding
// a postponed leaf keeps descending only inside the current instance: it
// stops at an instance entry or before popping CODE_EXIT (node =
// CODE_EXIT_NEXT), so the leaf is tested with its own instance's ray.
//
// Traces rays [0, n): each wavefront's first TRACE_BLOCK rays by its
// workgroup index, the rest [gridDim.x * TRACE_BLOCK, n) in granules of
// GRANULE rays dealt round robin to C = min(COUNTER_STRIPES, gridDim.x)
// chunks, chunk k's granules handed out by counter fetch[k * COUNTER_STRIDE]
// (all FETCH_WORDS words start at 0) to the wavefronts of workgroups
// b = k (mod C). Sink provides
//   uint32_t load(uint32_t base, uint32_t i, V3 &o, V3 &d, float &tmin, float &tmax)
//       (ray i of a refill that hands out rays base .. base + 63; returns
//       the key its result is stored under)
//   void store(uint32_t key, const HitRecord &h)
// ANY ends a ray at its first hit in [tmin, tmax]. POSTPONE selects the
// while-while schedule (leaves wait until every lane holds one); without it
// a lane tests a leaf as soon as it meets it, which ends occluded shadow
// rays earlier.
template<int TRI, bool ANY, bool POSTPONE, bool TWO, class Sink>
__device__ __forceinline__ void trace_persistent(
    const TraceScene &sc, Sink &sink, int32_t n, uint32_t *fetch, const TraceOverflow &ovf)
{
    __shared__ uint32_t s_code[SHORT_STACK * TRACE_BLOCK];
    __shared__ float s_t[SHORT_STACK * TRACE_BLOCK];
    const int lane = (int)threadIdx.x;
    const int gl = (int)(blockIdx.x * TRACE_BLOCK + threadIdx.x);
    const uint64_t lanes_below = (1ull << lane) - 1ull;

    bool active = false;
    uint32_t ray = 0; // the sink's key of the lane's ray
    V3 o = v3s(0.0f), d = v3s(0.0f);
    float tmin = 0.0f;
    float margin = 0.0f; // culling margin of the ray (cull_margin)
    RayBoxPre rb;
    uint32_t near_x = 0, near_y = 32, near_z = 64; // byte offsets of the near planes in a node
    WatertightPre wt;
    MTPre mt;
    HitRecord h;
    uint32_t node = CODE_DONE;
    int32_t sp = 0;
    // Rays [0, static_span) go out as one batch of TRACE_BLOCK per wavefront
    // (workgroup b's first refill takes rays TRACE_BLOCK b + 0..63), so a
    // small launch takes no atomics. The rest go out in granules of
    // GRANULE rays: granule g belongs to chunk g % C, and chunk k's
    // counter hands its granules out in order to the wavefronts of
    // workgroups b = k (mod C). Every chunk spans the whole queue (balanced
    // work) and the granules in flight stay close together (coherent rays).
    // A wavefront keeps the rest of its granule for its next refill.
    const uint32_t static_span = gridDim.x * (uint32_t)TRACE_BLOCK;
    const uint32_t chunks = min((uint32_t)COUNTER_STRIPES, gridDim.x); // every chunk has wavefronts
    const uint32_t home = blockIdx.x % chunks;
    uint32_t *counter = fetch + home * COUNTER_STRIDE;
    // wavefront-uniform: reserved rays not yet handed out, whether the home
    // chunk has run dry, and whether no ray is left for this wavefront
    uint32_t res_lo = min((uint32_t)n, blockIdx.x * (uint32_t)TRACE_BLOCK);
    uint32_t res_hi = min((uint32_t)n, res_lo + (uint32_t)TRACE_BLOCK);
    bool dry = static_span >= (uint32_t)n;
    bool exhausted = dry && res_lo == res_hi;
    uint32_t cur_inst = 0xffffffffu, cur_base = 0; // two levels: the instance being traversed

    auto set_ray = [&](V3 ro, V3 rd) {
        rb = raybox_pre(ro, rd);
        near_x = rb.neg_x ? 16u : 0u;   // hi_x sits 16 bytes after lo_x
        near_y = rb.neg_y ? 48u : 32u;
        near_z = rb.neg_z ? 80u : 64u;
        if(TRI == TRI_WATERTIGHT)
            wt = watertight_pre(ro, rd);
        else
            mt = mt_pre(ro, rd);
    };

    auto push = [&](uint32_t code, float t) {
        if(sp < SHORT_STACK)
        {
            s_code[sp * TRACE_BLOCK + lane] = code;
            s_t[sp * TRACE_BLOCK + lane] = t;
        }
        else
        {
            const size_t k = (size_t)(sp - SHORT_STACK) * ovf.lanes + gl;
            ovf.code[k] = code;
            ovf.t[k] = t;
        }
        ++sp;
    };
    // next entry whose subtree may still hold a closer hit, or CODE_DONE;
    // stay_inside: stop before leaving the current instance
    auto pop = [&](bool stay_inside) -> uint32_t {
        const float cull = cull_distance(h.t, margin);
        while(sp > 0)
        {
            --sp;
            uint32_t code;
            float t;
            if(sp < SHORT_STACK)
            {
                code = s_code[sp * TRACE_BLOCK + lane];
                t = s_t[sp * TRACE_BLOCK + lane];
            }
            else
            {
                const size_t k = (size_t)(sp - SHORT_STACK) * ovf.lanes + gl;
                code = ovf.code[k];
                t = ovf.t[k];
            }
            if(TWO && code == CODE_EXIT)
            {
                if(stay_inside)
                {
                    ++sp; // left for the lane to pop once its leaf is tested
                    return CODE_EXIT_NEXT;
                }
                set_ray(o, d); // back to the world ray
                cur_inst = 0xffffffffu;
                continue;
            }
            if(t <= cull)
                return code;
        }
        return CODE_DONE;
    };

    // tests the triangles of a leaf code; true on a hit (ANY stops at the first)
    auto test_leaf = [&](uint32_t code) -> bool {
        const int32_t first = (int32_t)(code & 0x0fffffffu);
        const int32_t cnt = (int32_t)((code >> 28) & 3u) + 1;
        bool any_hit = false;
        for(int32_t q = 0; q < cnt; ++q)
        {
            const float4 a = sc.tris[3 * (first + q) + 0];
            const float4 b = sc.tris[3 * (first + q) + 1];
            const float4 cc = sc.tris[3 * (first + q) + 2];
            const uint32_t gid = TWO ? cur_base + __float_as_uint(a.w) : __float_as_uint(a.w);
            float t, u, v;
            bool ok;
            if(TRI == TRI_WATERTIGHT)
                ok = watertight_hit(wt, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(cc.x, cc.y, cc.z), tmin, h.t, &t, &u, &v);
            else
                ok = mt_hit(mt, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(cc.x, cc.y, cc.z), tmin, h.t, &t, &u, &v);
            if(ok && (t < h.t || gid < h.gid))
            {
                h.t = t;
                h.b1 = u;
                h.b2 = v;
                h.gid = gid;
                h.inst = TWO ? cur_inst : __float_as_uint(b.w);
                h.prim = TWO ? __float_as_uint(a.w) : __float_as_uint(cc.w);
                any_hit = true;
                if(ANY)
                    break;
            }
        }
        return any_hit;
    };

    for(;;)
    {
        // ---- hand new rays to idle lanes
        const uint64_t idle = __ballot(!active);
        const int nidle = (int)__popcll(idle);
        if(nidle >= REFILL_MIN && !exhausted)
        {
            // first the rest of the reserved granule, then, for lanes still
            // idle, the home chunk's next granule (every lane runs this
            // loop, so lane 0 takes it for the wavefront)
            const uint32_t base_a = res_lo;
            const uint32_t take_a = min((uint32_t)nidle, res_hi - res_lo);
            res_lo += take_a;
            uint32_t base_b = 0, take_b = 0;
            if(take_a < (uint32_t)nidle && !dry)
            {
                uint32_t g = 0;
                if(lane == 0)
                    g = atomicAdd(counter, 1u);
                g = __builtin_amdgcn_readfirstlane(g);
                const uint64_t lo = (uint64_t)static_span + ((uint64_t)g * chunks + home) * GRANULE;
                if(lo < (uint64_t)n)
                {
                    res_lo = (uint32_t)lo;
                    res_hi = min((uint32_t)n, res_lo + (uint32_t)GRANULE);
                    base_b = res_lo;
                    take_b = min((uint32_t)nidle - take_a, res_hi - res_lo);
                    res_lo += take_b;
                }
                else
                    dry = true;
            }
            exhausted = dry && res_lo == res_hi;
            if(!active)
            {
                // idle lane number `rank` takes ray base_a + rank, or ray
                // base_b + rank - take_a of the new granule
                const uint32_t rank = (uint32_t)__popcll(idle & lanes_below);
                bool got = false;
                float tmax = 0.0f;
                if(take_a > 0 && rank < take_a)
                {
                    ray = sink.load(base_a, base_a + rank, o, d, tmin, tmax);
                    got = true;
                }
                if(take_b > 0 && rank >= take_a && rank < take_a + take_b)
                {
                    ray = sink.load(base_b, base_b + (rank - take_a), o, d, tmin, tmax);
                    got = true;
                }
                if(got)
                {
                    set_ray(o, d);
                    margin = cull_margin(sc.extent, o);
                    cur_inst = 0xffffffffu;
                    h.t = tmax;
                    h.b1 = h.b2 = 0.0f;
                    h.gid = h.inst = h.prim = 0xffffffffu;
                    node = 0;
                    sp = 0;
                    active = true;
                }
            }
        }
        if(!__any(active))
            break;

        // ---- inner nodes: step until every lane holds a leaf or has finished
        uint32_t leaf = 0;
        for(;;)
        {
            const bool step = active && node != CODE_DONE && (leaf == 0 || !(node & CODE_LEAF));
            if(!__any(step))
                break;
            if(step)
            {
                if(TWO && node == CODE_EXIT_NEXT)
                    node = pop(false);
                else if(TWO && (node & CODE_LEAF) && (node & CODE_INST))
                {
                    // enter an instance: world -> local, continue in its tree
                    const TraceInstance &ti = sc.inst[node & 0x0fffffffu];
                    push(CODE_EXIT, -INFINITY);
                    const float *r = ti.w2l;
                    const V3 lo = v3(r[0] * o.x + r[1] * o.y + r[2] * o.z + r[3],
                                     r[4] * o.x + r[5] * o.y + r[6] * o.z + r[7],
                                     r[8] * o.x + r[9] * o.y + r[10] * o.z + r[11]);
                    const V3 ld = v3(r[0] * d.x + r[1] * d.y + r[2] * d.z,
                                     r[4] * d.x + r[5] * d.y + r[6] * d.z,
                                     r[8] * d.x + r[9] * d.y + r[10] * d.z);
                    set_ray(lo, ld);
                    cur_inst = node & 0x0fffffffu;
                    cur_base = ti.prim_base;
                    node = ti.blas_root;
                }
                else if(node & CODE_LEAF)
                {
                    if(POSTPONE)
                    {
                        leaf = node;
                        node = pop(TWO);
                    }
                    else
                        node = (test_leaf(node) && ANY) ? CODE_DONE : pop(false);
                }
                else
                {
                    // the node stores lo then hi (16 bytes each) per axis: the
                    // near and far planes are picked by address from the
                    // ray's direction signs, not by selects after the load
                    const char *nb = reinterpret_cast<const char *>(sc.nodes);
                    const uint32_t off = node * (uint32_t)sizeof(BvhNode4);
                    const float4 nxv = *reinterpret_cast<const float4 *>(nb + (off + near_x));
                    const float4 fxv = *reinterpret_cast<const float4 *>(nb + (off + (near_x ^ 16u)));
                    const float4 nyv = *reinterpret_cast<const float4 *>(nb + (off + near_y));
                    const float4 fyv = *reinterpret_cast<const float4 *>(nb + (off + (near_y ^ 16u)));
                    const float4 nzv = *reinterpret_cast<const float4 *>(nb + (off + near_z));
                    const float4 fzv = *reinterpret_cast<const float4 *>(nb + (off + (near_z ^ 16u)));
                    const uint4 code = *reinterpret_cast<const uint4 *>(nb + (off + 96u));
                    const float bnx[4] = { nxv.x, nxv.y, nxv.z, nxv.w };
                    const float bfx[4] = { fxv.x, fxv.y, fxv.z, fxv.w };
                    const float bny[4] = { nyv.x, nyv.y, nyv.z, nyv.w };
                    const float bfy[4] = { fyv.x, fyv.y, fyv.z, fyv.w };
                    const float bnz[4] = { nzv.x, nzv.y, nzv.z, nzv.w };
                    const float bfz[4] = { fzv.x, fzv.y, fzv.z, fzv.w };
                    const uint32_t cc[4] = { code.x, code.y, code.z, code.w };
                    const float cull = cull_distance(h.t, margin);
                    float k[4];
                    uint32_t c[4];
#pragma unroll
                    for(int j = 0; j < 4; ++j)
                    {
                        float te;
                        const bool hit = raybox_nf(rb, bnx[j], bfx[j], bny[j], bfy[j], bnz[j], bfz[j],
                                                   tmin, cull, &te) && cc[j] != CODE_DONE;
                        k[j] = hit ? te : INFINITY;
                        c[j] = hit ? cc[j] : CODE_DONE;
                    }
#define BTRC_CSWAP(a, b)                                                  \
                    if(k[b] < k[a])                                       \
                    {                                                     \
                        const float tk = k[a]; k[a] = k[b]; k[b] = tk;    \
                        const uint32_t tc = c[a]; c[a] = c[b]; c[b] = tc; \
                    }
                    BTRC_CSWAP(0, 1)
                    BTRC_CSWAP(2, 3)
                    BTRC_CSWAP(0, 2)
                    BTRC_CSWAP(1, 3)
                    BTRC_CSWAP(1, 2)
#undef BTRC_CSWAP
                    if(c[0] != CODE_DONE)
                    {
                        if(c[3] != CODE_DONE) push(c[3], k[3]);
                        if(c[2] != CODE_DONE) push(c[2], k[2]);
                        if(c[1] != CODE_DONE) push(c[1], k[1]);
                        node = c[0];
                    }
                    else
                        node = pop(TWO && leaf != 0);
                }
            }
            if(POSTPONE)
            {
                if(__all(!active || node == CODE_DONE || leaf != 0))
                    break;
            }
            else if(!exhausted && (int)__popcll(__ballot(!active || node == CODE_DONE)) >= REFILL_MIN)
                break; // enough lanes to refill

        }

        // ---- leaf: test its triangles
        if(POSTPONE && active && leaf != 0)
        {
            if(test_leaf(leaf) && ANY)
                node = CODE_DONE;
        }

        // ---- finished rays leave their result
        if(active && node == CODE_DONE)
        {
            sink.store(ray, h);
            active = false;
        }
    }
}

} // namespace btrc
