// Device data layout of the wavefront path tracer and the launch wrappers of
// its stages. Mirrors the reference's SOA pools (renderer/wavefront/soa.h:
// 8-147) with compact records, and replaces its per-bounce host readback of
// the queue counters (renderer/wavefront.cpp:283-284) with counters that stay
// in device memory: every stage reads its own work count on the GPU.
#pragma once

#include <hip/hip_runtime.h>

#include "../core/sampler.h"
#include "../core/scene.h"
#include "nrc_kernels.h"
#include "trace.h"

namespace btrc {

// Path state pool, one side of the ping-pong pair.
struct PathSOA
{
    float4 *ray_o;    // origin.xyz, medium id bits
    float4 *ray_d;    // direction.xyz, t1
    float4 *beta;     // throughput.xyz, depth bits
    float4 *rad;      // path radiance.xyz, training path id bits (cache mode) or ~0
    float4 *bsdf_le;  // beta_le.xyz, bsdf pdf (negative: delta lobe)
    uint4 *smp;       // pixel index, then Halton: dimension, index lo, hi;
                      // independent: PCG state lo, hi, unused
};

struct HitSOA
{
    uint32_t *flag;   // bit 30: intersected, bit 31: scattered, low 30 bits: instance
    float4 *tuv;      // t, b1, b2, prim bits
};

struct ShadowSOA
{
    float4 *o;        // origin.xyz, pixel index bits
    float4 *d;        // direction.xyz, t1
    float4 *beta;     // beta_li.xyz, medium id bits
    uint2 *seed;      // transmittance stream seed (scenes with media), else null
    uint32_t *train;  // training path id or ~0 (cache mode), else null
    float *train_acc; // radiance sums of the training paths (cache mode)
};


constexpr uint32_t PATH_FLAG_HAS_INTERSECTION = 1u << 30; // wavefront/common.h:12
constexpr uint32_t PATH_FLAG_HAS_SCATTERING = 1u << 31;
constexpr uint32_t PATH_FLAG_INSTANCE_ID_MASK = ~0u >> 2;

struct WaveCounters
{
    int32_t active;       // live states in the current pool side
    int32_t gen_count;    // states generated this iteration
    int32_t gen_offset;   // dense index of the first generated state
    int32_t pad0;
    uint64_t generated;   // paths generated so far
    uint64_t total;       // spp * width * height
    uint64_t gen_base;    // path index of the first generated state
    uint64_t ext_rays;    // extension rays traced
    uint64_t shadow_rays; // shadow rays traced
    uint64_t iterations;
};

// Queues without a shared counter. A bounce's stage kernels take their
// inputs by dense index j (0 <= j < n) and push at most one state and one
// shadow ray per input; input j pushes into stripe k = j / stripe (stripe:
// a multiple of 64, so a wavefront's inputs share one stripe, and
// COUNTER_STRIPES * stripe >= n) at slots k * stripe + 0, 1, ... from its
// own counter. A pool side therefore holds the states of the last bounce in
// COUNTER_STRIPES segments and, after them, the states generated for this
// bounce at slot pool_gen_slot(state_count); pool_slot turns dense
// indices into slots. The shadow rays of bounce b are traced beside the
// next bounce's plan, generate and trace (on another stream), so what they
// use is kept per bounce parity: the stripe size, the shadow stripe counters
// and the shadow fetch counters.
constexpr int POOL_SEGMENTS = COUNTER_STRIPES + 1;

struct QueueState
{
    int32_t next_n[COUNTER_STRIPES * COUNTER_STRIDE];      // states pushed into the other pool side, per stripe
    int32_t shadow_n[2][COUNTER_STRIPES * COUNTER_STRIDE]; // shadow rays pushed, per parity and stripe
    uint32_t trace_fetch[FETCH_WORDS];                     // extension-ray chunk counters
    uint32_t shadow_fetch[2][FETCH_WORDS];                 // shadow-ray chunk counters, per parity
    uint32_t medium_fetch[FETCH_WORDS];                    // medium-tracking chunk counters
    int32_t stripe[2];                                     // inputs per stripe of a bounce, per parity
    int32_t pool_first[POOL_SEGMENTS];                  // slot of each segment's first state
    int32_t pool_prefix[POOL_SEGMENTS + 1];             // dense index of each segment's first state
};

// the counters the host reads back, then the queue bookkeeping (device only)
struct WaveState
{
    WaveCounters c;
    alignas(128) QueueState q;
};

__host__ __device__ inline QueueState *queue_state(WaveCounters *c) { return &reinterpret_cast<WaveState *>(c)->q; }
__host__ __device__ inline const QueueState *queue_state(const WaveCounters *c)
{
    return &reinterpret_cast<const WaveState *>(c)->q;
}

// slots of a pool side and of the shadow queue for a path pool of state_count
__host__ __device__ inline int32_t pool_gen_slot(int32_t state_count) { return state_count + 64 * COUNTER_STRIPES; }
inline size_t pool_capacity(size_t state_count) { return 2 * state_count + 64 * COUNTER_STRIPES; }
inline size_t shadow_capacity(size_t state_count) { return state_count + 64 * COUNTER_STRIPES; }

// Slot of the state with dense index j of the current pool side, for lanes
// of one wavefront whose indices lie in [j0, j0 + 63] (j0 wave-uniform):
// the segment of j0 is found with scalar loads and compares, then the lanes
// past its end (segments that start inside the range: rare) are moved on.
__device__ __forceinline__ int32_t pool_slot(const QueueState *q, int32_t j0, int32_t j)
{
    j0 = __builtin_amdgcn_readfirstlane(j0);
    int32_t s = 0;
    for(int k = 1; k < POOL_SEGMENTS; ++k)
        if(q->pool_prefix[k] <= j0)
            s = k;
    int32_t slot = j - q->pool_prefix[s] + q->pool_first[s];
    for(int t = s + 1; t < POOL_SEGMENTS && q->pool_prefix[t] <= j0 + 63; ++t)
        if(j >= q->pool_prefix[t])
            slot = j - q->pool_prefix[t] + q->pool_first[t];
    return slot;
}

// the same for the shadow rays of the bounce of this parity, whose segments
// (stripe k at slot k * stripe, in stripe order) follow from the counters
__device__ __forceinline__ int32_t shadow_slot(const QueueState *q, int parity, int32_t j0, int32_t j)
{
    j0 = __builtin_amdgcn_readfirstlane(j0);
    const int32_t stripe = q->stripe[parity];
    int32_t slot = j, first = 0;
    for(int k = 0; k < COUNTER_STRIPES && first <= j0 + 63; ++k)
    {
        const int32_t count = q->shadow_n[parity][k * COUNTER_STRIDE];
        if(first + count > j0 && j >= first)
            slot = j - first + k * stripe;
        first += count;
    }
    return slot;
}

// the number of shadow rays pushed by the bounce of this parity
__device__ __forceinline__ int32_t shadow_count(const QueueState *q, int parity)
{
    int32_t n = 0;
    for(int k = 0; k < COUNTER_STRIPES; ++k)
        n += q->shadow_n[parity][k * COUNTER_STRIDE];
    return n;
}

struct CameraConsts
{
    float eye[3];
    float lbc[3];    // left_bottom_corner, pinhole.cpp:79
    float film_x[3];
    float film_y[3];
};

enum : int32_t
{
    SAMPLER_HALTON = 0,
    SAMPLER_INDEPENDENT = 1,
    FILTER_BOX = 0,
    FILTER_GAUSSIAN = 1,
};

// GaussianFilter's importance table (film_filter/gaussian.cpp:7-52)
constexpr int32_t GAUSSIAN_GRID = 256;
constexpr int32_t ENVIR_TABLE_SAMPLES = 256;       // IBL::commit's n_samples (light/ibl.cpp:21-25)
constexpr int64_t MAX_ENVIR_TABLE = 1 << 24;       // IBL importance table tiles

struct FilterConsts
{
    int32_t type;
    float radius;
    float cell;               // 2 * radius / GAUSSIAN_GRID
    const AliasUnit *table;   // GAUSSIAN_GRID^2 cells, row-major (y, x)
};

// One view of a render: its camera and film. A render walks the views'
// paths in order: view v owns path indices [path_base, path_base + width *
// height * spp) and film pixels [pixel_base, pixel_base + width * height).
struct ViewRec
{
    CameraConsts camera;
    HaltonConsts halton;
    int32_t width, height;
    uint32_t pixel_base;
    uint32_t pad;
    uint64_t path_base;
};

constexpr int32_t MAX_VIEWS = 64;

struct RenderConsts
{
    const ViewRec *views;
    int32_t view_count;
    int32_t spp;
    int32_t min_depth, max_depth;
    float rr_threshold, rr_cont_prob;
    int32_t state_count;
    uint32_t vol_prim_medium;
    int32_t film_fp64;
    int32_t tile;
    int32_t sample_offset;  // first sample index of this render
    // cache mode: paths whose pixel % nrc_stride == nrc_phase are training
    // paths with id pixel / nrc_stride while that is below nrc_max_paths
    int32_t nrc_stride, nrc_phase, nrc_max_paths;
    FilterConsts filter;
    HaltonTables halton;    // first-digit tables of the Halton sampler
};

// the view holding film pixel `pixel`
__device__ __forceinline__ const ViewRec &view_of_pixel(const RenderConsts &rc, uint32_t pixel)
{
    int32_t v = rc.view_count - 1;
    while(v > 0 && pixel < rc.views[v].pixel_base)
        --v;
    return rc.views[v];
}

struct FilmView
{
    float *rgb;      // 4 floats per pixel (radiance xyz, unused)
    double *rgb64;   // 3 doubles per pixel, or null
    float *albedo;   // 4 floats per pixel (first-hit albedo xyz), or null
    float *normal;   // 4 floats per pixel (first-hit shading normal xyz), or null
};

enum : int32_t
{
    RESOLVE_COLOR = 0,   // value / weight
    RESOLVE_NORMAL = 1,  // value / weight, normalised unless shorter than 1e-3
};

// kernel timing buckets
enum : int
{
    STAGE_GENERATE = 0,
    STAGE_TRACE = 1,
    STAGE_SHADE = 2,
    STAGE_SHADOW = 3,
    STAGE_MEDIUM = 4,
    STAGE_CACHE = 5,   // cache inference
    STAGE_TRAIN = 6,   // cache training
    STAGE_COUNT = 7,
};

// parity: the bounce's parity (the pool side it reads)
void launch_plan(hipStream_t s, WaveCounters *c, int32_t state_count, int parity);
void launch_generate(hipStream_t s, int grid, int sampler, const RenderConsts *rc, const WaveCounters *c,
                     PathSOA path);
void launch_trace_extension(hipStream_t s, int grid, int sched, const TraceScene &ts, WaveCounters *c,
                            PathSOA path, HitSOA hit, const TraceOverflow &ovf);
void launch_shade(hipStream_t s, int grid, int sampler, const RenderConsts *rc, const SceneView &sv,
                  WaveCounters *c, PathSOA path, HitSOA hit, PathSOA next, ShadowSOA shadow, FilmView film,
                  const NrcView &nrc, int parity);
// wavefronts per CU of the persistent medium-tracking kernel
constexpr int MEDIUM_TRACK_WAVES_PER_CU = 20;
// track_grid: wavefronts of the persistent tracking kernel; scatter: two
// float4 per dense index (the tracking results of scattering states)
void launch_medium(hipStream_t s, int grid, int track_grid, int sampler, const RenderConsts *rc,
                   const SceneView &sv, WaveCounters *c, PathSOA path, HitSOA hit, float4 *scatter, PathSOA next,
                   ShadowSOA shadow, FilmView film, int parity);
void launch_shadow(hipStream_t s, int grid, int sched, bool media, const TraceScene &ts, const SceneView &sv,
                   WaveCounters *c, ShadowSOA shadow, FilmView film, const TraceOverflow &ovf, int parity);
void launch_end_iteration(hipStream_t s, WaveCounters *c, int parity);
void launch_resolve(hipStream_t s, int mode, const float *value, int32_t pixels, float inv_weight, float4 *out);

// ray batches of the traversal entry points: o.xyz/tmin and d.xyz/tmax
void launch_trace_batch(hipStream_t s, int grid, int sched, const TraceScene &ts, int tri_test, bool any_hit,
                        const float4 *rays, int32_t n, uint32_t *out, uint32_t *fetch, const TraceOverflow &ovf);

// `vectors` 16-byte vectors from src to dst (STREAM-style bandwidth measurement)
void launch_stream_copy(hipStream_t s, int grid, const void *src, void *dst, int64_t vectors);
// test hooks
void launch_halton_eval(hipStream_t s, const HaltonConsts &hc, const HaltonTables &tab, const uint32_t *px,
                        const uint32_t *py,
                        const int32_t *sample_index, int32_t n, int32_t dims, float *out);
// one-level structures: world-space triangle records of `geometry`'s
// instances from its new local positions (9 floats per triangle)
void launch_update_world_tris(hipStream_t s, float4 *tris, uint32_t n, const InstanceRec *insts, int32_t geometry,
                              const float *positions);
// EnvirLightSampler::preprocess's table kernel: lum and area per tile (w * h each)
void launch_envir_table(hipStream_t s, const TextureView &tv, int32_t tex, int32_t w, int32_t h, int32_t n_samples,
                        float *lum, float *area);
// the environment light's eval_le, pdf_li and sample_li: 12 floats per query
void launch_envir_eval(hipStream_t s, const SceneView &sv, const float *dirs, const float *sam, int32_t n, float *out);
void launch_texture_eval(hipStream_t s, const TextureView &tv, int32_t id, const float2 *uv, int32_t n, float4 *out);
void launch_volume_eval(hipStream_t s, const VolumeView &vv, const float4 *seg, const uint2 *seed, int32_t n,
                        float *out);
void launch_medium_eval(hipStream_t s, const SceneView &sv, int32_t med, const float4 *seg, const uint2 *seed,
                        int32_t n, float *out);
void launch_independent_eval(hipStream_t s, const uint32_t *px, const uint32_t *py, const int32_t *sample_index,
                             int32_t n, int32_t dims, float *out);
void launch_detmath_eval(hipStream_t s, int32_t fn, const float *x, const float *y, int32_t n, float *out);
// ACESToneMap (postprocess/tonemap.cpp:14-24,35-51) in place on resolved RGBA
void launch_tonemap_aces(hipStream_t s, float4 *rgba, int32_t pixels, float exposure);
// p[0, n) = value, by a kernel (a memset node misbehaves in captured graphs)
void launch_fill_u32(hipStream_t s, uint32_t *p, uint32_t n, uint32_t value);

} // namespace btrc
