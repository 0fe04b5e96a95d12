// Host side of the C ABI (include/btrc_hip.h): scene commit, device memory,
// the wavefront render loop and the traversal entry points.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/btrc_hip.h"
#include "../kernels/bvh_gpu.h"
#include "../kernels/wavefront.h"
#include "bvh_build.h"

using namespace btrc;

// Scratch of btrc_hip_trace_closest_noalloc, for callers that cannot
// allocate: the batch's ray counters in the first NOALLOC_COUNTER_BYTES, then the
// traversal's per-lane stack entries beyond the LDS part (codes, then entry
// distances, each [entry][lane]). Every GPU the library runs on holds one
// copy until the process ends. Host code never uses the array's name as an
// address: it resolves the device address once per device
// (hipGetSymbolAddress, at context creation) and uses that.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)8 << 20;
constexpr size_t NOALLOC_COUNTER_BYTES = FETCH_WORDS * sizeof(uint32_t);
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

// device address of `scratch` per device, resolved once per device with that
// device current; null with a message when the lookup failed
struct ScratchAddress
{
    std::mutex lock;
    static constexpr int MAX_DEVICES = 64;
    bool resolved[MAX_DEVICES] = {};
    void *ptr[MAX_DEVICES] = {};
    std::string error[MAX_DEVICES];
};

ScratchAddress &scratch_address()
{
    static ScratchAddress a;
    return a;
}

// called by btrc_hip_create with `device` current
void resolve_scratch(int device, void **ptr, std::string *error)
{
    ScratchAddress &a = scratch_address();
    std::lock_guard<std::mutex> g(a.lock);
    if(device < 0 || device >= ScratchAddress::MAX_DEVICES)
    {
        *ptr = nullptr;
        *error = "no-allocation scratch: device index beyond the table";
        return;
    }
    if(!a.resolved[device])
    {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if(e != hipSuccess || !p)
        {
            (void)hipGetLastError();
            a.ptr[device] = nullptr;
            a.error[device] = std::string("no-allocation scratch: hipGetSymbolAddress failed: ") +
                              hipGetErrorString(e);
        }
        else
            a.ptr[device] = p;
        a.resolved[device] = true;
    }
    *ptr = a.ptr[device];
    *error = a.error[device];
}

struct HipError : std::runtime_error
{
    using std::runtime_error::runtime_error;
};
struct InvalidArg : std::runtime_error
{
    using std::runtime_error::runtime_error;
};
struct StateError : std::runtime_error
{
    using std::runtime_error::runtime_error;
};
struct Stopped : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

#define BTRC_HIP_CHECK(expr)                                                                        \
    do                                                                                              \
    {                                                                                               \
        hipError_t e_ = (expr);                                                                     \
        if(e_ != hipSuccess)                                                                        \
            throw HipError(std::string(#expr) + ": " + hipGetErrorString(e_));                      \
    } while(0)

// RAII device allocation
template<typename T>
struct DevBuf
{
    T *ptr = nullptr;
    size_t count = 0;

    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }

    void release()
    {
        if(ptr)
            (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    void alloc(size_t n)
    {
        if(n == count && ptr)
            return;
        release();
        if(n == 0)
            return;
        if(hipMalloc(&ptr, n * sizeof(T)) != hipSuccess)
        {
            ptr = nullptr;
            throw std::bad_alloc();
        }
        count = n;
    }
    void upload(const T *src, size_t n, hipStream_t s)
    {
        alloc(n);
        if(n)
            BTRC_HIP_CHECK(hipMemcpyAsync(ptr, src, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

struct PoolSide
{
    DevBuf<float4> ray_o, ray_d, beta, rad, bsdf_le;
    DevBuf<uint4> smp;

    void alloc(size_t n)
    {
        ray_o.alloc(n); ray_d.alloc(n); beta.alloc(n); rad.alloc(n); bsdf_le.alloc(n); smp.alloc(n);
    }
    PathSOA view()
    {
        return PathSOA{ ray_o.ptr, ray_d.ptr, beta.ptr, rad.ptr, bsdf_le.ptr, smp.ptr };
    }
};

int64_t multiplicative_inverse(int64_t a, int64_t n)
{
    // extended Euclid, sampler/halton.cpp:18-37
    int64_t old_r = a, r = n, old_s = 1, s = 0;
    while(r != 0)
    {
        const int64_t q = old_r / r;
        int64_t t = old_r - q * r; old_r = r; r = t;
        t = old_s - q * s; old_s = s; s = t;
    }
    int64_t x = old_s % n;
    return x < 0 ? x + n : x;
}

HaltonConsts make_halton_consts(int32_t width, int32_t height)
{
    // HaltonSampler::initialize_consts, sampler/halton.cpp:118-133
    HaltonConsts hc;
    const int32_t res[2] = { width, height };
    for(int i = 0; i < 2; ++i)
    {
        const int32_t base = i == 0 ? 2 : 3;
        int32_t scale = 1, exp = 0;
        while(scale < std::min(res[i], MAX_HALTON_RESOLUTION))
        {
            scale *= base;
            ++exp;
        }
        hc.base_scales[i] = scale;
        hc.base_exponents[i] = exp;
    }
    hc.mult_inverse[0] = (int32_t)multiplicative_inverse(hc.base_scales[1], hc.base_scales[0]);
    hc.mult_inverse[1] = (int32_t)multiplicative_inverse(hc.base_scales[0], hc.base_scales[1]);
    return hc;
}

Xform make_xform(const float *mat, const float *inv)
{
    Xform x;
    for(int k = 0; k < 12; ++k)
    {
        x.m[k] = mat[k];
        x.n[k] = inv[k];
    }
    return x;
}

// triangle_dpdu, utils/math/triangle.h:8-23
V3 triangle_dpdu(V3 B_A, V3 C_A, float bax, float bay, float cax, float cay, V3 nor)
{
    const float m00 = bax, m01 = bay;
    const float m10 = cax, m11 = cay;
    const float det = m00 * m11 - m01 * m10;
    if(det == 0.0f)
        return frame_from_z(nor).x;
    const float inv_det = 1.0f / det;
    return normalize(m11 * inv_det * B_A - m01 * inv_det * C_A);
}

} // namespace

struct btrc_ctx
{
    int device = 0;
    hipStream_t stream = nullptr;
    // the cache's training runs on its own stream, overlapping the next
    // frame's bounces; the next inference waits for it (train_done)
    hipStream_t train_stream = nullptr;
    hipEvent_t infer_done = nullptr, train_done = nullptr;
    // a bounce's shadow rays run on their own stream beside the next
    // bounce's first stages (inside the captured bounce loop)
    hipStream_t shadow_stream = nullptr;
    hipEvent_t shaded = nullptr, shadow_done = nullptr;
    bool train_pending = false;       // training enqueued and not yet waited for on the host
    std::string last_error;
    std::atomic<bool> stop{ false };
    int cu_count = 256;
    std::string device_name;
    // this device's copy of `scratch` (btrc_hip_trace_closest_noalloc)
    void *noalloc_scratch = nullptr;
    std::string noalloc_error;
    // progress callbacks of renders (Reporter, core/reporter.h:7-36)
    btrc_reporter reporter{ nullptr, nullptr, nullptr, nullptr };

    // scene
    bool has_scene = false;
    DevBuf<BvhNode4> nodes;
    DevBuf<float4> tris;
    DevBuf<TraceInstance> trace_inst;
    bool two_level = false;
    DevBuf<InstanceRec> instances;
    DevBuf<GeometryRec> geometries;
    DevBuf<MaterialRec> materials;
    DevBuf<LightRec> lights;
    DevBuf<float4> gx_u_a, gy_u_ba, gz_u_ca, sn_v_a, sn_v_ba, sn_v_ca, light_pos;
    DevBuf<AliasUnit> alias;
    DevBuf<MediumRec> media;
    DevBuf<float> voxels;
    DevBuf<VolumeRec> volumes;
    int32_t volume_count = 0;
    float world_lo[3] = { 0, 0, 0 }, world_hi[3] = { 0, 0, 0 }; // triangles and volume primitives
    DevBuf<TextureRec> textures;
    DevBuf<uint32_t> texels8;
    DevBuf<float4> texels32;
    DevBuf<float> tex_decode;
    int32_t texture_count = 0;
    DevBuf<float> env_probs;          // the IBL's tile probabilities
    DevBuf<AliasUnit> env_alias;      // and their alias table
    std::vector<float> env_probs_host;
    int32_t medium_count = 0;
    int32_t light_count = 0;
    int32_t envir_light_index = -1;
    int32_t instance_count = 0;
    int64_t triangle_count = 0;
    int64_t node_count = 0;
    double build_ms = 0.0;            // the last upload's acceleration-structure build
    double refit_ms = 0.0;            // the last btrc_hip_update_geometry
    PlocBuffers ploc;                 // GPU builder scratch, released after a scene build
    // geometry updates of one-level structures: per-geometry ranges, the
    // 4-wide nodes level by level (root level first), the bounds of the
    // triangles and of the volume primitives
    std::vector<GeometryRec> geos_host;
    DevBuf<uint32_t> refit_order;
    std::vector<uint32_t> refit_levels;
    float tri_lo[3] = { 0, 0, 0 }, tri_hi[3] = { 0, 0, 0 };
    float vol_lo[3] = { 0, 0, 0 }, vol_hi[3] = { 0, 0, 0 };
    int32_t bvh_depth = 0;
    int32_t stack_size = 0;

    // camera
    bool has_camera = false;
    CameraConsts camera{};

    // render state
    PoolSide pool[2];
    size_t pool_size = 0;
    DevBuf<uint32_t> hit_flag;
    DevBuf<float4> hit_tuv;
    DevBuf<float4> sh_o, sh_d, sh_beta;
    DevBuf<uint2> sh_seed;
    DevBuf<float4> medium_scatter; // scenes with media: tracking results, two per dense index
    DevBuf<uint32_t> sh_train;
    DevBuf<WaveState> counters;       // the counters (read back) and the queue bookkeeping
    WaveCounters *wc() const { return &counters.ptr->c; }
    WaveCounters *host_counters = nullptr; // pinned
    DevBuf<float> film;
    DevBuf<double> film64;
    DevBuf<float> film_albedo, film_normal;
    DevBuf<float4> resolved;
    int64_t film_pixels = 0;            // all views of the last render
    std::vector<ViewRec> film_views;    // their layout
    DevBuf<ViewRec> views;
    int32_t film_spp = 0;
    bool film_has64 = false, film_has_albedo = false, film_has_normal = false;

    // first-digit tables of the Halton sampler (built on first use)
    DevBuf<uint16_t> halton_values;
    DevBuf<HaltonDim> halton_dims;

    // film filter
    FilterConsts filter{ FILTER_BOX, 0.0f, 0.0f, nullptr };
    DevBuf<AliasUnit> filter_table;

    // world box of the scene's triangles (the cache's position frame)
    float scene_lo[3] = { 0, 0, 0 }, scene_hi[3] = { 1, 1, 1 };

    // neural radiance cache
    struct Nrc
    {
        btrc_nrc_params params{ 0, 1, 2, 16, 4, 1e-2f, 128.0f, 0, 0.0f };
        bool ready = false;             // weights allocated and initialised
        DevBuf<float> w, m, v, grad;
        DevBuf<uint16_t> wb, wtb;
        DevBuf<float> ema;              // moving average of the weights, and its BF16 copy
        DevBuf<uint16_t> ema_b;
        DevBuf<NrcQuery> queries;
        // training records and path radiance sums, two sets: a frame fills
        // one while the previous frame's training reads the other
        DevBuf<NrcRecord> rec[2];
        DevBuf<float> acc[2];
        int set = 0;                    // the set of the last frame
        int pending_set = -1;           // the set the pending training reads
        int32_t pending_steps = 0;      // steps whose statistics are still to be read
        bool stats_pending = false;
        DevBuf<int32_t> counts;         // [0] queries, [1] valid records of a batch, [2] valid records of the frame
        DevBuf<int32_t> index;          // the frame's valid record slots, compacted
        DevBuf<uint64_t> valid_mask;    // their flags in bijection order, one bit each
        DevBuf<uint16_t> xt, dzt;
        DevBuf<double> loss_partial;
        DevBuf<int32_t> count_partial;
        DevBuf<float> grad_partial, step_scale;
        DevBuf<double> g2_partial;
        DevBuf<NrcStepStats> step_stats;
        int64_t steps_total = 0;
        btrc_nrc_stats last{};
    } nrc;

    // traversal: per-lane stack overflow and the batch fetch counter
    DevBuf<uint32_t> ovf_code;
    DevBuf<float> ovf_t;
    int32_t trace_grid = 0;
    DevBuf<uint32_t> fetch;
    btrc_tuning tuning{ 0, 0, TRACE_WAVES_PER_CU, 1, 0, 0, 0, 0 };

    // the bounce loop as a captured graph, kept while its launches stay the same
    RenderConsts *rc_host = nullptr;  // pinned
    DevBuf<RenderConsts> rc_dev;
    // (two: cache frames alternate between two record sets)
    hipGraphExec_t loop_exec[2] = { nullptr, nullptr };
    std::vector<uint8_t> loop_key[2];
    int loop_next = 0;                // the entry the next capture replaces
    void drop_loop_graph()
    {
        for(int k = 0; k < 2; ++k)
        {
            if(loop_exec[k])
                (void)hipGraphExecDestroy(loop_exec[k]);
            loop_exec[k] = nullptr;
            loop_key[k].clear();
        }
    }

    // resident ray batch
    DevBuf<float4> batch_rays;
    DevBuf<uint32_t> batch_out;
    int64_t batch_n = 0;

    float scene_extent = 0.0f; // largest |coordinate| of the world-space triangles
    TraceScene trace_scene() const { return TraceScene{ nodes.ptr, tris.ptr, trace_inst.ptr, scene_extent }; }
    // traversal schedule of closest-hit (any = false) or any-hit rays
    int trace_sched(bool any) const
    {
        if(two_level)
            return SCHED_TWO_LEVEL | ((tuning.two_level_postpone >> (any ? 1 : 0)) & 1 ? SCHED_POSTPONE : 0);
        const bool postpone = (any ? tuning.any_postpone_leaves : tuning.closest_postpone_leaves) != 0;
        return postpone ? SCHED_POSTPONE : SCHED_EAGER;
    }
    TraceOverflow trace_overflow() const
    {
        return TraceOverflow{ ovf_code.ptr, ovf_t.ptr, trace_grid * TRACE_BLOCK };
    }
    SceneView scene_view() const
    {
        SceneView v;
        std::memset(&v, 0, sizeof(v)); // padding too: the view is compared bytewise
        v.instances = instances.ptr;
        v.geometries = geometries.ptr;
        v.materials = materials.ptr;
        v.lights = lights.ptr;
        v.gx_u_a = gx_u_a.ptr; v.gy_u_ba = gy_u_ba.ptr; v.gz_u_ca = gz_u_ca.ptr;
        v.sn_v_a = sn_v_a.ptr; v.sn_v_ba = sn_v_ba.ptr; v.sn_v_ca = sn_v_ca.ptr;
        v.light_pos = light_pos.ptr;
        v.alias = alias.ptr;
        v.light_count = light_count;
        v.envir_light_index = envir_light_index;
        v.instance_count = instance_count;
        v.medium_count = medium_count;
        v.media = media.ptr;
        v.voxels = voxels.ptr;
        v.tex = TextureView{ textures.ptr, texels8.ptr, texels32.ptr, tex_decode.ptr };
        v.env_probs = env_probs.ptr;
        v.env_alias = env_alias.ptr;
        v.vol = VolumeView{ volumes.ptr, volume_count, voxels.ptr };
        v.world_diagonal = 0.0f; // set per render from the cameras
        return v;
    }
    int persistent_grid(int threads_per_block, int64_t work) const
    {
        const int64_t want = (work + threads_per_block - 1) / threads_per_block;
        const int64_t cap = (int64_t)cu_count * (tuning.stage_blocks_per_cu > 0 ? tuning.stage_blocks_per_cu : 32);
        return (int)std::max<int64_t>(1, std::min(want, cap));
    }
};

namespace {

template<typename F>
int guarded(btrc_ctx *ctx, F &&f)
{
    if(!ctx)
        return BTRC_ERR_INVALID;
    try
    {
        f();
        ctx->last_error.clear();
        return BTRC_OK;
    }
    catch(const InvalidArg &e)
    {
        ctx->last_error = e.what();
        return BTRC_ERR_INVALID;
    }
    catch(const StateError &e)
    {
        ctx->last_error = e.what();
        return BTRC_ERR_STATE;
    }
    catch(const Stopped &e)
    {
        ctx->last_error = e.what();
        return BTRC_ERR_STOPPED;
    }
    catch(const HipError &e)
    {
        ctx->last_error = e.what();
        return BTRC_ERR_HIP;
    }
    catch(const std::bad_alloc &)
    {
        ctx->last_error = "out of memory";
        return BTRC_ERR_OOM;
    }
    catch(const std::exception &e)
    {
        ctx->last_error = e.what();
        return BTRC_ERR_INVALID;
    }
}

// AliasTable::initialize (utils/math/alias.cpp:13-58). The sum runs in index
// order (the reference's std::reduce leaves the order to the library).
// Returns the sum of the weights.
float build_alias_table(const std::vector<float> &w, std::vector<AliasUnit> &units)
{
    float sum = 0.0f;
    for(float a : w)
        sum += a;
    const size_t n = w.size();
    units.assign(n, AliasUnit{});
    if(n == 0)
        return sum;
    const float ratio = (float)n / sum;
    std::vector<uint32_t> overs, unders;
    for(uint32_t i = 0; i < (uint32_t)n; ++i)
    {
        const float pr = w[i] * ratio;
        units[i].accept_prob = pr;
        units[i].another_idx = i;
        if(pr > 1)
            overs.push_back(i);
        else if(pr < 1)
            unders.push_back(i);
    }
    while(!overs.empty() && !unders.empty())
    {
        const uint32_t over = overs.back();
        const uint32_t under = unders.back();
        overs.pop_back();
        unders.pop_back();
        units[over].accept_prob -= 1 - units[under].accept_prob;
        units[under].another_idx = over;
        if(units[over].accept_prob > 1)
            overs.push_back(over);
        else if(units[over].accept_prob < 1)
            unders.push_back(over);
    }
    for(auto i : overs)
        units[i].accept_prob = 1;
    for(auto i : unders)
        units[i].accept_prob = 1;
    return sum;
}

// EnvirLightSampler::preprocess (light/env_sampler.cpp:72-122): per tile the
// mean luminance L and solid angle A (a kernel over the uploaded texture),
// then p = max(0, A (L - mean L over the sphere)), normalized when the sum
// exceeds 1e-3, and an alias table over p. A texture of one luminance leaves
// every p at zero; the alias table then degenerates as the reference's does
// (NaN acceptance, so every tile samples itself) and the light's pdf is 0.
void build_envir_table(btrc_ctx *ctx, const LightRec &l)
{
    const int32_t w = l.lut_w, h = l.lut_h, n = w * h;
    DevBuf<float> dlum, darea;
    dlum.alloc(n);
    darea.alloc(n);
    launch_envir_table(ctx->stream, ctx->scene_view().tex, l.texture, w, h, ENVIR_TABLE_SAMPLES, dlum.ptr, darea.ptr);
    BTRC_HIP_CHECK(hipGetLastError());
    std::vector<float> lum(n), area(n);
    BTRC_HIP_CHECK(hipMemcpyAsync(lum.data(), dlum.ptr, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    BTRC_HIP_CHECK(hipMemcpyAsync(area.data(), darea.ptr, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    float lum_area_sum = 0.0f, area_sum = 0.0f;
    for(int32_t i = 0; i < n; ++i)
    {
        lum_area_sum += lum[i] * area[i];
        area_sum += area[i];
    }
    const float avg_lum = lum_area_sum / area_sum;
    for(int32_t i = 0; i < n; ++i)
    {
        const float p = area[i] * (lum[i] - avg_lum);
        lum[i] = 0.0f < p ? p : 0.0f;
    }
    float table_sum = 0.0f;
    for(float p : lum)
        table_sum += p;
    if(table_sum > 1e-3f)
    {
        const float ratio = 1 / table_sum;
        for(float &p : lum)
            p *= ratio;
    }
    std::vector<AliasUnit> units;
    build_alias_table(lum, units);
    ctx->env_probs.upload(lum.data(), lum.size(), ctx->stream);
    ctx->env_alias.upload(units.data(), units.size(), ctx->stream);
    BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->env_probs_host = std::move(lum);
}

// GaussianFilter::commit (film_filter/gaussian.cpp:27-52): separable
// truncated Gaussian on a GAUSSIAN_GRID^2 grid over [-r, r]^2, zero outside
// the disc of radius r, turned into an alias table
void set_filter(btrc_ctx *ctx, const btrc_filter_desc *f)
{
    if(!f)
        throw InvalidArg("filter is null");
    if(f->type == BTRC_FILTER_BOX)
    {
        ctx->filter = FilterConsts{ FILTER_BOX, 0.0f, 0.0f, nullptr };
        return;
    }
    if(f->type != BTRC_FILTER_GAUSSIAN)
        throw InvalidArg("unknown filter type");
    if(!(f->radius > 0.0f) || !std::isfinite(f->radius) || !std::isfinite(f->alpha))
        throw InvalidArg("gaussian filter needs a finite positive radius and a finite alpha");
    const float r = f->radius, alpha = f->alpha;
    const float expv = std::exp(-alpha * r * r);
    auto g = [&](float d) { return std::max(0.0f, std::exp(-alpha * d * d) - expv); };
    const int G = GAUSSIAN_GRID;
    std::vector<float> value((size_t)G * G, 0.0f);
    for(int y = 0; y < G; ++y)
    {
        const float yf = 2 * r * (y + 0.5f) / G - r;
        for(int x = 0; x < G; ++x)
        {
            const float xf = 2 * r * (x + 0.5f) / G - r;
            if(xf * xf + yf * yf < r * r)
                value[(size_t)y * G + x] = g(xf) * g(yf);
        }
    }
    std::vector<AliasUnit> units;
    const float sum = build_alias_table(value, units);
    if(!(sum > 0.0f))
        throw InvalidArg("gaussian filter has no weight (alpha too large for the radius)");
    BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->filter_table.upload(units.data(), units.size(), ctx->stream);
    BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->filter = FilterConsts{ FILTER_GAUSSIAN, r, 2 * r / G, ctx->filter_table.ptr };
}

// persistent traversal grid and the per-lane stack overflow it needs
void size_traversal(btrc_ctx *ctx)
{
    ctx->trace_grid = ctx->cu_count * ctx->tuning.trace_waves_per_cu;
    const size_t deep = ctx->stack_size > SHORT_STACK ? (size_t)(ctx->stack_size - SHORT_STACK) : 0;
    ctx->ovf_code.alloc(std::max<size_t>(1, deep * ctx->trace_grid * TRACE_BLOCK));
    ctx->ovf_t.alloc(std::max<size_t>(1, deep * ctx->trace_grid * TRACE_BLOCK));
    ctx->fetch.alloc(FETCH_WORDS);
}

// device layout of the traversal structure
struct Accel
{
    std::vector<BvhNode4> nodes;     // child slots hold traversal codes
    std::vector<float4> tris;        // 3 per triangle in leaf order
    std::vector<TraceInstance> inst; // two levels only
    bool on_device = false;          // built on the GPU straight into the context's buffers
    int64_t device_nodes = 0;
    // one-level trees: node indices grouped by depth (root first) and the
    // group boundaries; empty order: the indices are already in that order
    std::vector<uint32_t> level_order;
    std::vector<uint32_t> level_offsets;
    int depth = 0;
    int stack_bound = 0;
    float lo[3] = { MAX_FLOAT, MAX_FLOAT, MAX_FLOAT }, hi[3] = { -MAX_FLOAT, -MAX_FLOAT, -MAX_FLOAT };
};

// child slots as traversal codes: inner index (+ node_base), leaf code of
// triangles from tri_base, or CODE_DONE
void append_nodes(Accel &acc, const Bvh4 &bvh, uint32_t tri_base)
{
    const uint32_t node_base = (uint32_t)acc.nodes.size();
    for(BvhNode4 nd : bvh.nodes)
    {
        for(int j = 0; j < 4; ++j)
        {
            uint32_t code;
            if(nd.child[j] >= 0)
                code = node_base + (uint32_t)nd.child[j];
            else if(nd.count[j] > 0)
                code = CODE_LEAF | ((uint32_t)(nd.count[j] - 1) << 28) | (tri_base + (uint32_t)~nd.child[j]);
            else
                code = CODE_DONE;
            nd.child[j] = (int32_t)code;
        }
        acc.nodes.push_back(nd);
    }
}

// One tree over the world-space triangles; a triangle's record carries its
// flattened index, instance and primitive.
Accel build_flat(const std::vector<float> &world, const std::vector<uint32_t> &tri_inst,
                 const std::vector<uint32_t> &tri_prim, uint32_t ntri)
{
    if(ntri > MAX_LEAF_FIRST)
        throw InvalidArg("too many triangles for the traversal's leaf encoding");
    BvhBuildOptions opt;
    opt.max_leaf_size = 4; // leaf codes carry counts 1..4
    const Bvh4 bvh = build_bvh4(world.data(), ntri, opt);
    Accel acc;
    acc.tris.resize(3 * (size_t)ntri);
    for(uint32_t k = 0; k < ntri; ++k)
    {
        const uint32_t gid = bvh.prim_order[k];
        const float *w = world.data() + 9 * (size_t)gid;
        acc.tris[3 * k + 0] = make_float4(w[0], w[1], w[2], u2f(gid));
        acc.tris[3 * k + 1] = make_float4(w[3], w[4], w[5], u2f(tri_inst[gid]));
        acc.tris[3 * k + 2] = make_float4(w[6], w[7], w[8], u2f(tri_prim[gid]));
    }
    append_nodes(acc, bvh, 0);
    acc.depth = bvh.depth;
    acc.stack_bound = bvh.stack_bound;
    // breadth-first levels of the 4-wide nodes, for refits
    acc.level_order.push_back(0);
    acc.level_offsets = { 0, 1 };
    while(acc.level_offsets.back() > acc.level_offsets[acc.level_offsets.size() - 2])
    {
        const uint32_t lo = acc.level_offsets[acc.level_offsets.size() - 2], hi = acc.level_offsets.back();
        for(uint32_t k = lo; k < hi; ++k)
            for(int j = 0; j < 4; ++j)
            {
                const uint32_t code = (uint32_t)acc.nodes[acc.level_order[k]].child[j];
                if(code != CODE_DONE && !(code & CODE_LEAF))
                    acc.level_order.push_back(code);
            }
        acc.level_offsets.push_back((uint32_t)acc.level_order.size());
    }
    acc.level_offsets.pop_back();
    for(size_t k = 0; k < world.size(); ++k)
    {
        acc.lo[k % 3] = std::min(acc.lo[k % 3], world[k]);
        acc.hi[k % 3] = std::max(acc.hi[k % 3], world[k]);
    }
    return acc;
}

// One tree over the world-space triangles built on the GPU: PLOC over the
// triangle boxes, leaves collapsed by SAH up to four triangles (the leaf
// code's limit), 4-wide nodes, leaf-ordered triangle records; nodes and
// records land in the context's buffers directly.
Accel build_flat_gpu(btrc_ctx *ctx, const std::vector<float> &world, const std::vector<uint32_t> &tri_inst,
                     const std::vector<uint32_t> &tri_prim, uint32_t ntri)
{
    if(ntri > MAX_LEAF_FIRST)
        throw InvalidArg("too many triangles for the traversal's leaf encoding");
    hipStream_t s = ctx->stream;
    DevBuf<float> dv, dbox;
    DevBuf<uint32_t> dinst, dprim;
    dv.upload(world.data(), world.size(), s);
    dinst.upload(tri_inst.data(), tri_inst.size(), s);
    dprim.upload(tri_prim.data(), tri_prim.size(), s);
    dbox.alloc(6 * (size_t)ntri);
    triangle_boxes(dv.ptr, ntri, dbox.ptr, s);
    const PlocResult r = ploc_build(ctx->ploc, dbox.ptr, ntri, 4, s);
    const Bvh4Device b4 = ploc_to_bvh4(ctx->ploc, r, s);
    ctx->nodes.alloc(b4.node_count);
    BTRC_HIP_CHECK(hipMemcpyAsync(ctx->nodes.ptr, ctx->ploc.nodes4, b4.node_count * sizeof(BvhNode4),
                                  hipMemcpyDeviceToDevice, s));
    ctx->tris.alloc(3 * (size_t)ntri);
    gather_triangles(dv.ptr, dinst.ptr, dprim.ptr, ctx->ploc.prims, ntri, ctx->tris.ptr, s);
    BTRC_HIP_CHECK(hipGetLastError());
    BTRC_HIP_CHECK(hipStreamSynchronize(s));
    ctx->ploc.release(); // the builder's scratch is ~250 bytes per triangle; refits do not need it
    Accel acc;
    acc.on_device = true;
    acc.device_nodes = b4.node_count;
    acc.level_offsets = b4.level_offsets;
    acc.depth = b4.depth;
    acc.stack_bound = b4.stack_bound;
    for(size_t k = 0; k < world.size(); ++k)
    {
        acc.lo[k % 3] = std::min(acc.lo[k % 3], world[k]);
        acc.hi[k % 3] = std::max(acc.hi[k % 3], world[k]);
    }
    return acc;
}

// Relative padding of the instance boxes: a ray moved into an instance's
// local space carries rounding of order 2^-24 times the transform's
// condition and the coordinates' magnitude; 2^-14 of the largest world
// coordinate keeps the instance tree conservative for transforms with a
// condition up to about 2^8 and ray origins within a few scene extents.
constexpr float INSTANCE_BOX_PAD = 1.0f / 16384.0f;

// The instance tree (node 0) over one tree per geometry: geometry trees
// hold local-space triangles tagged with their primitive index, instance
// leaves name a TraceInstance (world_to_local rows, geometry root, first
// flattened index). The stack bound adds the instance tree's, the exit
// marker and the deepest geometry tree's.
Accel build_two_level(const btrc_scene_desc *sd, const std::vector<InstanceRec> &insts)
{
    const int32_t G = sd->geometry_count;
    const int32_t I = sd->instance_count;
    if((uint32_t)I > 0x0fffffffu)
        throw InvalidArg("too many instances for the instance tree's leaf encoding");
    uint64_t total = 0;
    for(int32_t gi = 0; gi < G; ++gi)
        total += (uint64_t)sd->geometries[gi].triangle_count;
    if(total > MAX_LEAF_FIRST)
        throw InvalidArg("too many geometry triangles for the traversal's leaf encoding");

    // local bounds of each geometry
    std::vector<float> gbox(6 * (size_t)G);
    for(int32_t gi = 0; gi < G; ++gi)
    {
        const btrc_geometry_desc &g = sd->geometries[gi];
        float *b = &gbox[6 * (size_t)gi];
        b[0] = b[1] = b[2] = MAX_FLOAT;
        b[3] = b[4] = b[5] = -MAX_FLOAT;
        for(size_t k = 0; k < 9 * (size_t)g.triangle_count; ++k)
        {
            b[k % 3] = std::min(b[k % 3], g.positions[k]);
            b[3 + k % 3] = std::max(b[3 + k % 3], g.positions[k]);
        }
    }
    // world bounds of each instance: its geometry box's corners, transformed
    Accel acc;
    std::vector<float> ibox(6 * (size_t)I);
    float mag = 0.0f;
    for(int32_t i = 0; i < I; ++i)
    {
        const float *b = &gbox[6 * (size_t)insts[i].geometry_id];
        float *w = &ibox[6 * (size_t)i];
        w[0] = w[1] = w[2] = MAX_FLOAT;
        w[3] = w[4] = w[5] = -MAX_FLOAT;
        if(!(b[0] <= b[3]))
            continue; // no triangles: never hit
        for(int c = 0; c < 8; ++c)
        {
            const V3 p = xf_point(insts[i].xf, v3(b[(c & 1) ? 3 : 0], b[(c & 2) ? 4 : 1], b[(c & 4) ? 5 : 2]));
            const float pc[3] = { p.x, p.y, p.z };
            for(int a = 0; a < 3; ++a)
            {
                w[a] = std::min(w[a], pc[a]);
                w[3 + a] = std::max(w[3 + a], pc[a]);
            }
        }
        for(int a = 0; a < 6; ++a)
            mag = std::max(mag, std::fabs(w[a]));
        // the scene box over the transformed vertices, as one level has it
        const btrc_geometry_desc &g = sd->geometries[insts[i].geometry_id];
        for(size_t k = 0; k < 3 * (size_t)g.triangle_count; ++k)
        {
            const float *q = g.positions + 3 * k;
            const V3 p = xf_point(insts[i].xf, v3(q[0], q[1], q[2]));
            acc.lo[0] = std::min(acc.lo[0], p.x); acc.hi[0] = std::max(acc.hi[0], p.x);
            acc.lo[1] = std::min(acc.lo[1], p.y); acc.hi[1] = std::max(acc.hi[1], p.y);
            acc.lo[2] = std::min(acc.lo[2], p.z); acc.hi[2] = std::max(acc.hi[2], p.z);
        }
    }
    const float pad = mag * INSTANCE_BOX_PAD;
    for(int32_t i = 0; i < I; ++i)
    {
        float *w = &ibox[6 * (size_t)i];
        if(w[0] <= w[3])
            for(int a = 0; a < 3; ++a)
            {
                w[a] -= pad;
                w[3 + a] += pad;
            }
    }

    BvhBuildOptions topt;
    topt.max_leaf_size = 1; // a leaf is one instance
    const Bvh4 tlas = build_bvh4_boxes(ibox.data(), (uint32_t)I, topt);
    append_nodes(acc, tlas, 0);
    for(BvhNode4 &nd : acc.nodes)
        for(int j = 0; j < 4; ++j)
        {
            const uint32_t code = (uint32_t)nd.child[j];
            if(code != CODE_DONE && (code & CODE_LEAF))
                nd.child[j] = (int32_t)(CODE_LEAF | CODE_INST | tlas.prim_order[code & 0x0fffffffu]);
        }

    // one tree per geometry, local-space triangles
    std::vector<uint32_t> groot(G);
    int blas_stack = 0, blas_depth = 0;
    BvhBuildOptions bopt;
    bopt.max_leaf_size = 4;
    for(int32_t gi = 0; gi < G; ++gi)
    {
        const btrc_geometry_desc &g = sd->geometries[gi];
        const uint32_t n = (uint32_t)g.triangle_count;
        const Bvh4 blas = build_bvh4(g.positions, n, bopt);
        const uint32_t tri_base = (uint32_t)(acc.tris.size() / 3);
        groot[gi] = (uint32_t)acc.nodes.size();
        append_nodes(acc, blas, tri_base);
        for(uint32_t k = 0; k < n; ++k)
        {
            const uint32_t prim = blas.prim_order[k];
            const float *p = g.positions + 9 * (size_t)prim;
            acc.tris.push_back(make_float4(p[0], p[1], p[2], u2f(prim)));
            acc.tris.push_back(make_float4(p[3], p[4], p[5], 0.0f));
            acc.tris.push_back(make_float4(p[6], p[7], p[8], 0.0f));
        }
        blas_stack = std::max(blas_stack, blas.stack_bound);
        blas_depth = std::max(blas_depth, blas.depth);
    }
    if(acc.nodes.size() >= (size_t)CODE_INST)
        throw InvalidArg("too many tree nodes for the traversal's node encoding");
    acc.inst.resize(I);
    for(int32_t i = 0; i < I; ++i)
    {
        TraceInstance &t = acc.inst[i];
        std::memcpy(t.w2l, insts[i].xf.n, sizeof(t.w2l));
        t.blas_root = groot[insts[i].geometry_id];
        t.prim_base = insts[i].prim_base;
        t.pad[0] = t.pad[1] = 0;
    }
    acc.depth = tlas.depth + blas_depth;
    acc.stack_bound = tlas.stack_bound + 1 + blas_stack;
    return acc;
}

// TriangleMesh::commit's per-triangle records of one geometry
// (triangle_mesh.cpp:24-171), appended to r; its area-weighted alias table
// and total area as well
struct GeoRecords
{
    std::vector<float4> gx, gy, gz, sa, sba, sca, lpos;
    std::vector<AliasUnit> alias;
};

float geometry_records(const btrc_geometry_desc &g, GeoRecords &r)
{
    std::vector<float> areas(g.triangle_count);
    for(int32_t t = 0; t < g.triangle_count; ++t)
    {
        const float *p = g.positions + 9 * (size_t)t;
        const V3 a = v3(p[0], p[1], p[2]), b = v3(p[3], p[4], p[5]), c = v3(p[6], p[7], p[8]);
        const V3 geo_z = normalize(cross(b - a, c - a));
        V3 sz[3];
        for(int k = 0; k < 3; ++k)
        {
            if(g.normals)
            {
                const float *n = g.normals + 9 * (size_t)t + 3 * k;
                sz[k] = v3(n[0], n[1], n[2]);
                if(sz[k].x == 0.0f && sz[k].y == 0.0f && sz[k].z == 0.0f)
                    sz[k] = geo_z;
            }
            else
                sz[k] = geo_z;
        }
        float uv[6] = { 0, 0, 0, 0, 0, 0 };
        if(g.tex_coords)
            std::memcpy(uv, g.tex_coords + 6 * (size_t)t, sizeof(uv));
        const V3 ex = triangle_dpdu(b - a, c - a, uv[2] - uv[0], uv[3] - uv[1], uv[4] - uv[0], uv[5] - uv[1], geo_z);
        const V3 ey = normalize(cross(geo_z, ex));
        r.gx.push_back(make_float4(ex.x, ex.y, ex.z, uv[0]));
        r.gy.push_back(make_float4(ey.x, ey.y, ey.z, uv[2] - uv[0]));
        r.gz.push_back(make_float4(geo_z.x, geo_z.y, geo_z.z, uv[4] - uv[0]));
        r.sa.push_back(make_float4(sz[0].x, sz[0].y, sz[0].z, uv[1]));
        const V3 dba = sz[1] - sz[0], dca = sz[2] - sz[0];
        r.sba.push_back(make_float4(dba.x, dba.y, dba.z, uv[3] - uv[1]));
        r.sca.push_back(make_float4(dca.x, dca.y, dca.z, uv[5] - uv[1]));
        const V3 ba = b - a, ca = c - a;
        r.lpos.push_back(make_float4(a.x, a.y, a.z, 0.0f));
        r.lpos.push_back(make_float4(ba.x, ba.y, ba.z, 0.0f));
        r.lpos.push_back(make_float4(ca.x, ca.y, ca.z, 0.0f));
        areas[t] = 0.5f * length(cross(ba, ca));
    }
    std::vector<AliasUnit> units;
    const float total_area = build_alias_table(areas, units);
    r.alias.insert(r.alias.end(), units.begin(), units.end());
    return total_area;
}

// Scene::commit's bounding box (core/scene.cpp:225-237) over the triangles
// and the volume primitives, the triangles' largest coordinate (the
// traversal's culling margin) and the triangles' box for the radiance cache
void set_world_bounds(btrc_ctx *ctx)
{
    for(int a = 0; a < 3; ++a)
    {
        ctx->world_lo[a] = std::min(ctx->tri_lo[a], ctx->vol_lo[a]);
        ctx->world_hi[a] = std::max(ctx->tri_hi[a], ctx->vol_hi[a]);
    }
    ctx->scene_extent = 0.0f;
    for(int a = 0; a < 3; ++a)
        if(ctx->tri_lo[a] <= ctx->tri_hi[a])
            ctx->scene_extent = std::max(ctx->scene_extent, std::max(std::fabs(ctx->tri_lo[a]), std::fabs(ctx->tri_hi[a])));
    for(int a = 0; a < 3; ++a)
    {
        float lo = ctx->tri_lo[a], hi = ctx->tri_hi[a];
        if(!(hi > lo))
        {
            lo = lo <= hi ? lo : 0.0f;
            hi = lo + 1.0f;
        }
        ctx->scene_lo[a] = lo;
        ctx->scene_hi[a] = hi;
    }
}

void upload_scene(btrc_ctx *ctx, const btrc_scene_desc *sd)
{
    if(!sd)
        throw InvalidArg("scene is null");
    if(sd->geometry_count < 0 || sd->material_count < 0 || sd->instance_count < 0)
        throw InvalidArg("negative element count");
    if(sd->instance_count > 0 && (!sd->instances || !sd->geometries || !sd->materials))
        throw InvalidArg("missing scene arrays");
    if((uint32_t)sd->instance_count > PATH_FLAG_INSTANCE_ID_MASK)
        throw InvalidArg("too many instances");

    hipStream_t s = ctx->stream;

    // geometries: TriangleMesh::commit per-triangle records (triangle_mesh.cpp:24-171)
    std::vector<GeometryRec> geos(sd->geometry_count);
    GeoRecords rec;
    uint32_t prim_offset = 0;
    for(int32_t gi = 0; gi < sd->geometry_count; ++gi)
    {
        const btrc_geometry_desc &g = sd->geometries[gi];
        if(g.triangle_count < 0 || (g.triangle_count > 0 && !g.positions))
            throw InvalidArg("geometry " + std::to_string(gi) + " has no positions");
        geos[gi].prim_offset = prim_offset;
        geos[gi].prim_count = (uint32_t)g.triangle_count;
        geos[gi].total_area = geometry_records(g, rec);
        prim_offset += (uint32_t)g.triangle_count;
    }
    std::vector<float4> &gx = rec.gx, &gy = rec.gy, &gz = rec.gz, &sa = rec.sa, &sba = rec.sba, &sca = rec.sca,
                        &lpos = rec.lpos;
    std::vector<AliasUnit> &alias_units = rec.alias;

    // textures: RGBA8 texels as one uint32 each, RGBA32F as float4
    if(sd->texture_count < 0 || (sd->texture_count > 0 && !sd->textures))
        throw InvalidArg("bad texture list");
    std::vector<TextureRec> texs(sd->texture_count);
    std::vector<uint32_t> tex8;
    std::vector<float4> tex32;
    for(int32_t i = 0; i < sd->texture_count; ++i)
    {
        const btrc_texture_desc &t = sd->textures[i];
        const std::string name = "texture " + std::to_string(i);
        if(t.width < 1 || t.height < 1 || t.width > 65536 || t.height > 65536 || !t.texels)
            throw InvalidArg(name + ": size must lie in 1..65536 with texels given");
        if(t.address_u < BTRC_ADDRESS_CLAMP || t.address_u > BTRC_ADDRESS_BORDER || t.address_v < BTRC_ADDRESS_CLAMP ||
           t.address_v > BTRC_ADDRESS_BORDER)
            throw InvalidArg(name + ": unknown address mode");
        if(t.filter != BTRC_TEX_FILTER_LINEAR && t.filter != BTRC_TEX_FILTER_POINT)
            throw InvalidArg(name + ": unknown filter");
        const size_t n = (size_t)t.width * (size_t)t.height;
        TextureRec r{};
        r.width = t.width;
        r.height = t.height;
        r.format = t.format;
        r.address_u = t.address_u;
        r.address_v = t.address_v;
        r.filter = t.filter;
        r.srgb = t.srgb_to_linear ? 1 : 0;
        for(int k = 0; k < 4; ++k)
            r.border[k] = t.border[k];
        if(t.format == BTRC_TEXTURE_RGBA8)
        {
            if(tex8.size() + n > 0xffffffffull)
                throw InvalidArg(name + ": texel arrays too large");
            r.offset = (uint32_t)tex8.size();
            const uint8_t *p = static_cast<const uint8_t *>(t.texels);
            for(size_t k = 0; k < n; ++k)
                tex8.push_back((uint32_t)p[4 * k] | ((uint32_t)p[4 * k + 1] << 8) | ((uint32_t)p[4 * k + 2] << 16) |
                               ((uint32_t)p[4 * k + 3] << 24));
        }
        else if(t.format == BTRC_TEXTURE_RGBA32F)
        {
            if(tex32.size() + n > 0xffffffffull)
                throw InvalidArg(name + ": texel arrays too large");
            r.offset = (uint32_t)tex32.size();
            const float *p = static_cast<const float *>(t.texels);
            for(size_t k = 0; k < n; ++k)
                tex32.push_back(make_float4(p[4 * k], p[4 * k + 1], p[4 * k + 2], p[4 * k + 3]));
        }
        else
            throw InvalidArg(name + ": unknown format");
        texs[i] = r;
    }
    // 8-bit decode: value / 255, then the sRGB transfer function inverted
    std::vector<float> decode(512);
    for(int v = 0; v < 256; ++v)
    {
        const double c = v / 255.0;
        decode[v] = (float)c;
        decode[256 + v] = (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
    }
    auto texture_ref = [&](int32_t ref, int32_t mat) -> int32_t {
        if(ref < 0 || ref > sd->texture_count)
            throw InvalidArg("material " + std::to_string(mat) + ": bad texture index");
        return ref - 1;
    };

    // materials
    std::vector<MaterialRec> mats(sd->material_count);
    for(int32_t i = 0; i < sd->material_count; ++i)
    {
        const btrc_material_desc &m = sd->materials[i];
        if(m.type != BTRC_MATERIAL_DIFFUSE && m.type != BTRC_MATERIAL_DISNEY && m.type != BTRC_MATERIAL_INVISIBLE)
            throw InvalidArg("material " + std::to_string(i) + ": unknown type");
        MaterialRec r{};
        r.type = m.type;
        r.shadow_terminator = m.shadow_terminator;
        for(int k = 0; k < 3; ++k)
            r.color[k] = m.color[k];
        r.metallic = m.metallic; r.roughness = m.roughness; r.specular = m.specular;
        r.specular_tint = m.specular_tint; r.anisotropic = m.anisotropic; r.sheen = m.sheen;
        r.sheen_tint = m.sheen_tint; r.clearcoat = m.clearcoat; r.clearcoat_gloss = m.clearcoat_gloss;
        r.transmission = m.transmission; r.transmission_roughness = m.transmission_roughness; r.ior = m.ior;
        r.color_tex = texture_ref(m.color_texture, i);
        r.metallic_tex = texture_ref(m.metallic_texture, i);
        r.roughness_tex = texture_ref(m.roughness_texture, i);
        r.specular_tex = texture_ref(m.specular_texture, i);
        r.specular_tint_tex = texture_ref(m.specular_tint_texture, i);
        r.anisotropic_tex = texture_ref(m.anisotropic_texture, i);
        r.sheen_tex = texture_ref(m.sheen_texture, i);
        r.sheen_tint_tex = texture_ref(m.sheen_tint_texture, i);
        r.clearcoat_tex = texture_ref(m.clearcoat_texture, i);
        r.clearcoat_gloss_tex = texture_ref(m.clearcoat_gloss_texture, i);
        r.transmission_tex = texture_ref(m.transmission_texture, i);
        r.transmission_roughness_tex = texture_ref(m.transmission_roughness_texture, i);
        r.ior_tex = texture_ref(m.ior_texture, i);
        r.normal_tex = texture_ref(m.normal_texture, i);
        mats[i] = r;
    }

    // media: ids by descending priority (ties keep the input order); id R is
    // the volume-primitive medium, id R + 1 means none
    if(sd->medium_count < 0 || (sd->medium_count > 0 && !sd->media))
        throw InvalidArg("bad medium list");
    const int32_t R = sd->medium_count;
    std::vector<int32_t> order(R);
    for(int32_t i = 0; i < R; ++i)
        order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return sd->media[a].priority > sd->media[b].priority;
    });
    std::vector<uint32_t> medium_id(R);
    std::vector<MediumRec> media(R);
    std::vector<float> voxels;
    for(int32_t rank = 0; rank < R; ++rank)
    {
        const int32_t i = order[rank];
        const btrc_medium_desc &md = sd->media[i];
        const std::string name = "medium " + std::to_string(i);
        MediumRec m{};
        m.type = md.type;
        m.g = md.g;
        if(!(std::fabs(md.g) < 1.0f))
            throw InvalidArg(name + ": g must lie in (-1, 1)");
        for(int k = 0; k < 3; ++k)
        {
            if(!std::isfinite(md.albedo[k]))
                throw InvalidArg(name + ": albedo must be finite");
            m.albedo[k] = md.albedo[k];
        }
        if(md.type == BTRC_MEDIUM_HOMOGENEOUS)
        {
            if(!(md.sigma_t > 0.0f) || !std::isfinite(md.sigma_t))
                throw InvalidArg(name + ": sigma_t must be positive and finite");
            m.sigma_t = md.sigma_t;
        }
        else if(md.type == BTRC_MEDIUM_HETEROGENEOUS)
        {
            const int32_t nx = md.grid_size[0], ny = md.grid_size[1], nz = md.grid_size[2];
            if(nx <= 0 || ny <= 0 || nz <= 0 || !md.sigma_t_grid)
                throw InvalidArg(name + ": heterogeneous media need a density grid");
            const size_t nv = (size_t)nx * ny * nz;
            if(voxels.size() + nv * (md.albedo_grid ? 4 : 1) > 0xffffffffull)
                throw InvalidArg(name + ": voxel grids too large");
            m.nx = nx;
            m.ny = ny;
            m.nz = nz;
            float mx = -MAX_FLOAT;
            m.density_offset = (uint32_t)voxels.size();
            for(size_t v = 0; v < nv; ++v)
            {
                const float x = md.sigma_t_grid[v];
                if(!std::isfinite(x))
                    throw InvalidArg(name + ": density grid holds a non-finite value");
                mx = std::max(mx, x);
                voxels.push_back(x);
            }
            // HetergeneousMedium::sample, hetergeneous.cpp:34-35
            m.inv_max_density = 1.0f / std::max(mx, 0.001f);
            if(md.albedo_grid)
            {
                m.albedo_grid = 1;
                m.albedo_offset = (uint32_t)voxels.size();
                for(size_t v = 0; v < 3 * nv; ++v)
                {
                    if(!std::isfinite(md.albedo_grid[v]))
                        throw InvalidArg(name + ": albedo grid holds a non-finite value");
                    voxels.push_back(md.albedo_grid[v]);
                }
            }
        }
        else
            throw InvalidArg(name + ": unknown type");
        if(md.has_transform)
        {
            m.has_xform = 1;
            for(int k = 0; k < 12; ++k)
                m.world_to_local[k] = md.world_to_local[k];
        }
        media[rank] = m;
        medium_id[i] = (uint32_t)rank;
    }
    // volume primitives (core/volume.cpp:21-79): grids appended to the voxels
    if(sd->volume_count < 0 || (sd->volume_count > 0 && !sd->volumes))
        throw InvalidArg("bad volume list");
    std::vector<VolumeRec> vols(sd->volume_count);
    for(int32_t i = 0; i < sd->volume_count; ++i)
    {
        const btrc_volume_desc &vd = sd->volumes[i];
        const std::string name = "volume " + std::to_string(i);
        VolumeRec v{};
        const float *e[3] = { vd.x, vd.y, vd.z };
        float *xd[3] = { v.xd, v.yd, v.zd };
        float *dst[3] = { v.x, v.y, v.z };
        float *inv[3] = { &v.inv_x2, &v.inv_y2, &v.inv_z2 };
        for(int k = 0; k < 3; ++k)
        {
            v.o[k] = vd.o[k];
            const float l2 = e[k][0] * e[k][0] + e[k][1] * e[k][1] + e[k][2] * e[k][2];
            if(!(l2 > 0.0f) || !std::isfinite(l2) || !std::isfinite(vd.o[k]))
                throw InvalidArg(name + ": edges must be finite and non-zero");
            *inv[k] = 1.0f / l2;
            for(int c = 0; c < 3; ++c)
            {
                dst[k][c] = e[k][c];
                xd[k][c] = e[k][c] / l2;
            }
        }
        if(vd.sigma_t_size[0] > 0)
        {
            const int32_t nx = vd.sigma_t_size[0], ny = vd.sigma_t_size[1], nz = vd.sigma_t_size[2];
            if(ny <= 0 || nz <= 0 || !vd.sigma_t_grid)
                throw InvalidArg(name + ": bad density grid");
            const size_t nv = (size_t)nx * ny * nz;
            if(voxels.size() + nv > 0xffffffffull)
                throw InvalidArg(name + ": voxel grids too large");
            v.nx = nx; v.ny = ny; v.nz = nz;
            v.density_offset = (uint32_t)voxels.size();
            float mx = vd.sigma_t_grid[0];
            for(size_t k = 0; k < nv; ++k)
            {
                const float x = vd.sigma_t_grid[k];
                if(!std::isfinite(x) || x < 0.0f)
                    throw InvalidArg(name + ": densities must be finite and non-negative");
                mx = std::max(mx, x);
                voxels.push_back(x);
            }
            v.max_density = mx;
        }
        else
        {
            if(!std::isfinite(vd.sigma_t) || vd.sigma_t < 0.0f)
                throw InvalidArg(name + ": sigma_t must be finite and non-negative");
            v.sigma_t = vd.sigma_t;
            v.max_density = vd.sigma_t;
        }
        for(int k = 0; k < 3; ++k)
            v.albedo[k] = vd.albedo[k];
        if(vd.albedo_size[0] > 0)
        {
            const int32_t ax = vd.albedo_size[0], ay = vd.albedo_size[1], az = vd.albedo_size[2];
            if(ay <= 0 || az <= 0 || !vd.albedo_grid)
                throw InvalidArg(name + ": bad albedo grid");
            const size_t nv = 3 * (size_t)ax * ay * az;
            if(voxels.size() + nv > 0xffffffffull)
                throw InvalidArg(name + ": voxel grids too large");
            v.ax = ax; v.ay = ay; v.az = az;
            v.albedo_offset = (uint32_t)voxels.size();
            voxels.insert(voxels.end(), vd.albedo_grid, vd.albedo_grid + nv);
        }
        vols[i] = v;
    }

    auto resolve_medium = [&](int32_t idx, int32_t inst) -> uint32_t {
        if(idx < 0)
            return (uint32_t)R + 1;
        if(idx >= R)
            throw InvalidArg("instance " + std::to_string(inst) + ": bad medium index");
        return medium_id[idx];
    };

    if(sd->accel != BTRC_ACCEL_FLAT && sd->accel != BTRC_ACCEL_TWO_LEVEL && sd->accel != BTRC_ACCEL_FLAT_GPU)
        throw InvalidArg("unknown acceleration structure type");
    const bool two_level = sd->accel == BTRC_ACCEL_TWO_LEVEL;

    // instances, lights and (one level) the flattened world-space triangle list
    std::vector<InstanceRec> insts(sd->instance_count);
    std::vector<LightRec> lights;
    std::vector<float> world;
    std::vector<uint32_t> tri_inst, tri_prim;
    uint32_t prim_base = 0;
    for(int32_t i = 0; i < sd->instance_count; ++i)
    {
        const btrc_instance_desc &d = sd->instances[i];
        if(d.geometry < 0 || d.geometry >= sd->geometry_count)
            throw InvalidArg("instance " + std::to_string(i) + ": bad geometry index");
        if(d.material < 0 || d.material >= sd->material_count)
            throw InvalidArg("instance " + std::to_string(i) + ": bad material index");
        InstanceRec r{};
        r.geometry_id = d.geometry;
        r.material_id = d.material;
        r.light_id = -1;
        r.inner_medium = resolve_medium(d.inner_medium, i);
        r.outer_medium = resolve_medium(d.outer_medium, i);
        r.flag = 0;
        r.prim_base = prim_base;
        r.xf = make_xform(d.local_to_world, d.world_to_local);
        if(d.has_light)
        {
            LightRec l{};
            l.type = LIGHT_MESH;
            l.instance = i;
            l.geometry = d.geometry;
            l.scale = length(xf_point(r.xf, v3(1, 0, 0)) - xf_point(r.xf, v3(0, 0, 0)));
            for(int k = 0; k < 3; ++k)
                l.intensity[k] = d.light_intensity[k];
            r.light_id = (int32_t)lights.size();
            lights.push_back(l);
        }
        insts[i] = r;
        const btrc_geometry_desc &g = sd->geometries[d.geometry];
        for(int32_t t = 0; t < (two_level ? 0 : g.triangle_count); ++t)
        {
            const float *p = g.positions + 9 * (size_t)t;
            for(int k = 0; k < 3; ++k)
            {
                const V3 w = xf_point(r.xf, v3(p[3 * k], p[3 * k + 1], p[3 * k + 2]));
                world.push_back(w.x);
                world.push_back(w.y);
                world.push_back(w.z);
            }
            tri_inst.push_back((uint32_t)i);
            tri_prim.push_back((uint32_t)t);
        }
        prim_base += (uint32_t)g.triangle_count;
    }
    int32_t envir_index = -1;
    if(sd->envir_light.type == BTRC_ENVIR_GRADIENT_SKY)
    {
        LightRec l{};
        l.type = LIGHT_GRADIENT_SKY;
        l.instance = -1;
        l.geometry = -1;
        const V3 up = normalize(v3(sd->envir_light.up[0], sd->envir_light.up[1], sd->envir_light.up[2]));
        l.up[0] = up.x; l.up[1] = up.y; l.up[2] = up.z;
        for(int k = 0; k < 3; ++k)
        {
            l.lower[k] = sd->envir_light.lower[k];
            l.upper[k] = sd->envir_light.upper[k];
        }
        envir_index = (int32_t)lights.size();
        lights.push_back(l);
    }
    else if(sd->envir_light.type == BTRC_ENVIR_IBL)
    {
        const btrc_envir_light_desc &e = sd->envir_light;
        if(e.texture < 0 || e.texture >= sd->texture_count)
            throw InvalidArg("ibl: bad texture index");
        if(e.lut_width < 1 || e.lut_height < 1 || (int64_t)e.lut_width * e.lut_height > MAX_ENVIR_TABLE)
            throw InvalidArg("ibl: importance table size out of range");
        const V3 up = v3(e.up[0], e.up[1], e.up[2]);
        if(!(dot(up, up) > 0.0f) || !std::isfinite(dot(up, up)))
            throw InvalidArg("ibl: bad up direction");
        LightRec l{};
        l.type = LIGHT_IBL;
        l.instance = -1;
        l.geometry = -1;
        // Frame::from_z (utils/math/frame.cpp:45-57), normalized once more by CFrame(Frame)
        const Frame f0 = frame_from_z(up);
        const Frame f = make_frame(f0.x, f0.y, f0.z);
        l.frame_x[0] = f.x.x; l.frame_x[1] = f.x.y; l.frame_x[2] = f.x.z;
        l.frame_y[0] = f.y.x; l.frame_y[1] = f.y.y; l.frame_y[2] = f.y.z;
        l.up[0] = f.z.x; l.up[1] = f.z.y; l.up[2] = f.z.z;
        l.texture = e.texture;
        l.lut_w = e.lut_width;
        l.lut_h = e.lut_height;
        envir_index = (int32_t)lights.size();
        lights.push_back(l);
    }
    else if(sd->envir_light.type != BTRC_ENVIR_NONE)
        throw InvalidArg("unknown environment light type");

    const uint32_t ntri = prim_base;
    ctx->has_scene = false;
    BTRC_HIP_CHECK(hipStreamSynchronize(s));
    const auto build_start = std::chrono::steady_clock::now();
    Accel acc = two_level ? build_two_level(sd, insts)
              : (sd->accel == BTRC_ACCEL_FLAT_GPU && ntri > 0) ? build_flat_gpu(ctx, world, tri_inst, tri_prim, ntri)
                                                                : build_flat(world, tri_inst, tri_prim, ntri);
    ctx->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - build_start).count();

    if(!acc.on_device)
    {
        ctx->nodes.upload(acc.nodes.data(), acc.nodes.size(), s);
        ctx->tris.upload(acc.tris.data(), acc.tris.size(), s);
    }
    ctx->trace_inst.upload(acc.inst.data(), acc.inst.size(), s);
    ctx->instances.upload(insts.data(), insts.size(), s);
    ctx->geometries.upload(geos.data(), geos.size(), s);
    ctx->materials.upload(mats.data(), mats.size(), s);
    ctx->lights.upload(lights.data(), lights.size(), s);
    ctx->gx_u_a.upload(gx.data(), gx.size(), s);
    ctx->gy_u_ba.upload(gy.data(), gy.size(), s);
    ctx->gz_u_ca.upload(gz.data(), gz.size(), s);
    ctx->sn_v_a.upload(sa.data(), sa.size(), s);
    ctx->sn_v_ba.upload(sba.data(), sba.size(), s);
    ctx->sn_v_ca.upload(sca.data(), sca.size(), s);
    ctx->light_pos.upload(lpos.data(), lpos.size(), s);
    ctx->alias.upload(alias_units.data(), alias_units.size(), s);
    ctx->media.upload(media.data(), media.size(), s);
    ctx->voxels.upload(voxels.data(), voxels.size(), s);
    ctx->volumes.upload(vols.data(), vols.size(), s);
    ctx->textures.upload(texs.data(), texs.size(), s);
    ctx->texels8.upload(tex8.data(), tex8.size(), s);
    ctx->texels32.upload(tex32.data(), tex32.size(), s);
    ctx->tex_decode.upload(decode.data(), decode.size(), s);
    BTRC_HIP_CHECK(hipStreamSynchronize(s));
    ctx->texture_count = sd->texture_count;
    if(envir_index >= 0 && lights[envir_index].type == LIGHT_IBL)
        build_envir_table(ctx, lights[envir_index]);
    else
    {
        ctx->env_probs.release();
        ctx->env_alias.release();
        ctx->env_probs_host.clear();
    }

    ctx->medium_count = R;
    ctx->volume_count = sd->volume_count;
    // each volume's o, o + x, o + y, o + z (VolumePrimitive::get_bounding_box)
    for(int a = 0; a < 3; ++a)
    {
        ctx->tri_lo[a] = acc.lo[a];
        ctx->tri_hi[a] = acc.hi[a];
        ctx->vol_lo[a] = MAX_FLOAT;
        ctx->vol_hi[a] = -MAX_FLOAT;
    }
    for(const VolumeRec &v : vols)
        for(int c = 0; c < 4; ++c)
            for(int a = 0; a < 3; ++a)
            {
                const float p = c == 0 ? v.o[a] : v.o[a] + (c == 1 ? v.x[a] : c == 2 ? v.y[a] : v.z[a]);
                ctx->vol_lo[a] = std::min(ctx->vol_lo[a], p);
                ctx->vol_hi[a] = std::max(ctx->vol_hi[a], p);
            }
    set_world_bounds(ctx);
    ctx->geos_host = geos;
    ctx->refit_levels = acc.level_offsets;
    if(!two_level)
    {
        std::vector<uint32_t> order = acc.level_order;
        if(order.empty())
        {
            order.resize(acc.on_device ? (size_t)acc.device_nodes : acc.nodes.size());
            for(size_t k = 0; k < order.size(); ++k)
                order[k] = (uint32_t)k;
        }
        ctx->refit_order.upload(order.data(), order.size(), s);
    }
    else
        ctx->refit_order.release();
    ctx->light_count = (int32_t)lights.size();
    ctx->envir_light_index = envir_index;
    ctx->instance_count = sd->instance_count;
    ctx->triangle_count = ntri;
    ctx->node_count = acc.on_device ? acc.device_nodes : (int64_t)acc.nodes.size();
    ctx->bvh_depth = acc.depth;
    ctx->stack_size = acc.stack_bound + 1;
    ctx->two_level = two_level;
    size_traversal(ctx);
    ctx->has_scene = true;
}

// PinholeCamera::commit, camera/pinhole.cpp:55-77
CameraConsts camera_consts(const btrc_camera_desc *cd)
{
    const V3 eye = v3(cd->eye[0], cd->eye[1], cd->eye[2]);
    const V3 dst = v3(cd->dst[0], cd->dst[1], cd->dst[2]);
    const V3 up = v3(cd->up[0], cd->up[1], cd->up[2]);
    const V3 forward = normalize(dst - eye);
    const V3 ex = normalize(cross(forward, up));
    const V3 ey = normalize(cross(forward, ex));
    const float y_len = 2.0f * std::tan(0.5f * cd->fov_y_deg * PI / 180.0f);
    const float x_len = cd->w_over_h * y_len;
    const V3 lbc = eye + forward - 0.5f * (x_len * ex + y_len * ey);
    const V3 fx = x_len * ex;
    const V3 fy = y_len * ey;
    CameraConsts c;
    c.eye[0] = eye.x; c.eye[1] = eye.y; c.eye[2] = eye.z;
    c.lbc[0] = lbc.x; c.lbc[1] = lbc.y; c.lbc[2] = lbc.z;
    c.film_x[0] = fx.x; c.film_x[1] = fx.y; c.film_x[2] = fx.z;
    c.film_y[0] = fy.x; c.film_y[1] = fy.y; c.film_y[2] = fy.z;
    return c;
}

void set_camera(btrc_ctx *ctx, const btrc_camera_desc *cd)
{
    if(!cd)
        throw InvalidArg("camera is null");
    ctx->camera = camera_consts(cd);
    ctx->has_camera = true;
}

struct StageTimer
{
    bool enabled = false;
    hipStream_t stream = nullptr;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events[STAGE_COUNT];
    std::vector<hipEvent_t> pool;
    size_t used = 0;

    hipEvent_t get()
    {
        if(used == pool.size())
        {
            hipEvent_t e;
            BTRC_HIP_CHECK(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[used++];
    }
    template<typename F>
    void run(int stage, F &&f)
    {
        if(!enabled)
        {
            f();
            return;
        }
        hipEvent_t a = get(), b = get();
        BTRC_HIP_CHECK(hipEventRecord(a, stream));
        f();
        BTRC_HIP_CHECK(hipEventRecord(b, stream));
        events[stage].push_back({ a, b });
    }
    void collect(double *ms, uint64_t *launches)
    {
        for(int s = 0; s < STAGE_COUNT; ++s)
        {
            double total = 0;
            for(auto &p : events[s])
            {
                float t = 0;
                BTRC_HIP_CHECK(hipEventElapsedTime(&t, p.first, p.second));
                total += t;
            }
            ms[s] = total;
            launches[s] = events[s].size();
        }
    }
    ~StageTimer()
    {
        for(auto e : pool)
            (void)hipEventDestroy(e);
    }
};

// ---- neural radiance cache

constexpr int64_t NRC_RECORDS = 1 << 16;     // training vertices per frame
constexpr int64_t NRC_BATCH = 1 << 14;       // vertices per optimizer step

NrcWeights nrc_weights(btrc_ctx *ctx)
{
    auto &n = ctx->nrc;
    return NrcWeights{ n.w.ptr, n.m.ptr, n.v.ptr, n.wb.ptr, n.wtb.ptr, n.ema.ptr, n.ema_b.ptr };
}

// the weights rendered frames query: the trained ones, or their moving average
NrcWeights nrc_render_weights(btrc_ctx *ctx)
{
    NrcWeights w = nrc_weights(ctx);
    if(ctx->nrc.params.ema_decay > 0.0f)
        w.w_bf16 = ctx->nrc.ema_b.ptr;
    return w;
}

NrcBox nrc_box(btrc_ctx *ctx)
{
    NrcBox b;
    for(int a = 0; a < 3; ++a)
    {
        b.lo[a] = ctx->scene_lo[a];
        b.inv_extent[a] = 1.0f / (ctx->scene_hi[a] - ctx->scene_lo[a]);
    }
    return b;
}

void nrc_sync(btrc_ctx *ctx);

void nrc_alloc(btrc_ctx *ctx)
{
    auto &n = ctx->nrc;
    n.w.alloc(NRC_PARAMS);
    n.m.alloc(NRC_PARAMS);
    n.v.alloc(NRC_PARAMS);
    n.grad.alloc(NRC_PARAMS);
    n.wb.alloc(NRC_PARAMS);
    n.wtb.alloc(NRC_PARAMS);
    n.ema.alloc(NRC_PARAMS);
    n.ema_b.alloc(NRC_PARAMS);
    for(int k = 0; k < 2; ++k)
    {
        n.rec[k].alloc(NRC_RECORDS);
        n.acc[k].alloc(4 * NRC_RECORDS);
    }
    n.counts.alloc(3);
    n.index.alloc(NRC_RECORDS);
    n.valid_mask.alloc(NRC_RECORDS / 64);
    n.xt.alloc((size_t)NRC_LAYERS * NRC_IN * NRC_RECORDS);
    n.dzt.alloc((size_t)NRC_LAYERS * NRC_IN * NRC_RECORDS);
    n.loss_partial.alloc(NRC_RECORDS / 32);
    n.grad_partial.alloc((size_t)NRC_KSPLIT * NRC_PARAMS);
    n.count_partial.alloc(NRC_RECORDS / 32);
    n.g2_partial.alloc(NRC_REDUCE_BLOCKS);
    n.step_scale.alloc(1);
    n.step_stats.alloc(64);
}

// uniform He initialisation: U(-sqrt(6 / fan_in), sqrt(6 / fan_in)); the
// output layer uses sqrt(6 / (fan_in + fan_out)); padding rows are zero
void nrc_reset(btrc_ctx *ctx, uint64_t seed)
{
    nrc_sync(ctx);
    nrc_alloc(ctx);
    auto &n = ctx->nrc;
    std::vector<float> w(NRC_PARAMS, 0.0f);
    Pcg32 rng = pcg32_seed(seed, mix_bits(seed));
    for(int l = 0; l < NRC_LAYERS; ++l)
    {
        const bool out = l == NRC_LAYERS - 1;
        const float a = out ? std::sqrt(6.0f / (float)(NRC_IN + NRC_OUT)) : std::sqrt(6.0f / (float)NRC_IN);
        const int rows = out ? NRC_OUT : NRC_WIDTH;
        for(int r = 0; r < rows; ++r)
            for(int c = 0; c < NRC_IN; ++c)
                w[nrc_layer_offset(l) + r * NRC_IN + c] = a * (2.0f * pcg32_uniform_float(rng) - 1.0f);
    }
    hipStream_t s = ctx->stream;
    n.w.upload(w.data(), w.size(), s);
    BTRC_HIP_CHECK(hipMemsetAsync(n.m.ptr, 0, NRC_PARAMS * sizeof(float), s));
    BTRC_HIP_CHECK(hipMemsetAsync(n.v.ptr, 0, NRC_PARAMS * sizeof(float), s));
    launch_nrc_publish(s, nrc_weights(ctx));
    BTRC_HIP_CHECK(hipGetLastError());
    BTRC_HIP_CHECK(hipStreamSynchronize(s));
    n.steps_total = 0;
    n.last = btrc_nrc_stats{};
    n.ready = true;
}

void nrc_require(btrc_ctx *ctx)
{
    if(!ctx->nrc.ready)
        nrc_reset(ctx, 1);
}

NrcAdam nrc_adam(btrc_ctx *ctx)
{
    // Adam settings of the cache paper: beta1 0.9, beta2 0.99, eps 1e-15
    auto &n = ctx->nrc;
    const double t = (double)(n.steps_total + 1);
    NrcAdam a;
    a.lr = n.params.learning_rate;
    a.beta1 = 0.9f;
    a.beta2 = 0.99f;
    a.eps = 1e-15f;
    a.inv_bias1 = (float)(1.0 / (1.0 - std::pow(0.9, t)));
    a.inv_bias2 = (float)(1.0 / (1.0 - std::pow(0.99, t)));
    a.loss_scale = n.params.loss_scale;
    a.ema_decay = n.params.ema_decay;
    return a;
}

NrcScratch nrc_scratch(btrc_ctx *ctx, int64_t batch)
{
    auto &n = ctx->nrc;
    (void)batch;
    return NrcScratch{ n.xt.ptr, n.dzt.ptr, n.loss_partial.ptr, n.count_partial.ptr, n.grad_partial.ptr,
                       n.step_scale.ptr, n.g2_partial.ptr };
}

// `steps` optimizer steps over the records, `batch` at a time: step k
// takes the k-th batch (rotating through all records), or always the first
void nrc_train(btrc_ctx *ctx, hipStream_t s, int set, int32_t steps, int64_t batch, bool rotate)
{
    auto &n = ctx->nrc;
    const int64_t batches = NRC_RECORDS / batch;
    for(int32_t k = 0; k < steps; ++k)
    {
        // a rendered frame's steps take equal parts of its valid records
        // (compacted after the bounce loop); fits take the loaded batch
        NrcBatch b{ n.rec[set].ptr, rotate ? (k % batches) * batch : 0, batch, n.acc[set].ptr, n.counts.ptr + 1,
                    n.params.loss_scale, rotate ? 1 : 0, rotate ? n.index.ptr : nullptr,
                    rotate ? n.counts.ptr + 2 : nullptr, k, steps };
        launch_nrc_train_step(s, nrc_weights(ctx), nrc_box(ctx), b, nrc_scratch(ctx, batch), nrc_adam(ctx),
                              n.grad.ptr, n.step_stats.ptr + k);
        BTRC_HIP_CHECK(hipGetLastError());
        ++n.steps_total;
    }
}

void nrc_collect(btrc_ctx *ctx, int32_t steps)
{
    auto &n = ctx->nrc;
    std::vector<NrcStepStats> st(std::max(steps, 0));
    int32_t counts[2] = { 0, 0 };
    BTRC_HIP_CHECK(hipMemcpyAsync(counts, n.counts.ptr, sizeof(counts), hipMemcpyDeviceToHost, ctx->stream));
    if(steps > 0)
        BTRC_HIP_CHECK(hipMemcpyAsync(st.data(), n.step_stats.ptr, steps * sizeof(NrcStepStats), hipMemcpyDeviceToHost,
                                      ctx->stream));
    BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    btrc_nrc_stats r{};
    r.queries = counts[0];
    r.steps_total = n.steps_total;
    for(int32_t k = 0; k < steps; ++k)
    {
        r.records += st[k].samples;
        if(k < 4)
        {
            r.loss[k] = st[k].loss;
            r.grad_norm[k] = st[k].grad_norm;
        }
        r.skipped += st[k].skipped;
    }
    n.last = r;
}

// the Halton sampler's first-digit tables in device memory (a pure function
// of the prime table: 226,691 values, 443 KB)
HaltonTables halton_tables(btrc_ctx *ctx)
{
    if(!ctx->halton_values.ptr)
    {
        constexpr PrimeTable primes = make_prime_table();
        std::vector<HaltonDim> dims(HALTON_TABLE_DIMS);
        uint32_t total = 0;
        for(int d = 0; d < HALTON_TABLE_DIMS; ++d)
        {
            dims[d] = halton_dim_shape((uint32_t)primes.p[d], total);
            total += dims[d].span;
        }
        std::vector<uint16_t> values(total);
        for(int d = 0; d < HALTON_TABLE_DIMS; ++d)
            fill_halton_table((uint32_t)primes.p[d], (uint32_t)d, dims[d], values.data() + dims[d].first);
        ctx->halton_dims.upload(dims.data(), dims.size(), ctx->stream);
        ctx->halton_values.upload(values.data(), values.size(), ctx->stream);
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
    return HaltonTables{ ctx->halton_values.ptr, ctx->halton_dims.ptr };
}

// Waits for the cache training still running on the training stream and
// reads the last frame's statistics; every cache entry point calls it first
void nrc_sync(btrc_ctx *ctx)
{
    auto &n = ctx->nrc;
    if(ctx->train_pending)
    {
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->train_stream));
        ctx->train_pending = false;
        n.pending_set = -1;
    }
    if(n.stats_pending)
    {
        n.stats_pending = false;
        nrc_collect(ctx, n.pending_steps);
    }
}

// The views of a render: `n` cameras with film sizes, or (views null) the
// context's camera at params' size. Films are laid out view after view.
std::vector<ViewRec> make_views(btrc_ctx *ctx, const btrc_render_params *p, int32_t n, const btrc_view_desc *views)
{
    std::vector<ViewRec> out;
    if(!views)
    {
        if(!ctx->has_camera)
            throw StateError("render before btrc_hip_set_camera");
        ViewRec v{};
        v.camera = ctx->camera;
        v.width = p->width;
        v.height = p->height;
        out.push_back(v);
    }
    else
    {
        if(n < 1 || n > MAX_VIEWS)
            throw InvalidArg("view count must lie in 1.." + std::to_string(MAX_VIEWS));
        for(int32_t i = 0; i < n; ++i)
        {
            ViewRec v{};
            v.camera = camera_consts(&views[i].camera);
            v.width = views[i].width;
            v.height = views[i].height;
            out.push_back(v);
        }
    }
    uint64_t pixel_base = 0, path_base = 0;
    for(ViewRec &v : out)
    {
        if(v.width <= 0 || v.height <= 0)
            throw InvalidArg("width and height must be positive");
        v.halton = make_halton_consts(v.width, v.height);
        v.pixel_base = (uint32_t)pixel_base;
        v.path_base = path_base;
        const uint64_t pixels = (uint64_t)v.width * (uint64_t)v.height;
        pixel_base += pixels;
        path_base += pixels * (uint64_t)std::max(p->spp, 0);
        if(pixel_base > 0x7fffffffull)
            throw InvalidArg("film too large");
    }
    return out;
}

void render(btrc_ctx *ctx, const btrc_render_params *p, int32_t view_count, const btrc_view_desc *view_desc,
            btrc_render_stats *stats, int32_t capture_bounces, btrc_ray *capture, int64_t capture_max,
            int64_t *capture_count)
{
    if(!p)
        throw InvalidArg("params is null");
    if(!ctx->has_scene)
        throw StateError("render before btrc_hip_upload_scene");
    if(p->spp <= 0 || p->state_count <= 0)
        throw InvalidArg("spp and state_count must be positive");
    if(p->state_count > (1 << 28))
        throw InvalidArg("state_count above 2^28");
    if(p->min_depth < 0 || p->max_depth < 0)
        throw InvalidArg("depths must be non-negative");
    if(p->sampler != BTRC_SAMPLER_HALTON && p->sampler != BTRC_SAMPLER_INDEPENDENT)
        throw InvalidArg("unknown sampler");
    if(p->sample_offset < 0)
        throw InvalidArg("sample_offset must be non-negative");
    if(((uint64_t)p->spp + (uint64_t)p->sample_offset) * 31104ull > 0x7fffffffull)
        throw InvalidArg("spp too large");
    const std::vector<ViewRec> views = make_views(ctx, p, view_count, view_desc);
    const ViewRec &last = views.back();
    const uint64_t pixels = (uint64_t)last.pixel_base + (uint64_t)last.width * (uint64_t)last.height;
    const uint64_t total = last.path_base + (uint64_t)last.width * (uint64_t)last.height * (uint64_t)p->spp;
    const bool nrc_on = ctx->nrc.params.enabled != 0;
    if(nrc_on)
    {
        if(p->spp != 1)
            throw InvalidArg("the radiance cache renders one sample per pixel per frame");
        if(ctx->medium_count > 0 || ctx->volume_count > 0)
            throw InvalidArg("the radiance cache does not support participating media");
        if(p->max_depth + 1 > NRC_RECORDS)
            throw InvalidArg("max_depth too large for the radiance cache");
        nrc_require(ctx);
    }

    hipStream_t s = ctx->stream;
    const size_t P = (size_t)p->state_count;
    // pool sides: the last bounce's stripes, then the generated states;
    // shadow queue: one stripe per COUNTER_STRIPES-th of the inputs
    const size_t PS = pool_capacity(P), SS = shadow_capacity(P);
    if(ctx->pool_size != P)
    {
        ctx->pool[0].alloc(PS);
        ctx->pool[1].alloc(PS);
        ctx->hit_flag.alloc(P);
        ctx->hit_tuv.alloc(P);
        ctx->sh_o.alloc(SS);
        ctx->sh_d.alloc(SS);
        ctx->sh_beta.alloc(SS);
        ctx->pool_size = P;
    }
    const bool media = ctx->medium_count > 0 || ctx->volume_count > 0;
    if(media)
    {
        ctx->sh_seed.alloc(SS);
        ctx->medium_scatter.alloc(2 * P);
    }
    ctx->counters.alloc(1);
    if(!ctx->host_counters)
        BTRC_HIP_CHECK(hipHostMalloc((void **)&ctx->host_counters, sizeof(WaveCounters), hipHostMallocDefault));
    // a render adds to the previous film when asked and the film matches
    bool same_views = ctx->film_views.size() == views.size();
    for(size_t v = 0; same_views && v < views.size(); ++v)
        same_views = ctx->film_views[v].width == views[v].width && ctx->film_views[v].height == views[v].height;
    const bool keep = p->accumulate && same_views && ctx->film_has64 == (p->film_fp64 != 0) &&
                      ctx->film_has_albedo == (p->albedo != 0) && ctx->film_has_normal == (p->normal != 0) &&
                      ctx->film_spp > 0;
    ctx->film.alloc(4 * pixels);
    if(p->film_fp64)
        ctx->film64.alloc(3 * pixels);
    if(p->albedo)
        ctx->film_albedo.alloc(4 * pixels);
    if(p->normal)
        ctx->film_normal.alloc(4 * pixels);
    if(!keep)
    {
        BTRC_HIP_CHECK(hipMemsetAsync(ctx->film.ptr, 0, 4 * pixels * sizeof(float), s));
        if(p->film_fp64)
            BTRC_HIP_CHECK(hipMemsetAsync(ctx->film64.ptr, 0, 3 * pixels * sizeof(double), s));
        if(p->albedo)
            BTRC_HIP_CHECK(hipMemsetAsync(ctx->film_albedo.ptr, 0, 4 * pixels * sizeof(float), s));
        if(p->normal)
            BTRC_HIP_CHECK(hipMemsetAsync(ctx->film_normal.ptr, 0, 4 * pixels * sizeof(float), s));
    }
    ctx->film_has_albedo = p->albedo != 0;
    ctx->film_has_normal = p->normal != 0;
    ctx->film_views = views;
    ctx->film_pixels = (int64_t)pixels;
    ctx->views.upload(views.data(), views.size(), s);
    ctx->film_spp = keep ? ctx->film_spp + p->spp : p->spp;
    ctx->film_has64 = p->film_fp64 != 0;

    // Reporter calls of WavefrontPathTracer::render (wavefront.cpp:221,287,314):
    // a new stage now, progress at every counter read-back, the stage
    // completed when the loop ends (also when it stops early)
    struct ReporterStage
    {
        const btrc_reporter &r;
        explicit ReporterStage(const btrc_reporter &rep) : r(rep)
        {
            if(r.new_stage)
                r.new_stage(r.user, "");
        }
        ~ReporterStage()
        {
            if(r.complete_stage)
                r.complete_stage(r.user);
        }
        void progress(const WaveCounters &c) const
        {
            if(r.progress && c.total > 0)
                r.progress(r.user, (float)(100.0 * (double)(c.generated - (uint64_t)c.active) / (double)c.total));
        }
    } reporter_stage(ctx->reporter);

    WaveCounters init{};
    init.total = total;
    *ctx->host_counters = init;
    BTRC_HIP_CHECK(hipMemsetAsync(ctx->counters.ptr, 0, sizeof(WaveState), s));
    BTRC_HIP_CHECK(hipMemcpyAsync(ctx->wc(), ctx->host_counters, sizeof(WaveCounters), hipMemcpyHostToDevice, s));

    RenderConsts rc{};
    rc.views = ctx->views.ptr;
    rc.view_count = (int32_t)views.size();
    rc.spp = p->spp;
    rc.min_depth = p->min_depth;
    rc.max_depth = p->max_depth;
    rc.rr_threshold = p->rr_threshold;
    rc.rr_cont_prob = p->rr_cont_prob;
    rc.state_count = p->state_count;
    rc.vol_prim_medium = (uint32_t)ctx->medium_count;
    rc.film_fp64 = p->film_fp64;
    rc.tile = p->tile != 0;
    rc.sample_offset = p->sample_offset;
    NrcView nv{ -1, 0, -1, nullptr, nullptr, nullptr, nullptr };
    if(nrc_on)
    {
        auto &n = ctx->nrc;
        // record slots per training path: its vertices at depths 0..max_depth,
        // or 0..train_depth - 1 when the path ends at train_depth with a query
        const int32_t slots = n.params.train_depth > 0 ? n.params.train_depth : p->max_depth + 1;
        rc.nrc_stride = std::max(1, n.params.train ? n.params.train_stride : 0);
        rc.nrc_phase = p->sample_offset % rc.nrc_stride;
        rc.nrc_max_paths = n.params.train ? (int32_t)(NRC_RECORDS / slots) : 0;
        if(!n.params.train)
            rc.nrc_stride = 0;
        n.queries.alloc(std::max<size_t>(1, (size_t)pixels));
        // this frame's record set; a training still reading it finishes first
        n.set ^= 1;
        if(ctx->train_pending && n.pending_set == n.set)
            BTRC_HIP_CHECK(hipStreamWaitEvent(s, ctx->train_done, 0));
        nv = NrcView{ n.params.query_depth, slots, n.params.train_depth > 0 ? n.params.train_depth : -1,
                      n.queries.ptr, n.counts.ptr, n.rec[n.set].ptr, n.acc[n.set].ptr };
        launch_nrc_clear(s, n.rec[n.set].ptr, NRC_RECORDS, n.acc[n.set].ptr, NRC_RECORDS);
        BTRC_HIP_CHECK(hipMemsetAsync(n.counts.ptr, 0, sizeof(int32_t), s)); // the query count
    }
    rc.filter = ctx->filter;
    rc.halton = halton_tables(ctx);

    const TraceScene ts = ctx->trace_scene();
    SceneView sv = ctx->scene_view();
    {
        // 1.2 |bbox of the scene and the cameras' eyes| (wavefront.cpp:138-139)
        V3 lo = v3(ctx->world_lo[0], ctx->world_lo[1], ctx->world_lo[2]);
        V3 hi = v3(ctx->world_hi[0], ctx->world_hi[1], ctx->world_hi[2]);
        for(const ViewRec &v : views)
        {
            const V3 e = v3(v.camera.eye[0], v.camera.eye[1], v.camera.eye[2]);
            lo = v3(std::min(lo.x, e.x), std::min(lo.y, e.y), std::min(lo.z, e.z));
            hi = v3(std::max(hi.x, e.x), std::max(hi.y, e.y), std::max(hi.z, e.z));
        }
        sv.world_diagonal = 1.2f * length(hi - lo);
    }
    FilmView fv{ ctx->film.ptr, p->film_fp64 ? ctx->film64.ptr : nullptr, p->albedo ? ctx->film_albedo.ptr : nullptr,
                 p->normal ? ctx->film_normal.ptr : nullptr };
    HitSOA hit{ ctx->hit_flag.ptr, ctx->hit_tuv.ptr };
    if(nrc_on)
        ctx->sh_train.alloc(SS);
    ShadowSOA sh{ ctx->sh_o.ptr, ctx->sh_d.ptr, ctx->sh_beta.ptr, media ? ctx->sh_seed.ptr : nullptr,
                  nrc_on ? ctx->sh_train.ptr : nullptr, nrc_on ? ctx->nrc.acc[ctx->nrc.set].ptr : nullptr };

    const int grid256 = ctx->persistent_grid(256, (int64_t)P);
    const TraceOverflow ovf = ctx->trace_overflow();

    // the constants live in device memory so that a captured loop stays
    // valid from one render to the next
    if(!ctx->rc_host)
        BTRC_HIP_CHECK(hipHostMalloc((void **)&ctx->rc_host, sizeof(RenderConsts), hipHostMallocDefault));
    ctx->rc_dev.alloc(1);
    *ctx->rc_host = rc;
    BTRC_HIP_CHECK(hipMemcpyAsync(ctx->rc_dev.ptr, ctx->rc_host, sizeof(RenderConsts), hipMemcpyHostToDevice, s));
    const RenderConsts *rcd = ctx->rc_dev.ptr;

    StageTimer timer;
    timer.enabled = p->time_stages != 0;
    timer.stream = s;

    ctx->stop = false;
    int64_t captured = 0;
    const auto t0 = std::chrono::steady_clock::now();
    int cur = 0;
    // even: a block of bounces ends on the pool side it began with
    const int check_every = ctx->tuning.graph_bounces > 0 ? ctx->tuning.graph_bounces : 4;

    // one bounce: plan, generate, trace, [medium], shade, shadow. With
    // `overlap` the shadow rays go to the shadow stream and are traced beside
    // the next bounce's plan, generate and trace; the next bounce's medium
    // and shade stages (which write the film, the radiance sums and the
    // shadow queue again) wait for them.
    bool shadow_open = false; // a shadow launch on the shadow stream not yet joined
    auto join_shadow = [&] {
        if(shadow_open)
            BTRC_HIP_CHECK(hipStreamWaitEvent(s, ctx->shadow_done, 0));
        shadow_open = false;
    };
    auto enqueue_bounce = [&](int side, bool timed, bool overlap) {
        PathSOA path = ctx->pool[side].view();
        PathSOA next = ctx->pool[side ^ 1].view();
        launch_plan(s, ctx->wc(), p->state_count, side);
        auto run = [&](int stage, auto &&fn) {
            if(timed)
                timer.run(stage, fn);
            else
                fn();
        };
        run(STAGE_GENERATE, [&] { launch_generate(s, grid256, p->sampler, rcd, ctx->wc(), path); });
        run(STAGE_TRACE, [&] { launch_trace_extension(s, ctx->trace_grid, ctx->trace_sched(false), ts, ctx->wc(), path, hit, ovf); });
        join_shadow();
        if(media)
            run(STAGE_MEDIUM, [&] {
                launch_medium(s, grid256, ctx->cu_count * MEDIUM_TRACK_WAVES_PER_CU, p->sampler, rcd, sv, ctx->wc(),
                              path, hit, ctx->medium_scatter.ptr, next, sh, fv, side);
            });
        run(STAGE_SHADE, [&] { launch_shade(s, grid256, p->sampler, rcd, sv, ctx->wc(), path, hit, next, sh, fv, nv, side); });
        if(overlap)
        {
            BTRC_HIP_CHECK(hipEventRecord(ctx->shaded, s));
            BTRC_HIP_CHECK(hipStreamWaitEvent(ctx->shadow_stream, ctx->shaded, 0));
            launch_shadow(ctx->shadow_stream, ctx->trace_grid, ctx->trace_sched(true), media, ts, sv, ctx->wc(), sh,
                          fv, ovf, side);
            BTRC_HIP_CHECK(hipEventRecord(ctx->shadow_done, ctx->shadow_stream));
            shadow_open = true;
        }
        else
            run(STAGE_SHADOW, [&] { launch_shadow(s, ctx->trace_grid, ctx->trace_sched(true), media, ts, sv, ctx->wc(), sh, fv, ovf, side); });
        launch_end_iteration(s, ctx->wc(), side);
    };

    if(ctx->tuning.launch_graph && !timer.enabled && !capture)
    {
        // blocks of check_every bounces from a captured graph; recaptured
        // when any launch argument differs from the cached graph's
        std::vector<uint8_t> key;
        auto add = [&](const auto &v) {
            const uint8_t *b = reinterpret_cast<const uint8_t *>(&v);
            key.insert(key.end(), b, b + sizeof(v));
        };
        add(ts); add(sv); add(hit); add(sh); add(fv); add(nv); add(rcd);
        add(ovf.code); add(ovf.t); add(ovf.lanes);
        add(ctx->pool[0].view()); add(ctx->pool[1].view()); add(ctx->wc());
        // the shadow stream pays off once a bounce fills the GPU: the event
        // hand-offs cost ~6 us per bounce (256^2 frames: +8 %), the overlap
        // saves ~7 % of a 1080p frame
        const bool overlap = std::min<uint64_t>((uint64_t)p->state_count, total) >= (1u << 20);
        const int32_t ints[9] = { grid256, ctx->trace_grid, ctx->trace_sched(false), ctx->trace_sched(true),
                                  p->sampler, p->state_count, media ? 1 : 0, check_every, overlap ? 1 : 0 };
        add(ints);
        int g = -1;
        for(int k = 0; k < 2; ++k)
            if(ctx->loop_exec[k] && key == ctx->loop_key[k])
                g = k;
        if(g < 0)
        {
            g = ctx->loop_next;
            ctx->loop_next ^= 1;
            if(ctx->loop_exec[g])
                (void)hipGraphExecDestroy(ctx->loop_exec[g]);
            ctx->loop_exec[g] = nullptr;
            ctx->loop_key[g].clear();
            hipGraph_t graph = nullptr;
            BTRC_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for(int k = 0; k < check_every; ++k)
                enqueue_bounce(k & 1, false, overlap);
            join_shadow();
            BTRC_HIP_CHECK(hipStreamEndCapture(s, &graph));
            const hipError_t e = hipGraphInstantiate(&ctx->loop_exec[g], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if(e != hipSuccess)
            {
                ctx->loop_exec[g] = nullptr;
                throw HipError(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            }
            ctx->loop_key[g] = key;
        }
        for(;;)
        {
            if(ctx->stop.load())
                throw Stopped("render stopped");
            BTRC_HIP_CHECK(hipGraphLaunch(ctx->loop_exec[g], s));
            BTRC_HIP_CHECK(hipMemcpyAsync(ctx->host_counters, ctx->wc(), sizeof(WaveCounters), hipMemcpyDeviceToHost, s));
            BTRC_HIP_CHECK(hipStreamSynchronize(s));
            const WaveCounters &hc = *ctx->host_counters;
            reporter_stage.progress(hc);
            if(hc.generated == hc.total && hc.active == 0)
                break;
        }
    }
    else
    for(uint64_t iter = 0;; ++iter)
    {
        if(ctx->stop.load())
            throw Stopped("render stopped");
        PathSOA path = ctx->pool[cur].view();
        PathSOA next = ctx->pool[cur ^ 1].view();
        launch_plan(s, ctx->wc(), p->state_count, cur);
        timer.run(STAGE_GENERATE, [&] { launch_generate(s, grid256, p->sampler, rcd, ctx->wc(), path); });
        timer.run(STAGE_TRACE, [&] { launch_trace_extension(s, ctx->trace_grid, ctx->trace_sched(false), ts, ctx->wc(), path, hit, ovf); });
        if(capture && (int64_t)iter < capture_bounces && captured < capture_max)
        {
            BTRC_HIP_CHECK(hipMemcpyAsync(ctx->host_counters, ctx->wc(), sizeof(WaveCounters), hipMemcpyDeviceToHost, s));
            BTRC_HIP_CHECK(hipStreamSynchronize(s));
            const int64_t n = std::min<int64_t>(ctx->host_counters->active, capture_max - captured);
            std::vector<float4> o(n), d(n);
            if(n > 0)
            {
                // the rays in dense order: segment by segment of the pool side
                QueueState qs;
                BTRC_HIP_CHECK(hipMemcpy(&qs, queue_state(ctx->wc()), sizeof(QueueState), hipMemcpyDeviceToHost));
                for(int k = 0; k < POOL_SEGMENTS; ++k)
                {
                    const int64_t lo = qs.pool_prefix[k], hi = std::min<int64_t>(qs.pool_prefix[k + 1], n);
                    if(hi > lo)
                    {
                        BTRC_HIP_CHECK(hipMemcpy(o.data() + lo, path.ray_o + qs.pool_first[k], (hi - lo) * sizeof(float4),
                                                 hipMemcpyDeviceToHost));
                        BTRC_HIP_CHECK(hipMemcpy(d.data() + lo, path.ray_d + qs.pool_first[k], (hi - lo) * sizeof(float4),
                                                 hipMemcpyDeviceToHost));
                    }
                }
            }
            for(int64_t k = 0; k < n; ++k)
            {
                btrc_ray &r = capture[captured + k];
                r.o[0] = o[k].x; r.o[1] = o[k].y; r.o[2] = o[k].z; r.tmin = 0.0f;
                r.d[0] = d[k].x; r.d[1] = d[k].y; r.d[2] = d[k].z; r.tmax = d[k].w;
            }
            captured += n;
        }
        if(media)
            timer.run(STAGE_MEDIUM, [&] {
                launch_medium(s, grid256, ctx->cu_count * MEDIUM_TRACK_WAVES_PER_CU, p->sampler, rcd, sv, ctx->wc(),
                              path, hit, ctx->medium_scatter.ptr, next, sh, fv, cur);
            });
        timer.run(STAGE_SHADE, [&] { launch_shade(s, grid256, p->sampler, rcd, sv, ctx->wc(), path, hit, next, sh, fv, nv, cur); });
        timer.run(STAGE_SHADOW, [&] { launch_shadow(s, ctx->trace_grid, ctx->trace_sched(true), media, ts, sv, ctx->wc(), sh, fv, ovf, cur); });
        launch_end_iteration(s, ctx->wc(), cur);
        BTRC_HIP_CHECK(hipGetLastError());
        cur ^= 1;
        if((iter + 1) % check_every == 0)
        {
            BTRC_HIP_CHECK(hipMemcpyAsync(ctx->host_counters, ctx->wc(), sizeof(WaveCounters), hipMemcpyDeviceToHost, s));
            BTRC_HIP_CHECK(hipStreamSynchronize(s));
            const WaveCounters &hc = *ctx->host_counters;
            reporter_stage.progress(hc);
            if(hc.generated == hc.total && hc.active == 0)
                break;
        }
    }
    int32_t nrc_steps = 0;
    if(nrc_on)
    {
        // the cache answers the queries of this frame, then learns from its
        // training paths
        auto &n = ctx->nrc;
        // the cache renders one sample per pixel and a path makes at most
        // one query: no two queries share a pixel
        // the weights the previous frame's training writes
        if(ctx->train_pending)
            BTRC_HIP_CHECK(hipStreamWaitEvent(s, ctx->train_done, 0));
        timer.run(STAGE_CACHE, [&] {
            launch_nrc_infer(s, ctx->cu_count, (int64_t)pixels, nrc_render_weights(ctx), nrc_box(ctx), n.queries.ptr,
                             n.counts.ptr, ctx->film.ptr, p->film_fp64 ? ctx->film64.ptr : nullptr,
                             n.acc[n.set].ptr);
        });
        if(n.params.train)
        {
            nrc_steps = std::min(n.params.steps, 64);
            const int set = n.set;
            auto train = [&](hipStream_t ts) {
                launch_nrc_compact(ts, n.rec[set].ptr, n.valid_mask.ptr, n.index.ptr, n.counts.ptr + 2);
                nrc_train(ctx, ts, set, nrc_steps, NRC_BATCH, true);
            };
            if(timer.enabled)
                timer.run(STAGE_TRAIN, [&] { train(s); });
            else
            {
                // on the training stream, after this frame's inference; the
                // render returns without waiting for it
                BTRC_HIP_CHECK(hipEventRecord(ctx->infer_done, s));
                BTRC_HIP_CHECK(hipStreamWaitEvent(ctx->train_stream, ctx->infer_done, 0));
                train(ctx->train_stream);
                BTRC_HIP_CHECK(hipEventRecord(ctx->train_done, ctx->train_stream));
                ctx->train_pending = true;
                n.pending_set = set;
            }
        }
        n.pending_steps = nrc_steps;
        n.stats_pending = true;
        BTRC_HIP_CHECK(hipGetLastError());
    }
    BTRC_HIP_CHECK(hipStreamSynchronize(s));
    const auto t1 = std::chrono::steady_clock::now();
    if(capture_count)
        *capture_count = captured;
    if(stats)
    {
        std::memset(stats, 0, sizeof(*stats));
        const WaveCounters &hc = *ctx->host_counters;
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
        stats->extension_rays = hc.ext_rays;
        stats->shadow_rays = hc.shadow_rays;
        stats->paths = hc.generated;
        stats->iterations = hc.iterations;
        timer.collect(stats->stage_ms, stats->stage_launches);
    }
}

void film_readback(btrc_ctx *ctx, float *rgba, double *rgb64)
{
    if(ctx->film_pixels <= 0)
        throw StateError("no film: render first");
    const int32_t pixels = (int32_t)ctx->film_pixels;
    hipStream_t s = ctx->stream;
    if(rgba)
    {
        ctx->resolved.alloc(pixels);
        launch_resolve(s, RESOLVE_COLOR, ctx->film.ptr, pixels, 1.0f / (float)ctx->film_spp, ctx->resolved.ptr);
        BTRC_HIP_CHECK(hipGetLastError());
        BTRC_HIP_CHECK(hipMemcpyAsync(rgba, ctx->resolved.ptr, pixels * sizeof(float4), hipMemcpyDeviceToHost, s));
    }
    if(rgb64)
    {
        if(!ctx->film_has64)
            throw StateError("the last render had no FP64 film");
        BTRC_HIP_CHECK(hipMemcpyAsync(rgb64, ctx->film64.ptr, 3 * (size_t)pixels * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    BTRC_HIP_CHECK(hipStreamSynchronize(s));
    if(rgb64)
    {
        const double inv = 1.0 / (double)ctx->film_spp;
        for(size_t i = 0; i < 3 * (size_t)pixels; ++i)
            rgb64[i] *= inv;
    }
}

// ACESToneMap::process (postprocess/tonemap.cpp:86-97) on the resolved colour
// film, then the RGBA readback
void tonemap_readback(btrc_ctx *ctx, float exposure, float *rgba)
{
    if(ctx->film_pixels <= 0)
        throw StateError("no film: render first");
    if(!rgba)
        throw InvalidArg("rgba is null");
    if(!std::isfinite(exposure))
        throw InvalidArg("exposure is not finite");
    const int32_t pixels = (int32_t)ctx->film_pixels;
    hipStream_t s = ctx->stream;
    ctx->resolved.alloc(pixels);
    launch_resolve(s, RESOLVE_COLOR, ctx->film.ptr, pixels, 1.0f / (float)ctx->film_spp, ctx->resolved.ptr);
    BTRC_HIP_CHECK(hipGetLastError());
    launch_tonemap_aces(s, ctx->resolved.ptr, pixels, exposure);
    BTRC_HIP_CHECK(hipGetLastError());
    BTRC_HIP_CHECK(hipMemcpyAsync(rgba, ctx->resolved.ptr, pixels * sizeof(float4), hipMemcpyDeviceToHost, s));
    BTRC_HIP_CHECK(hipStreamSynchronize(s));
}

// preview.cpp generate_albedo / generate_normal
void aov_readback(btrc_ctx *ctx, float *albedo, float *normal)
{
    if(ctx->film_pixels <= 0)
        throw StateError("no film: render first");
    if((albedo && !ctx->film_has_albedo) || (normal && !ctx->film_has_normal))
        throw StateError("the last render did not record that output");
    const int32_t pixels = (int32_t)ctx->film_pixels;
    hipStream_t s = ctx->stream;
    ctx->resolved.alloc(pixels);
    const float inv = 1.0f / (float)ctx->film_spp;
    if(albedo)
    {
        launch_resolve(s, RESOLVE_COLOR, ctx->film_albedo.ptr, pixels, inv, ctx->resolved.ptr);
        BTRC_HIP_CHECK(hipGetLastError());
        BTRC_HIP_CHECK(hipMemcpyAsync(albedo, ctx->resolved.ptr, pixels * sizeof(float4), hipMemcpyDeviceToHost, s));
        BTRC_HIP_CHECK(hipStreamSynchronize(s));
    }
    if(normal)
    {
        launch_resolve(s, RESOLVE_NORMAL, ctx->film_normal.ptr, pixels, inv, ctx->resolved.ptr);
        BTRC_HIP_CHECK(hipGetLastError());
        BTRC_HIP_CHECK(hipMemcpyAsync(normal, ctx->resolved.ptr, pixels * sizeof(float4), hipMemcpyDeviceToHost, s));
        BTRC_HIP_CHECK(hipStreamSynchronize(s));
    }
}

void check_batch_args(btrc_ctx *ctx, const void *rays, int64_t n, int32_t tri_test)
{
    if(!ctx->has_scene)
        throw StateError("trace before btrc_hip_upload_scene");
    if(n < 0 || (n > 0 && !rays) || n > 0x7fffffff)
        throw InvalidArg("bad ray batch");
    if(tri_test != BTRC_TRI_WATERTIGHT && tri_test != BTRC_TRI_MOLLER_TRUMBORE)
        throw InvalidArg("unknown triangle test");
}

void trace_batch(btrc_ctx *ctx, int32_t tri_test, bool any_hit, const float4 *rays, int64_t n, uint32_t *out)
{
    BTRC_HIP_CHECK(hipMemsetAsync(ctx->fetch.ptr, 0, FETCH_WORDS * sizeof(uint32_t), ctx->stream));
    launch_trace_batch(ctx->stream, ctx->trace_grid, ctx->trace_sched(any_hit != 0), ctx->trace_scene(), tri_test, any_hit, rays,
                       (int32_t)n, out, ctx->fetch.ptr, ctx->trace_overflow());
    BTRC_HIP_CHECK(hipGetLastError());
}

} // namespace

extern "C" {

int btrc_hip_create(int device, btrc_ctx **out)
{
    if(!out)
        return BTRC_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if(hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return BTRC_ERR_HIP;
    btrc_ctx *ctx = new(std::nothrow) btrc_ctx;
    if(!ctx)
        return BTRC_ERR_OOM;
    ctx->device = device;
    if(hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
       hipStreamCreateWithFlags(&ctx->train_stream, hipStreamNonBlocking) != hipSuccess ||
       hipStreamCreateWithFlags(&ctx->shadow_stream, hipStreamNonBlocking) != hipSuccess ||
       hipEventCreateWithFlags(&ctx->infer_done, hipEventDisableTiming) != hipSuccess ||
       hipEventCreateWithFlags(&ctx->train_done, hipEventDisableTiming) != hipSuccess ||
       hipEventCreateWithFlags(&ctx->shaded, hipEventDisableTiming) != hipSuccess ||
       hipEventCreateWithFlags(&ctx->shadow_done, hipEventDisableTiming) != hipSuccess)
    {
        btrc_hip_destroy(ctx);
        return BTRC_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if(hipGetDeviceProperties(&prop, device) == hipSuccess)
    {
        ctx->cu_count = prop.multiProcessorCount;
        ctx->device_name = prop.gcnArchName;
    }
    resolve_scratch(device, &ctx->noalloc_scratch, &ctx->noalloc_error);
    *out = ctx;
    return BTRC_OK;
}

void btrc_hip_destroy(btrc_ctx *ctx)
{
    if(!ctx)
        return;
    (void)hipSetDevice(ctx->device);
    if(ctx->stream)
        (void)hipStreamSynchronize(ctx->stream);
    if(ctx->train_stream)
        (void)hipStreamSynchronize(ctx->train_stream);
    if(ctx->shadow_stream)
        (void)hipStreamSynchronize(ctx->shadow_stream);
    if(ctx->host_counters)
        (void)hipHostFree(ctx->host_counters);
    if(ctx->rc_host)
        (void)hipHostFree(ctx->rc_host);
    ctx->drop_loop_graph();
    if(ctx->infer_done)
        (void)hipEventDestroy(ctx->infer_done);
    if(ctx->train_done)
        (void)hipEventDestroy(ctx->train_done);
    if(ctx->train_stream)
        (void)hipStreamDestroy(ctx->train_stream);
    for(hipEvent_t e : { ctx->shaded, ctx->shadow_done })
        if(e)
            (void)hipEventDestroy(e);
    if(ctx->shadow_stream)
        (void)hipStreamDestroy(ctx->shadow_stream);
    if(ctx->stream)
        (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *btrc_hip_last_error(const btrc_ctx *ctx)
{
    return ctx ? ctx->last_error.c_str() : "null context";
}

int btrc_hip_device_name(btrc_ctx *ctx, char *buf, size_t len)
{
    return guarded(ctx, [&] {
        if(!buf || len == 0)
            throw InvalidArg("bad buffer");
        std::strncpy(buf, ctx->device_name.c_str(), len - 1);
        buf[len - 1] = 0;
    });
}

int btrc_hip_set_tuning(btrc_ctx *ctx, const btrc_tuning *t)
{
    return guarded(ctx, [&] {
        if(!t || t->trace_waves_per_cu < 1 || t->trace_waves_per_cu > TRACE_WAVES_PER_CU_MAX ||
           t->stage_blocks_per_cu < 0 || t->stage_blocks_per_cu > 64 || t->graph_bounces < 0 ||
           t->graph_bounces > 16 || (t->graph_bounces & 1))
            throw InvalidArg("bad tuning");
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        ctx->tuning = *t;
        if(ctx->has_scene)
            size_traversal(ctx);
    });
}

int btrc_hip_get_tuning(btrc_ctx *ctx, btrc_tuning *t)
{
    return guarded(ctx, [&] {
        if(!t)
            throw InvalidArg("tuning is null");
        *t = ctx->tuning;
    });
}

int btrc_hip_upload_scene(btrc_ctx *ctx, const btrc_scene_desc *scene)
{
    return guarded(ctx, [&] { upload_scene(ctx, scene); });
}

int btrc_hip_set_camera(btrc_ctx *ctx, const btrc_camera_desc *camera)
{
    return guarded(ctx, [&] { set_camera(ctx, camera); });
}

// New vertex data for one geometry of a one-level structure: its shading
// and light-sampling records are rebuilt, the world-space triangle records
// of its instances recomputed and the tree refitted bottom-up (same
// topology, boxes the exact unions of what they hold). The reference builds
// its structures once per commit (Scene::commit, core/scene.cpp:92-239);
// refitting in place serves callers that animate vertices between frames.
void update_geometry(btrc_ctx *ctx, int32_t gi, const btrc_geometry_desc *g)
{
    if(!ctx->has_scene)
        throw StateError("geometry update before btrc_hip_upload_scene");
    if(ctx->two_level)
        throw InvalidArg("geometry updates need a one-level structure (BTRC_ACCEL_FLAT or BTRC_ACCEL_FLAT_GPU)");
    if(gi < 0 || gi >= (int32_t)ctx->geos_host.size() || !g || (g->triangle_count > 0 && !g->positions))
        throw InvalidArg("bad geometry update");
    GeometryRec &gr = ctx->geos_host[gi];
    if((uint32_t)g->triangle_count != gr.prim_count)
        throw InvalidArg("a geometry update must keep the triangle count");
    hipStream_t s = ctx->stream;
    BTRC_HIP_CHECK(hipStreamSynchronize(s));
    const auto t0 = std::chrono::steady_clock::now();
    GeoRecords rec;
    gr.total_area = geometry_records(*g, rec);
    const size_t off = gr.prim_offset, n = gr.prim_count;
    if(n > 0)
    {
        auto put = [&](DevBuf<float4> &dst, const std::vector<float4> &src, size_t scale) {
            BTRC_HIP_CHECK(hipMemcpyAsync(dst.ptr + scale * off, src.data(), src.size() * sizeof(float4),
                                          hipMemcpyHostToDevice, s));
        };
        put(ctx->gx_u_a, rec.gx, 1);
        put(ctx->gy_u_ba, rec.gy, 1);
        put(ctx->gz_u_ca, rec.gz, 1);
        put(ctx->sn_v_a, rec.sa, 1);
        put(ctx->sn_v_ba, rec.sba, 1);
        put(ctx->sn_v_ca, rec.sca, 1);
        put(ctx->light_pos, rec.lpos, 3);
        BTRC_HIP_CHECK(hipMemcpyAsync(ctx->alias.ptr + off, rec.alias.data(), n * sizeof(AliasUnit),
                                      hipMemcpyHostToDevice, s));
    }
    BTRC_HIP_CHECK(hipMemcpyAsync(ctx->geometries.ptr + gi, &gr, sizeof(GeometryRec), hipMemcpyHostToDevice, s));
    if(n > 0 && ctx->triangle_count > 0)
    {
        DevBuf<float> dp;
        dp.upload(g->positions, 9 * n, s);
        launch_update_world_tris(s, ctx->tris.ptr, (uint32_t)ctx->triangle_count, ctx->instances.ptr, gi, dp.ptr);
        BTRC_HIP_CHECK(hipGetLastError());
        refit_bvh4(ctx->nodes.ptr, ctx->refit_order.ptr, ctx->refit_levels, ctx->tris.ptr, s);
        BvhNode4 root;
        BTRC_HIP_CHECK(hipMemcpyAsync(&root, ctx->nodes.ptr, sizeof(root), hipMemcpyDeviceToHost, s));
        BTRC_HIP_CHECK(hipStreamSynchronize(s));
        for(int a = 0; a < 3; ++a)
        {
            ctx->tri_lo[a] = MAX_FLOAT;
            ctx->tri_hi[a] = -MAX_FLOAT;
        }
        for(int j = 0; j < 4; ++j)
        {
            const float lo[3] = { root.lo_x[j], root.lo_y[j], root.lo_z[j] };
            const float hi[3] = { root.hi_x[j], root.hi_y[j], root.hi_z[j] };
            for(int a = 0; a < 3; ++a)
            {
                ctx->tri_lo[a] = std::min(ctx->tri_lo[a], lo[a]);
                ctx->tri_hi[a] = std::max(ctx->tri_hi[a], hi[a]);
            }
        }
        set_world_bounds(ctx);
    }
    BTRC_HIP_CHECK(hipStreamSynchronize(s));
    ctx->drop_loop_graph();
    ctx->refit_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

int btrc_hip_update_geometry(btrc_ctx *ctx, int32_t geometry, const btrc_geometry_desc *desc, double *ms)
{
    return guarded(ctx, [&] {
        update_geometry(ctx, geometry, desc);
        if(ms)
            *ms = ctx->refit_ms;
    });
}

int btrc_hip_scene_build_ms(btrc_ctx *ctx, double *ms)
{
    return guarded(ctx, [&] {
        if(!ctx->has_scene)
            throw StateError("no scene");
        if(!ms)
            throw InvalidArg("null output");
        *ms = ctx->build_ms;
    });
}

int btrc_hip_build_bvh(btrc_ctx *ctx, const float *boxes, int32_t n, int32_t max_leaf_size, const float *refit_boxes,
                       btrc_bvh_node *nodes, uint32_t *primitive_indices, int32_t *node_count, double *build_ms)
{
    return guarded(ctx, [&] {
        if(n <= 0 || !boxes || !nodes || !primitive_indices || !node_count || n > (1 << 28))
            throw InvalidArg("bad bvh build request");
        hipStream_t s = ctx->stream;
        DevBuf<float> db, dr;
        db.upload(boxes, 6 * (size_t)n, s);
        if(refit_boxes)
            dr.upload(refit_boxes, 6 * (size_t)n, s);
        BTRC_HIP_CHECK(hipStreamSynchronize(s));
        const auto t0 = std::chrono::steady_clock::now();
        const PlocResult r = ploc_build(ctx->ploc, db.ptr, (uint32_t)n, max_leaf_size, s);
        BTRC_HIP_CHECK(hipStreamSynchronize(s));
        if(build_ms)
            *build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if(refit_boxes)
            ploc_refit(ctx->ploc, r, dr.ptr, s);
        static_assert(sizeof(btrc_bvh_node) == sizeof(PlocNode), "node layouts differ");
        BTRC_HIP_CHECK(hipMemcpyAsync(nodes, ctx->ploc.nodes, r.node_count * sizeof(PlocNode), hipMemcpyDeviceToHost, s));
        BTRC_HIP_CHECK(hipMemcpyAsync(primitive_indices, ctx->ploc.prims, (size_t)n * sizeof(uint32_t),
                                      hipMemcpyDeviceToHost, s));
        BTRC_HIP_CHECK(hipStreamSynchronize(s));
        *node_count = (int32_t)r.node_count;
    });
}

int btrc_hip_scene_info(btrc_ctx *ctx, int64_t *nodes, int32_t *depth, int32_t *stack_bound, int64_t *triangles)
{
    return guarded(ctx, [&] {
        if(!ctx->has_scene)
            throw StateError("no scene");
        if(nodes) *nodes = ctx->node_count;
        if(depth) *depth = ctx->bvh_depth;
        if(stack_bound) *stack_bound = ctx->stack_size;
        if(triangles) *triangles = ctx->triangle_count;
    });
}

int btrc_hip_render(btrc_ctx *ctx, const btrc_render_params *params, btrc_render_stats *stats)
{
    return guarded(ctx, [&] { render(ctx, params, 0, nullptr, stats, 0, nullptr, 0, nullptr); });
}

int btrc_hip_render_views(btrc_ctx *ctx, const btrc_render_params *params, int32_t view_count,
                          const btrc_view_desc *views, btrc_render_stats *stats)
{
    return guarded(ctx, [&] {
        if(!views)
            throw InvalidArg("views is null");
        render(ctx, params, view_count, views, stats, 0, nullptr, 0, nullptr);
    });
}

int btrc_hip_view_readback(btrc_ctx *ctx, int32_t view, float *rgba)
{
    return guarded(ctx, [&] {
        if(view < 0 || view >= (int32_t)ctx->film_views.size() || !rgba)
            throw InvalidArg("bad view readback");
        const ViewRec &v = ctx->film_views[view];
        const int32_t pixels = v.width * v.height;
        hipStream_t s = ctx->stream;
        ctx->resolved.alloc(pixels);
        launch_resolve(s, RESOLVE_COLOR, ctx->film.ptr + 4 * (size_t)v.pixel_base, pixels,
                       1.0f / (float)ctx->film_spp, ctx->resolved.ptr);
        BTRC_HIP_CHECK(hipGetLastError());
        BTRC_HIP_CHECK(hipMemcpyAsync(rgba, ctx->resolved.ptr, pixels * sizeof(float4), hipMemcpyDeviceToHost, s));
        BTRC_HIP_CHECK(hipStreamSynchronize(s));
    });
}

int btrc_hip_capture_rays(btrc_ctx *ctx, const btrc_render_params *params, int32_t bounces,
                          btrc_ray *rays, int64_t max_rays, int64_t *count)
{
    return guarded(ctx, [&] {
        if(!rays || max_rays < 0 || bounces < 0)
            throw InvalidArg("bad capture buffer");
        render(ctx, params, 0, nullptr, nullptr, bounces, rays, max_rays, count);
    });
}

int btrc_hip_film_readback(btrc_ctx *ctx, float *rgba, double *rgb64)
{
    return guarded(ctx, [&] { film_readback(ctx, rgba, rgb64); });
}

int btrc_hip_tonemap_readback(btrc_ctx *ctx, float exposure, float *rgba)
{
    return guarded(ctx, [&] { tonemap_readback(ctx, exposure, rgba); });
}

int btrc_hip_aov_readback(btrc_ctx *ctx, float *albedo_rgba, float *normal_rgba)
{
    return guarded(ctx, [&] { aov_readback(ctx, albedo_rgba, normal_rgba); });
}

int btrc_hip_set_filter(btrc_ctx *ctx, const btrc_filter_desc *filter)
{
    return guarded(ctx, [&] { set_filter(ctx, filter); });
}

int btrc_hip_nrc_configure(btrc_ctx *ctx, const btrc_nrc_params *params)
{
    return guarded(ctx, [&] {
        if(!params)
            throw InvalidArg("cache parameters are null");
        const btrc_nrc_params &q = *params;
        if(q.query_depth < 0 || q.train_stride < 1 || q.steps < 0 || q.steps > 64 || !(q.learning_rate > 0.0f) ||
           !(q.loss_scale > 0.0f) || !std::isfinite(q.learning_rate) || !std::isfinite(q.loss_scale) ||
           q.train_depth < 0 || (q.train_depth > 0 && q.train_depth < q.query_depth) || !(q.ema_decay >= 0.0f) ||
           !(q.ema_decay < 1.0f))
            throw InvalidArg("bad cache parameters");
        nrc_sync(ctx);
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        ctx->nrc.params = q;
        if(q.enabled)
            nrc_require(ctx);
    });
}

int btrc_hip_nrc_reset(btrc_ctx *ctx, uint64_t seed)
{
    return guarded(ctx, [&] { nrc_reset(ctx, seed); });
}

int btrc_hip_nrc_stats(btrc_ctx *ctx, btrc_nrc_stats *stats)
{
    return guarded(ctx, [&] {
        if(!stats)
            throw InvalidArg("stats is null");
        nrc_sync(ctx);
        *stats = ctx->nrc.last;
    });
}

int btrc_hip_nrc_get_weights(btrc_ctx *ctx, float *weights)
{
    return guarded(ctx, [&] {
        if(!weights)
            throw InvalidArg("weights is null");
        nrc_require(ctx);
        nrc_sync(ctx);
        BTRC_HIP_CHECK(hipMemcpyAsync(weights, ctx->nrc.w.ptr, NRC_PARAMS * sizeof(float), hipMemcpyDeviceToHost,
                                      ctx->stream));
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int btrc_hip_nrc_get_ema_weights(btrc_ctx *ctx, float *weights)
{
    return guarded(ctx, [&] {
        if(!weights)
            throw InvalidArg("weights is null");
        nrc_require(ctx);
        nrc_sync(ctx);
        BTRC_HIP_CHECK(hipMemcpyAsync(weights, ctx->nrc.ema.ptr, NRC_PARAMS * sizeof(float), hipMemcpyDeviceToHost,
                                      ctx->stream));
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int btrc_hip_nrc_set_weights(btrc_ctx *ctx, const float *weights)
{
    return guarded(ctx, [&] {
        if(!weights)
            throw InvalidArg("weights is null");
        nrc_require(ctx);
        nrc_sync(ctx);
        // the output layer's padding rows stay zero
        std::vector<float> w(weights, weights + NRC_PARAMS);
        for(int k = nrc_layer_offset(NRC_LAYERS - 1) + NRC_OUT * NRC_IN; k < NRC_PARAMS; ++k)
            w[k] = 0.0f;
        ctx->nrc.w.upload(w.data(), w.size(), ctx->stream);
        launch_nrc_publish(ctx->stream, nrc_weights(ctx));
        BTRC_HIP_CHECK(hipGetLastError());
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int btrc_hip_nrc_eval(btrc_ctx *ctx, const float *vertices, int32_t n, float *out)
{
    return guarded(ctx, [&] {
        if(n < 0 || (n > 0 && (!vertices || !out)))
            throw InvalidArg("bad cache query");
        nrc_require(ctx);
        nrc_sync(ctx);
        if(n == 0)
            return;
        DevBuf<NrcVertex> dv;
        DevBuf<float> dout;
        dv.upload(reinterpret_cast<const NrcVertex *>(vertices), n, ctx->stream);
        dout.alloc(3 * (size_t)n);
        launch_nrc_eval(ctx->stream, nrc_weights(ctx), nrc_box(ctx), dv.ptr, n, dout.ptr);
        BTRC_HIP_CHECK(hipGetLastError());
        BTRC_HIP_CHECK(hipMemcpyAsync(out, dout.ptr, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

// records whose target is `targets`: throughput 1, nothing gathered before,
// one training path per record carrying the target as its radiance sum
void nrc_load_batch(btrc_ctx *ctx, const float *vertices, const float *targets, int32_t n)
{
    auto &nc = ctx->nrc;
    std::vector<NrcRecord> rec(NRC_RECORDS);
    std::vector<float> acc(4 * (size_t)NRC_RECORDS, 0.0f);
    for(int32_t i = 0; i < n; ++i)
    {
        NrcRecord &r = rec[i];
        std::memcpy(&r.v, vertices + 16 * (size_t)i, sizeof(NrcVertex));
        r.beta[0] = r.beta[1] = r.beta[2] = 1.0f;
        r.gathered[0] = r.gathered[1] = r.gathered[2] = 0.0f;
        r.path = (uint32_t)i;
        r.valid = 1;
        for(int c = 0; c < 3; ++c)
            acc[4 * (size_t)i + c] = targets[3 * (size_t)i + c];
    }
    nc.rec[0].upload(rec.data(), rec.size(), ctx->stream);
    nc.acc[0].upload(acc.data(), acc.size(), ctx->stream);
}

int btrc_hip_nrc_gradients(btrc_ctx *ctx, const float *vertices, const float *targets, int32_t n, float *grad,
                           double *loss)
{
    return guarded(ctx, [&] {
        if(n <= 0 || n > NRC_RECORDS || !vertices || !targets || !grad)
            throw InvalidArg("bad cache batch");
        nrc_require(ctx);
        nrc_sync(ctx);
        nrc_load_batch(ctx, vertices, targets, n);
        auto &nc = ctx->nrc;
        const int64_t batch = ((int64_t)n + 63) / 64 * 64;
        NrcBatch b{ nc.rec[0].ptr, 0, batch, nc.acc[0].ptr, nc.counts.ptr + 1, nc.params.loss_scale, 0, nullptr, nullptr,
                    0, 1 };
        launch_nrc_gradients(ctx->stream, nrc_weights(ctx), nrc_box(ctx), b, nrc_scratch(ctx, batch), nc.grad.ptr);
        BTRC_HIP_CHECK(hipGetLastError());
        std::vector<double> partial(batch / 32);
        BTRC_HIP_CHECK(hipMemcpyAsync(grad, nc.grad.ptr, NRC_PARAMS * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        BTRC_HIP_CHECK(hipMemcpyAsync(partial.data(), nc.loss_partial.ptr, partial.size() * sizeof(double),
                                      hipMemcpyDeviceToHost, ctx->stream));
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        // the kernels sum the loss-scaled per-sample gradients
        const double scale = 1.0 / (3.0 * (double)n * (double)nc.params.loss_scale);
        for(int k = 0; k < NRC_PARAMS; ++k)
            grad[k] = (float)((double)grad[k] * scale);
        if(loss)
        {
            double sum = 0.0;
            for(double v : partial)
                sum += v;
            *loss = sum / (3.0 * (double)n);
        }
    });
}

int btrc_hip_nrc_fit(btrc_ctx *ctx, const float *vertices, const float *targets, int32_t n, int32_t steps,
                     double *loss)
{
    return guarded(ctx, [&] {
        if(n <= 0 || n > NRC_BATCH || steps < 0 || steps > 64 || !vertices || !targets)
            throw InvalidArg("bad cache batch");
        nrc_require(ctx);
        nrc_sync(ctx);
        nrc_load_batch(ctx, vertices, targets, n);
        nrc_train(ctx, ctx->stream, 0, steps, NRC_BATCH, false);
        nrc_collect(ctx, steps);
        if(loss)
        {
            std::vector<NrcStepStats> st(steps);
            if(steps > 0)
                BTRC_HIP_CHECK(hipMemcpy(st.data(), ctx->nrc.step_stats.ptr, steps * sizeof(NrcStepStats),
                                         hipMemcpyDeviceToHost));
            for(int32_t k = 0; k < steps; ++k)
                loss[k] = st[k].loss;
        }
    });
}

int btrc_hip_set_reporter(btrc_ctx *ctx, const btrc_reporter *reporter)
{
    return guarded(ctx, [&] {
        ctx->reporter = reporter ? *reporter : btrc_reporter{ nullptr, nullptr, nullptr, nullptr };
    });
}

int btrc_hip_request_stop(btrc_ctx *ctx)
{
    if(!ctx)
        return BTRC_ERR_INVALID;
    ctx->stop = true;
    return BTRC_OK;
}

int btrc_hip_trace_closest(btrc_ctx *ctx, const btrc_ray *rays, int64_t n, int32_t tri_test, btrc_hit *hits)
{
    return guarded(ctx, [&] {
        check_batch_args(ctx, rays, n, tri_test);
        if(n > 0 && !hits)
            throw InvalidArg("hits is null");
        if(n == 0)
            return;
        DevBuf<float4> r;
        DevBuf<uint32_t> o;
        r.upload(reinterpret_cast<const float4 *>(rays), 2 * (size_t)n, ctx->stream);
        o.alloc(5 * (size_t)n);
        trace_batch(ctx, tri_test, false, r.ptr, n, o.ptr);
        BTRC_HIP_CHECK(hipGetLastError());
        BTRC_HIP_CHECK(hipMemcpyAsync(hits, o.ptr, 5 * (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int btrc_hip_trace_any(btrc_ctx *ctx, const btrc_ray *rays, int64_t n, int32_t tri_test, uint8_t *occluded)
{
    return guarded(ctx, [&] {
        check_batch_args(ctx, rays, n, tri_test);
        if(n > 0 && !occluded)
            throw InvalidArg("occluded is null");
        if(n == 0)
            return;
        DevBuf<float4> r;
        DevBuf<uint32_t> o;
        r.upload(reinterpret_cast<const float4 *>(rays), 2 * (size_t)n, ctx->stream);
        o.alloc((size_t)n);
        trace_batch(ctx, tri_test, true, r.ptr, n, o.ptr);
        BTRC_HIP_CHECK(hipGetLastError());
        std::vector<uint32_t> tmp(n);
        BTRC_HIP_CHECK(hipMemcpyAsync(tmp.data(), o.ptr, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for(int64_t i = 0; i < n; ++i)
            occluded[i] = (uint8_t)tmp[i];
    });
}

int btrc_hip_trace_closest_noalloc(btrc_ctx *ctx, const btrc_ray *rays, int64_t n, int32_t tri_test, btrc_hit *hits)
{
    return guarded(ctx, [&] {
        check_batch_args(ctx, rays, n, tri_test);
        if(n > 0 && !hits)
            throw InvalidArg("hits is null");
        if(n == 0)
            return;
        if(!ctx->noalloc_scratch)
            throw StateError(ctx->noalloc_error);
        // rays and hits must be device allocations that hold n entries
        auto check_range = [&](const void *p, size_t bytes, const char *what) {
            hipPointerAttribute_t attr;
            if(hipPointerGetAttributes(&attr, p) != hipSuccess ||
               (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged))
            {
                (void)hipGetLastError();
                throw InvalidArg(std::string(what) + " is not device memory");
            }
            hipDeviceptr_t base = nullptr;
            size_t size = 0;
            if(hipMemGetAddressRange(&base, &size, const_cast<void *>(p)) != hipSuccess)
            {
                (void)hipGetLastError();
                throw InvalidArg(std::string(what) + ": allocation range unknown");
            }
            const uintptr_t b = reinterpret_cast<uintptr_t>(base), q = reinterpret_cast<uintptr_t>(p);
            if(q < b || q + bytes > b + size)
                throw InvalidArg(std::string(what) + " allocation is smaller than n entries");
        };
        check_range(rays, (size_t)n * sizeof(btrc_ray), "rays");
        check_range(hits, (size_t)n * sizeof(btrc_hit), "hits");
        // the persistent grid, capped so that its stack entries fit the scratch
        const size_t deep = ctx->stack_size > SHORT_STACK ? (size_t)(ctx->stack_size - SHORT_STACK) : 0;
        const size_t per_wave = deep * TRACE_BLOCK * (sizeof(uint32_t) + sizeof(float));
        const size_t room = NOALLOC_SCRATCH_BYTES - NOALLOC_COUNTER_BYTES;
        int waves = ctx->trace_grid;
        if(per_wave > 0 && (size_t)waves * per_wave > room)
            waves = (int)(room / per_wave);
        if(waves < 1)
            throw InvalidArg("scene tree too deep for the no-allocation scratch: " + std::to_string(ctx->stack_size) +
                             " stack entries need " + std::to_string(per_wave) + " bytes per wavefront beyond the " +
                             std::to_string(room) + " the scratch holds");
        uint8_t *base = static_cast<uint8_t *>(ctx->noalloc_scratch);
        uint32_t *fetch = reinterpret_cast<uint32_t *>(base);
        const size_t lanes = (size_t)waves * TRACE_BLOCK;
        uint32_t *ovf_code = reinterpret_cast<uint32_t *>(base + NOALLOC_COUNTER_BYTES);
        float *ovf_t = reinterpret_cast<float *>(base + NOALLOC_COUNTER_BYTES + deep * lanes * sizeof(uint32_t));
        launch_fill_u32(ctx->stream, fetch, FETCH_WORDS, 0u);
        BTRC_HIP_CHECK(hipGetLastError());
        launch_trace_batch(ctx->stream, waves, ctx->trace_sched(false), ctx->trace_scene(), tri_test, false,
                           reinterpret_cast<const float4 *>(rays), (int32_t)n, reinterpret_cast<uint32_t *>(hits),
                           fetch, TraceOverflow{ ovf_code, ovf_t, (int32_t)lanes });
        BTRC_HIP_CHECK(hipGetLastError());
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int btrc_hip_batch_upload(btrc_ctx *ctx, const btrc_ray *rays, int64_t n)
{
    return guarded(ctx, [&] {
        if(n < 0 || (n > 0 && !rays))
            throw InvalidArg("bad ray batch");
        ctx->batch_rays.upload(reinterpret_cast<const float4 *>(rays), 2 * (size_t)n, ctx->stream);
        ctx->batch_out.alloc(5 * (size_t)std::max<int64_t>(n, 1));
        ctx->batch_n = n;
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int btrc_hip_batch_trace(btrc_ctx *ctx, int32_t tri_test, int32_t any_hit, int32_t repeats, double *kernel_ms)
{
    return guarded(ctx, [&] {
        check_batch_args(ctx, ctx->batch_rays.ptr, ctx->batch_n, tri_test);
        if(repeats <= 0)
            throw InvalidArg("repeats must be positive");
        hipEvent_t a, b;
        BTRC_HIP_CHECK(hipEventCreate(&a));
        BTRC_HIP_CHECK(hipEventCreate(&b));
        double total = 0;
        for(int32_t k = 0; k < repeats; ++k)
        {
            BTRC_HIP_CHECK(hipEventRecord(a, ctx->stream));
            trace_batch(ctx, tri_test, any_hit != 0, ctx->batch_rays.ptr, ctx->batch_n, ctx->batch_out.ptr);
            BTRC_HIP_CHECK(hipEventRecord(b, ctx->stream));
            BTRC_HIP_CHECK(hipEventSynchronize(b));
            float ms = 0;
            BTRC_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
            total += ms;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if(kernel_ms)
            *kernel_ms = total / repeats;
    });
}

int btrc_hip_measure_copy(btrc_ctx *ctx, int64_t bytes, int32_t repeats, double *gb_per_s)
{
    return guarded(ctx, [&] {
        if(bytes <= 0 || bytes % 16 != 0)
            throw InvalidArg("bytes must be a positive multiple of 16");
        if(repeats <= 0)
            throw InvalidArg("repeats must be positive");
        if(!gb_per_s)
            throw InvalidArg("gb_per_s is null");
        const int64_t vectors = bytes / 16;
        DevBuf<float4> src, dst;
        src.alloc((size_t)vectors);
        dst.alloc((size_t)vectors);
        BTRC_HIP_CHECK(hipMemsetAsync(src.ptr, 0, (size_t)bytes, ctx->stream));
        hipEvent_t a, b;
        BTRC_HIP_CHECK(hipEventCreate(&a));
        BTRC_HIP_CHECK(hipEventCreate(&b));
        std::vector<float> ms;
        const int grid = ctx->cu_count * 8;
        for(int32_t k = 0; k <= repeats; ++k) // the first copy is not timed
        {
            BTRC_HIP_CHECK(hipEventRecord(a, ctx->stream));
            launch_stream_copy(ctx->stream, grid, src.ptr, dst.ptr, vectors);
            BTRC_HIP_CHECK(hipGetLastError());
            BTRC_HIP_CHECK(hipEventRecord(b, ctx->stream));
            BTRC_HIP_CHECK(hipEventSynchronize(b));
            float t = 0;
            BTRC_HIP_CHECK(hipEventElapsedTime(&t, a, b));
            if(k > 0)
                ms.push_back(t);
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        std::sort(ms.begin(), ms.end());
        const double med = ms.size() % 2 ? ms[ms.size() / 2] : 0.5 * (ms[ms.size() / 2 - 1] + ms[ms.size() / 2]);
        *gb_per_s = 2.0 * (double)bytes / (med * 1e-3) / 1e9;
    });
}

int btrc_hip_batch_download(btrc_ctx *ctx, btrc_hit *hits)
{
    return guarded(ctx, [&] {
        if(!hits)
            throw InvalidArg("hits is null");
        BTRC_HIP_CHECK(hipMemcpy(hits, ctx->batch_out.ptr, 5 * (size_t)ctx->batch_n * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost));
    });
}

int btrc_hip_eval_halton(btrc_ctx *ctx, int32_t width, int32_t height, const uint32_t *px, const uint32_t *py,
                         const int32_t *sample_index, int32_t n, int32_t dims, float *out)
{
    return guarded(ctx, [&] {
        if(n < 0 || dims <= 0 || width <= 0 || height <= 0 || (n > 0 && (!px || !py || !sample_index || !out)))
            throw InvalidArg("bad halton query");
        if(n == 0)
            return;
        DevBuf<uint32_t> dx, dy;
        DevBuf<int32_t> ds;
        DevBuf<float> dout;
        dx.upload(px, n, ctx->stream);
        dy.upload(py, n, ctx->stream);
        ds.upload(sample_index, n, ctx->stream);
        dout.alloc((size_t)n * dims);
        launch_halton_eval(ctx->stream, make_halton_consts(width, height), halton_tables(ctx), dx.ptr, dy.ptr, ds.ptr,
                           n, dims, dout.ptr);
        BTRC_HIP_CHECK(hipGetLastError());
        BTRC_HIP_CHECK(hipMemcpyAsync(out, dout.ptr, (size_t)n * dims * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int btrc_hip_eval_independent(btrc_ctx *ctx, const uint32_t *px, const uint32_t *py, const int32_t *sample_index,
                              int32_t n, int32_t dims, float *out)
{
    return guarded(ctx, [&] {
        if(n < 0 || dims <= 0 || (n > 0 && (!px || !py || !sample_index || !out)))
            throw InvalidArg("bad sampler query");
        if(n == 0)
            return;
        DevBuf<uint32_t> dx, dy;
        DevBuf<int32_t> ds;
        DevBuf<float> dout;
        dx.upload(px, n, ctx->stream);
        dy.upload(py, n, ctx->stream);
        ds.upload(sample_index, n, ctx->stream);
        dout.alloc((size_t)n * dims);
        launch_independent_eval(ctx->stream, dx.ptr, dy.ptr, ds.ptr, n, dims, dout.ptr);
        BTRC_HIP_CHECK(hipGetLastError());
        BTRC_HIP_CHECK(hipMemcpyAsync(out, dout.ptr, (size_t)n * dims * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int btrc_hip_eval_volumes(btrc_ctx *ctx, const btrc_ray *segments, const uint64_t *seeds, int32_t n, float *out)
{
    return guarded(ctx, [&] {
        if(!ctx->has_scene)
            throw StateError("volume query before btrc_hip_upload_scene");
        if(n < 0 || (n > 0 && (!segments || !seeds || !out)))
            throw InvalidArg("bad volume query");
        if(n == 0)
            return;
        DevBuf<float4> dseg;
        DevBuf<uint2> dseed;
        DevBuf<float> dout;
        dseg.upload(reinterpret_cast<const float4 *>(segments), 2 * (size_t)n, ctx->stream);
        dseed.upload(reinterpret_cast<const uint2 *>(seeds), n, ctx->stream);
        dout.alloc(10 * (size_t)n);
        launch_volume_eval(ctx->stream, ctx->scene_view().vol, dseg.ptr, dseed.ptr, n, dout.ptr);
        BTRC_HIP_CHECK(hipGetLastError());
        BTRC_HIP_CHECK(hipMemcpyAsync(out, dout.ptr, 10 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int btrc_hip_eval_texture(btrc_ctx *ctx, int32_t texture, const float *uv, int32_t n, float *out)
{
    return guarded(ctx, [&] {
        if(!ctx->has_scene)
            throw StateError("texture query before btrc_hip_upload_scene");
        if(texture < 0 || texture >= ctx->texture_count || n < 0 || (n > 0 && (!uv || !out)))
            throw InvalidArg("bad texture query");
        if(n == 0)
            return;
        DevBuf<float2> duv;
        DevBuf<float4> dout;
        duv.upload(reinterpret_cast<const float2 *>(uv), n, ctx->stream);
        dout.alloc(n);
        launch_texture_eval(ctx->stream, ctx->scene_view().tex, texture, duv.ptr, n, dout.ptr);
        BTRC_HIP_CHECK(hipGetLastError());
        BTRC_HIP_CHECK(hipMemcpyAsync(out, dout.ptr, 4 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int btrc_hip_eval_envir(btrc_ctx *ctx, const float *dirs, const float *samples, int32_t n, float *out)
{
    return guarded(ctx, [&] {
        if(!ctx->has_scene)
            throw StateError("environment query before btrc_hip_upload_scene");
        if(ctx->envir_light_index < 0)
            throw InvalidArg("the scene has no environment light");
        if(n < 0 || (n > 0 && (!dirs || !samples || !out)))
            throw InvalidArg("bad environment query");
        if(n == 0)
            return;
        DevBuf<float> dd, ds, dout;
        dd.upload(dirs, 3 * (size_t)n, ctx->stream);
        ds.upload(samples, 3 * (size_t)n, ctx->stream);
        dout.alloc(12 * (size_t)n);
        launch_envir_eval(ctx->stream, ctx->scene_view(), dd.ptr, ds.ptr, n, dout.ptr);
        BTRC_HIP_CHECK(hipGetLastError());
        BTRC_HIP_CHECK(hipMemcpyAsync(out, dout.ptr, 12 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost,
                                      ctx->stream));
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int btrc_hip_envir_table(btrc_ctx *ctx, float *probs, int32_t capacity, int32_t *count)
{
    return guarded(ctx, [&] {
        if(!ctx->has_scene)
            throw StateError("environment query before btrc_hip_upload_scene");
        if(!count || capacity < 0 || (capacity > 0 && !probs))
            throw InvalidArg("bad environment table query");
        const int32_t n = (int32_t)ctx->env_probs_host.size();
        *count = n;
        std::memcpy(probs, ctx->env_probs_host.data(), sizeof(float) * (size_t)std::min(n, capacity));
    });
}

int btrc_hip_eval_medium(btrc_ctx *ctx, int32_t medium, const btrc_ray *segments, const uint64_t *seeds, int32_t n,
                         float *out)
{
    return guarded(ctx, [&] {
        if(!ctx->has_scene)
            throw StateError("medium query before btrc_hip_upload_scene");
        if(medium < 0 || medium >= ctx->medium_count || n < 0 || (n > 0 && (!segments || !seeds || !out)))
            throw InvalidArg("bad medium query");
        if(n == 0)
            return;
        DevBuf<float4> dseg;
        DevBuf<uint2> dseed;
        DevBuf<float> dout;
        dseg.upload(reinterpret_cast<const float4 *>(segments), 2 * (size_t)n, ctx->stream);
        dseed.upload(reinterpret_cast<const uint2 *>(seeds), n, ctx->stream);
        dout.alloc(10 * (size_t)n);
        launch_medium_eval(ctx->stream, ctx->scene_view(), medium, dseg.ptr, dseed.ptr, n, dout.ptr);
        BTRC_HIP_CHECK(hipGetLastError());
        BTRC_HIP_CHECK(hipMemcpyAsync(out, dout.ptr, 10 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int btrc_hip_eval_detmath(btrc_ctx *ctx, int32_t fn, const float *x, const float *y, int32_t n, float *out)
{
    return guarded(ctx, [&] {
        if(n < 0 || fn < 0 || fn > 9 || (n > 0 && (!x || !y || !out)))
            throw InvalidArg("bad detmath query");
        if(n == 0)
            return;
        DevBuf<float> dx, dy, dout;
        dx.upload(x, n, ctx->stream);
        dy.upload(y, n, ctx->stream);
        dout.alloc(n);
        launch_detmath_eval(ctx->stream, fn, dx.ptr, dy.ptr, n, dout.ptr);
        BTRC_HIP_CHECK(hipGetLastError());
        BTRC_HIP_CHECK(hipMemcpyAsync(out, dout.ptr, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        BTRC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

} // extern "C"
