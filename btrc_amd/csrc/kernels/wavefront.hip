// Wavefront path-tracing stages for gfx950.
//
// Stage semantics follow the reference kernels: generate_kernel
// (renderer/wavefront/generate.cpp:33-87, uniform mode), the OptiX trace
// programs (trace.cpp:38-78), shade_kernel (shade.cpp:29-236) and the shadow
// programs (shadow.cpp:24-76). Work counts are read from device memory and
// every kernel is a grid-stride loop over a persistent grid, so the host
// launches a fixed sequence per bounce and never waits on the GPU between
// bounces. Queue pushes are aggregated per wavefront: one atomic per wave.
#include "wavefront.h"

namespace btrc {

__constant__ PrimeTable c_primes = make_prime_table();

namespace {

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Wave-aggregated slot reservation: every lane of the wavefront calls this
// with its predicate; lanes with pred receive consecutive slots.
__device__ __forceinline__ int32_t wave_push(int32_t *counter, bool pred)
{
    const uint64_t mask = __ballot(pred);
    if(mask == 0)
        return -1;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    int32_t base = 0;
    if(lane_id() == leader)
        base = atomicAdd(counter, (int32_t)__popcll(mask));
    base = __shfl(base, leader);
    const uint64_t below = mask & ((1ull << lane_id()) - 1ull);
    return pred ? base + (int32_t)__popcll(below) : -1;
}

// Wave-aggregated push into stripe k of a striped queue (counters k *
// COUNTER_STRIDE apart): lanes with pred receive consecutive slots from
// k * stripe on.
__device__ __forceinline__ int32_t stripe_push(int32_t *counters, int32_t k, int32_t stripe, bool pred)
{
    const uint64_t mask = __ballot(pred);
    if(mask == 0)
        return -1;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    int32_t base = 0;
    if(lane_id() == leader)
        base = atomicAdd(counters + k * COUNTER_STRIDE, (int32_t)__popcll(mask));
    base = __shfl(base, leader);
    const uint64_t below = mask & ((1ull << lane_id()) - 1ull);
    return pred ? k * stripe + base + (int32_t)__popcll(below) : -1;
}

__device__ __forceinline__ void film_splat(const FilmView &film, uint32_t pixel, V3 v)
{
    // Film::splat_atomic drops non-finite values (core/film.cpp:145-163)
    if(!isfinite3(v))
        return;
    float *p = film.rgb + 4 * (size_t)pixel;
    atomicAdd(p + 0, v.x);
    atomicAdd(p + 1, v.y);
    atomicAdd(p + 2, v.z);
    if(film.rgb64)
    {
        double *q = film.rgb64 + 3 * (size_t)pixel;
        atomicAdd(q + 0, (double)v.x);
        atomicAdd(q + 1, (double)v.y);
        atomicAdd(q + 2, (double)v.z);
    }
}

__device__ __forceinline__ void aov_splat(float *buf, uint32_t pixel, V3 v)
{
    if(!isfinite3(v))
        return;
    float *p = buf + 4 * (size_t)pixel;
    atomicAdd(p + 0, v.x);
    atomicAdd(p + 1, v.y);
    atomicAdd(p + 2, v.z);
}

__device__ __forceinline__ V3 xyz(float4 a) { return v3(a.x, a.y, a.z); }
__device__ __forceinline__ float4 f4(V3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// The sampler of a path, kept in PathSOA::smp between stages (x = pixel).
template<int SAMPLER> struct PathSampler;

template<> struct PathSampler<SAMPLER_HALTON>
{
    HaltonState st;

    __device__ static PathSampler start(const ViewRec &view, uint32_t px, uint32_t py, int32_t sample)
    {
        return PathSampler{ halton_start(view.halton, px, py, sample) };
    }
    __device__ static PathSampler load(const RenderConsts &, uint4 s)
    {
        PathSampler r;
        r.st.dimension = s.y;
        r.st.index = (uint64_t)s.z | ((uint64_t)s.w << 32);
        return r;
    }
    __device__ uint4 pack(uint32_t pixel) const
    {
        return make_uint4(pixel, st.dimension, (uint32_t)st.index, (uint32_t)(st.index >> 32));
    }
    __device__ float get1d() { return halton_get1d(c_primes.p, st); }
    __device__ void get3d(float *a, float *b, float *c) { halton_get3d(c_primes.p, st, a, b, c); }
};

// Sampler::get2d/get3d default to consecutive get1d (core/sampler.h:15-28);
// the stream increment depends only on the pixel, so it is rederived on load
template<> struct PathSampler<SAMPLER_INDEPENDENT>
{
    Pcg32 rng;

    __device__ static PathSampler start(const ViewRec &, uint32_t px, uint32_t py, int32_t sample)
    {
        return PathSampler{ independent_start(px, py, sample) };
    }
    __device__ static PathSampler load(const RenderConsts &rc, uint4 s)
    {
        const ViewRec &view = view_of_pixel(rc, s.x);
        const uint32_t local = s.x - view.pixel_base;
        const uint32_t px = local % (uint32_t)view.width, py = local / (uint32_t)view.width;
        PathSampler r;
        r.rng.inc = (murmur_hash_pixel(px, py) << 1) | 1ull;
        r.rng.state = (uint64_t)s.y | ((uint64_t)s.z << 32);
        return r;
    }
    __device__ uint4 pack(uint32_t pixel) const
    {
        return make_uint4(pixel, (uint32_t)rng.state, (uint32_t)(rng.state >> 32), 0u);
    }
    __device__ float get1d() { return pcg32_uniform_float(rng); }
    __device__ void get3d(float *a, float *b, float *c)
    {
        *a = get1d();
        *b = get1d();
        *c = get1d();
    }
};

// generate's sampler: the Halton sampler reads the camera samples' first
// digits from the first-digit tables (same values). The later stages run
// the digit loop alone: in shade and k_medium_scatter the lookups cost more
// registers than they save, and in k_medium_track (two draws per tracking
// step) the dependent table loads cost more time than the digits they save.
template<int SAMPLER> struct TabledSampler
{
    PathSampler<SAMPLER> smp;

    __device__ static TabledSampler start(const RenderConsts &, const ViewRec &view, uint32_t px, uint32_t py,
                                          int32_t sample)
    {
        return TabledSampler{ PathSampler<SAMPLER>::start(view, px, py, sample) };
    }
    __device__ float get1d() { return smp.get1d(); }
    __device__ uint4 pack(uint32_t pixel) const { return smp.pack(pixel); }
};

template<> struct TabledSampler<SAMPLER_HALTON>
{
    HaltonState st;
    HaltonTables tab;

    __device__ static TabledSampler start(const RenderConsts &rc, const ViewRec &view, uint32_t px, uint32_t py,
                                          int32_t sample)
    {
        return TabledSampler{ halton_start(view.halton, px, py, sample), rc.halton };
    }
    __device__ float get1d() { return halton_get1d(c_primes.p, st, tab); }
    __device__ uint4 pack(uint32_t pixel) const
    {
        return PathSampler<SAMPLER_HALTON>{ st }.pack(pixel);
    }
};

// Seed of a shadow ray's transmittance stream: a hash of its path's sampler
// state when the ray is made, so ratio tracking along the ray is a function
// of the path alone.
__device__ __forceinline__ uint2 shadow_seed(uint4 s)
{
    const uint64_t h = mix_bits((((uint64_t)s.x << 32) | s.y) ^ mix_bits(((uint64_t)s.z << 32) | s.w));
    return make_uint2((uint32_t)h, (uint32_t)(h >> 32));
}

// PCG32 stream of a shadow ray's ratio tracking
struct TrSampler
{
    Pcg32 rng;
    __device__ static TrSampler start(uint2 seed)
    {
        const uint64_t v = (uint64_t)seed.x | ((uint64_t)seed.y << 32);
        return TrSampler{ pcg32_seed(v, mix_bits(v)) };
    }
    __device__ float get1d() { return pcg32_uniform_float(rng); }
};

// ---------------------------------------------------------------- plan

// GeneratePipeline::generate's refill rule (generate.cpp:140-176), evaluated
// on the device: refill when at least half of the pool is free or when all
// remaining paths fit.
__global__ void k_plan(WaveCounters *c, int32_t state_count, int parity)
{
    QueueState *q = queue_state(c);
    const int32_t active = c->active;
    const int64_t rest = (int64_t)state_count - active;
    int64_t fresh = 0;
    if(c->generated < c->total && rest > 0)
    {
        const int64_t unfinished = (int64_t)(c->total - c->generated);
        if(!(unfinished > rest && rest + rest < state_count))
            fresh = unfinished < rest ? unfinished : rest;
    }
    for(int k = 0; k < FETCH_WORDS; k += COUNTER_STRIDE)
    {
        q->trace_fetch[k] = 0;
        q->medium_fetch[k] = 0;
    }
    // the generated states form the last segment of the current side
    const int32_t n = active + (int32_t)fresh;
    q->pool_first[COUNTER_STRIPES] = pool_gen_slot(state_count);
    q->pool_prefix[COUNTER_STRIPES] = active;
    q->pool_prefix[COUNTER_STRIPES + 1] = n;
    // this bounce's pushes: COUNTER_STRIPES stripes over its n inputs
    const int32_t per = (n + COUNTER_STRIPES - 1) / COUNTER_STRIPES;
    q->stripe[parity] = per > 0 ? (per + 63) / 64 * 64 : 64;
    c->gen_offset = active;
    c->gen_count = (int32_t)fresh;
    c->gen_base = c->generated;
    c->generated += (uint64_t)fresh;
    c->active = n;
    c->ext_rays += (uint64_t)n;
}

// ---------------------------------------------------------------- generate

template<int SAMPLER>
__global__ void __launch_bounds__(256) k_generate(const RenderConsts *__restrict__ rcp, const WaveCounters *c,
                                                  PathSOA path)
{
    const RenderConsts &rc = *rcp;
    const int32_t n = c->gen_count;
    const int32_t gen_slot = queue_state(c)->pool_first[COUNTER_STRIPES]; // the generated segment
    const uint64_t base = c->gen_base;
    for(int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const uint64_t path_index = base + (uint64_t)i;
        int32_t v = rc.view_count - 1;
        while(v > 0 && path_index < rc.views[v].path_base)
            --v;
        const ViewRec &view = rc.views[v];
        const uint64_t local = path_index - view.path_base;
        const uint64_t pixels = (uint64_t)view.width * (uint64_t)view.height;
        // generate.cpp:46-62: uniform order walks all pixels once per sample
        // index; tile order finishes a pixel's samples before the next pixel
        uint32_t local_pixel;
        int32_t sample_index;
        if(rc.tile)
        {
            local_pixel = (uint32_t)(local / (uint64_t)rc.spp);
            sample_index = (int32_t)(local % (uint64_t)rc.spp);
        }
        else
        {
            local_pixel = (uint32_t)(local % pixels);
            sample_index = (int32_t)(local / pixels);
        }
        sample_index += rc.sample_offset;
        const uint32_t pixel_index = view.pixel_base + local_pixel; // film pixel
        uint32_t train = NO_TRAIN;
        if(rc.nrc_stride > 0 && pixel_index % (uint32_t)rc.nrc_stride == (uint32_t)rc.nrc_phase &&
           pixel_index / (uint32_t)rc.nrc_stride < (uint32_t)rc.nrc_max_paths)
            train = pixel_index / (uint32_t)rc.nrc_stride;
        const uint32_t px = local_pixel % (uint32_t)view.width;
        const uint32_t py = local_pixel / (uint32_t)view.width;

        TabledSampler<SAMPLER> smp = TabledSampler<SAMPLER>::start(rc, view, px, py, sample_index);
        float fx, fy;
        if(rc.filter.type == FILTER_GAUSSIAN)
        {
            // GaussianFilter::sample, gaussian.cpp:54-66
            const uint32_t cell = alias_sample(rc.filter.table, (uint32_t)(GAUSSIAN_GRID * GAUSSIAN_GRID), smp.get1d());
            const uint32_t gy = cell / (uint32_t)GAUSSIAN_GRID, gx = cell % (uint32_t)GAUSSIAN_GRID;
            const float r = rc.filter.radius, w = rc.filter.cell;
            const float x0 = -r + w * (float)gx, x1 = -r + w * (float)(gx + 1);
            const float y0 = -r + w * (float)gy, y1 = -r + w * (float)(gy + 1);
            fx = lerpf(x0, x1, smp.get1d());
            fy = lerpf(y0, y1, smp.get1d());
        }
        else
        {
            // BoxFilter::sample, film_filter/box.cpp:5-9
            fx = smp.get1d() - 0.5f;
            fy = smp.get1d() - 0.5f;
        }
        const float film_x = ((float)px + 0.5f + fx) / (float)view.width;
        const float film_y = ((float)py + 0.5f + fy) / (float)view.height;
        (void)smp.get1d(); // time sample, generate.cpp:70

        // PinholeCamera::generate_ray_inline, camera/pinhole.cpp:84-99
        const CameraConsts &cam = view.camera;
        const V3 x = film_x * v3(cam.film_x[0], cam.film_x[1], cam.film_x[2]);
        const V3 y = film_y * v3(cam.film_y[0], cam.film_y[1], cam.film_y[2]);
        const V3 dst = v3(cam.lbc[0], cam.lbc[1], cam.lbc[2]) + x + y;
        const V3 eye = v3(cam.eye[0], cam.eye[1], cam.eye[2]);
        const V3 dir = normalize(dst - eye);

        const int32_t s = gen_slot + i;
        path.ray_o[s] = f4(eye, __uint_as_float(rc.vol_prim_medium));
        path.ray_d[s] = f4(dir, MAX_FLOAT);
        path.beta[s] = make_float4(1.0f, 1.0f, 1.0f, __int_as_float(0));
        path.rad[s] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(train));
        path.bsdf_le[s] = make_float4(1.0f, 1.0f, 1.0f, -1.0f);
        path.smp[s] = smp.pack(pixel_index);
    }
}

// ---------------------------------------------------------------- trace

// rays of the path pool in, IntersectionSOA-style records out (trace.cpp:38-78)
struct ExtensionSink
{
    PathSOA path;
    HitSOA hit;
    const QueueState *q;

    // the state's ray from its pool slot; the hit goes to dense index i
    __device__ __forceinline__ uint32_t load(uint32_t base, uint32_t i, V3 &o, V3 &d, float &tmin, float &tmax) const
    {
        const int32_t slot = pool_slot(q, (int32_t)base, (int32_t)i);
        const float4 a = path.ray_o[slot];
        const float4 b = path.ray_d[slot];
        o = xyz(a);
        d = xyz(b);
        tmin = 0.0f;
        tmax = b.w;
        return i;
    }
    __device__ __forceinline__ void store(uint32_t i, const HitRecord &h) const
    {
        if(h.gid != 0xffffffffu)
        {
            hit.flag[i] = PATH_FLAG_HAS_INTERSECTION | h.inst;
            hit.tuv[i] = make_float4(h.t, h.b1, h.b2, __uint_as_float(h.prim));
        }
        else
            hit.flag[i] = 0;
    }
};

template<int SCHED>
__global__ void __launch_bounds__(TRACE_BLOCK) k_trace_extension(
    TraceScene ts, WaveCounters *c, PathSOA path, HitSOA hit, TraceOverflow ovf)
{
    QueueState *q = queue_state(c);
    ExtensionSink sink{ path, hit, q };
    trace_persistent<TRI_WATERTIGHT, false, (SCHED & SCHED_POSTPONE) != 0, (SCHED & SCHED_TWO_LEVEL) != 0>(
        ts, sink, c->active, q->trace_fetch, ovf);
}

// ---------------------------------------------------------------- shade

// the cache's view of a surface vertex
__device__ __forceinline__ NrcVertex nrc_vertex(const SceneView &sv, const InstanceRec &inst, const SurfacePoint &sp,
                                                V3 d)
{
    NrcVertex v;
    float rough;
    V3 dif, spe;
    nrc_material_inputs(material_at(sv, inst.material_id, sp), &rough, &dif, &spe);
    const V3 wo = -d;
    v.p[0] = sp.position.x; v.p[1] = sp.position.y; v.p[2] = sp.position.z;
    v.wo[0] = wo.x; v.wo[1] = wo.y; v.wo[2] = wo.z;
    v.n[0] = sp.interp_z.x; v.n[1] = sp.interp_z.y; v.n[2] = sp.interp_z.z;
    v.roughness = rough;
    v.diffuse[0] = dif.x; v.diffuse[1] = dif.y; v.diffuse[2] = dif.z;
    v.specular[0] = spe.x; v.specular[1] = spe.y; v.specular[2] = spe.z;
    return v;
}

// adds v to a training path's radiance sum; a path has at most one thread in
// any launch, so no atomics are needed
__device__ __forceinline__ void train_add(float *acc, uint32_t tid, V3 v)
{
    if(!isfinite3(v))
        return;
    float *a = acc + 4 * (size_t)tid;
    a[0] += v.x;
    a[1] += v.y;
    a[2] += v.z;
}

template<int SAMPLER>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) k_shade(
    const RenderConsts *__restrict__ rcp, SceneView sv, WaveCounters *c, PathSOA path, HitSOA hit,
    PathSOA next, ShadowSOA shadow, FilmView film, NrcView nrc, int parity)
{
    const RenderConsts &rc = *rcp;
    QueueState *q = queue_state(c);
    const int32_t n = c->active;
    const int32_t stripe = q->stripe[parity];
    const int32_t stride = gridDim.x * blockDim.x;
    for(int32_t base = blockIdx.x * blockDim.x; base < n; base += stride)
    {
        const int32_t i = base + (int32_t)threadIdx.x;
        const int32_t w0 = base + (int32_t)(threadIdx.x & ~63u); // this wavefront's first input
        const int32_t k = w0 / stripe;                            // its stripe
        // Queue pushes happen where their payload is made: a wave-aggregated
        // atomic among the lanes that reach that point together, so no
        // payload stays live across the rest of the shading code.
        if(i < n)
        {
            const int32_t ps = pool_slot(q, w0, i); // the state's pool slot
            const uint32_t flag = hit.flag[i];
            const float4 ro = path.ray_o[ps];
            const float4 rd = path.ray_d[ps];
            const float4 beta_depth = path.beta[ps];
            const float4 rad4 = path.rad[ps];
            const float4 le4 = path.bsdf_le[ps];
            const uint4 smp = path.smp[ps];
            const uint32_t pixel = smp.x;
            PathSampler<SAMPLER> hs = PathSampler<SAMPLER>::load(rc, smp);

            const V3 o = xyz(ro), d = xyz(rd);
            V3 beta = xyz(beta_depth);
            const int32_t depth = __float_as_int(beta_depth.w);
            V3 path_rad = xyz(rad4);
            const uint32_t train = __float_as_uint(rad4.w);
            const V3 beta_le = xyz(le4);
            const float bsdf_pdf = le4.w;

            if(!(flag & PATH_FLAG_HAS_INTERSECTION))
            {
                const V3 miss = eval_miss_le(sv, d, beta_le, bsdf_pdf);
                V3 splat = miss;
                if(!(flag & PATH_FLAG_HAS_SCATTERING))
                    splat = splat + path_rad;
                film_splat(film, pixel, splat);
                if(train != NO_TRAIN)
                    train_add(nrc.acc, train, miss);
            }
            else
            {
                const float4 tuv = hit.tuv[i];
                const uint32_t inst_id = flag & PATH_FLAG_INSTANCE_ID_MASK;
                const InstanceRec &inst = sv.instances[inst_id];
                const SurfacePoint sp = get_hitinfo(sv, o, d, inst, __float_as_uint(tuv.w), tuv.x, tuv.y, tuv.z);

                const V3 le_rad = handle_intersected_light(sv, o, d, sp, beta_le, bsdf_pdf, inst.light_id);
                bool done = false;
                if(flag & PATH_FLAG_HAS_SCATTERING)
                {
                    film_splat(film, pixel, le_rad);
                    done = true;
                }
                if(!done && train != NO_TRAIN)
                {
                    // training path: the radiance sum so far, then the vertex
                    // record (throughput before Russian roulette, so that the
                    // target stays unbiased)
                    train_add(nrc.acc, train, le_rad);
                    if(depth == nrc.train_depth && nrc.queries)
                    {
                        // self-training: the cache answers for the tail of
                        // the training path; its answer joins the path's sum
                        NrcQuery query;
                        query.v = nrc_vertex(sv, inst, sp, d);
                        query.beta[0] = beta.x; query.beta[1] = beta.y; query.beta[2] = beta.z;
                        query.pixel = pixel;
                        query.train = train;
                        nrc.queries[wave_push(nrc.query_count, true)] = query;
                        film_splat(film, pixel, path_rad + le_rad);
                        done = true;
                    }
                    else if(depth < nrc.slots_per_path)
                    {
                        NrcRecord &r = nrc.rec[(size_t)train * nrc.slots_per_path + depth];
                        r.v = nrc_vertex(sv, inst, sp, d);
                        const float *a = nrc.acc + 4 * (size_t)train;
                        r.beta[0] = beta.x; r.beta[1] = beta.y; r.beta[2] = beta.z;
                        r.gathered[0] = a[0]; r.gathered[1] = a[1]; r.gathered[2] = a[2];
                        r.path = train;
                        r.valid = 1;
                    }
                }
                else if(!done && depth == nrc.query_depth)
                {
                    // the prefix ends here: the cache answers for the rest
                    if(nrc.queries)
                    {
                        NrcQuery query;
                        query.v = nrc_vertex(sv, inst, sp, d);
                        query.beta[0] = beta.x; query.beta[1] = beta.y; query.beta[2] = beta.z;
                        query.pixel = pixel;
                        query.train = NO_TRAIN;
                        nrc.queries[wave_push(nrc.query_count, true)] = query;
                    }
                    film_splat(film, pixel, path_rad + le_rad);
                    done = true;
                }
                if(!done)
                {
                    path_rad = path_rad + le_rad;
                    // simple_russian_roulette, helper.cpp:67-97
                    bool rr_exit = false;
                    if(depth >= rc.min_depth)
                    {
                        if(depth >= rc.max_depth)
                            rr_exit = true;
                        else
                        {
                            const float u = hs.get1d();
                            if(luminance(beta) < rc.rr_threshold)
                            {
                                if(u > rc.rr_cont_prob)
                                    rr_exit = true;
                                else
                                    beta = beta / rc.rr_cont_prob;
                            }
                        }
                    }
                    if(rr_exit)
                    {
                        film_splat(film, pixel, path_rad);
                        done = true;
                    }
                }
                if(!done)
                {
                    float l0, l1, l2;
                    hs.get3d(&l0, &l1, &l2);
                    const float lsel = hs.get1d();
                    ShadowSample ss = sample_surface_light_li(sv, sp.position, sp.frame.z, l0, l1, l2, lsel);

                    const Shader shader = make_shader(material_at(sv, inst.material_id, sp), sp.frame, sp.interp_z);
                    if(depth == 0)
                    {
                        // g-buffer of the first hit: Shader::albedo and the
                        // shading normal (shade.cpp:133-139,196-204)
                        if(film.albedo)
                            aov_splat(film.albedo, pixel, shader.albedo);
                        if(film.normal)
                            aov_splat(film.normal, pixel, shader_normal(shader));
                    }
                    // draws in the reference's order; the light connection
                    // is finished (its shadow ray pushed) before the BSDF is
                    // sampled, then the continuation
                    float s0, s1, s2;
                    hs.get3d(&s0, &s1, &s2);
                    V3 shadow_f = v3s(0.0f);
                    float shadow_pdf = 0.0f;
                    if(ss.success)
                    {
                        shadow_f = shader_eval(shader, ss.d, -d);
                        ss.success = !is_zero(shadow_f);
                        if(ss.success)
                            shadow_pdf = shader_pdf(shader, ss.d, -d);
                    }
                    if(ss.success)
                    {
                        // vacuum: the shadow ray keeps the medium of its side
                        const bool shadow_out = dot(sp.frame.z, ss.d) > 0.0f;
                        const bool last_out = dot(sp.frame.z, d) < 0.0f;
                        const uint32_t med = shadow_out == last_out ? __float_as_uint(ro.w)
                                           : (shadow_out ? inst.outer_medium : inst.inner_medium);
                        const float cs = fabs_(dot(normalize(ss.d), sp.frame.z));
                        const V3 beta_li = ss.li * beta * shadow_f * cs / (shadow_pdf + ss.light_pdf);
                        const int32_t si = stripe_push(q->shadow_n[parity], k, stripe, true);
                        shadow.o[si] = f4(ss.o, __uint_as_float(pixel));
                        shadow.d[si] = f4(ss.d, ss.t1);
                        shadow.beta[si] = f4(beta_li, __uint_as_float(med));
                        if(shadow.seed)
                            shadow.seed[si] = shadow_seed(hs.pack(pixel));
                        if(shadow.train)
                            shadow.train[si] = train;
                    }

                    BSDFSample bs;
                    shader_sample(shader, -d, s0, s1, s2, bs);
                    if(!is_zero(bs.bsdf))
                    {
                        const V3 dir = normalize(bs.dir);
                        const float cs = fabs_(dot(dir, sp.frame.z));
                        const V3 new_beta_le = beta * cs * bs.bsdf;
                        beta = new_beta_le / bs.pdf;
                        const V3 next_o = intersection_offset(sp.position, sp.frame.z, dir);
                        const bool next_out = dot(sp.frame.z, dir) > 0.0f;
                        const uint32_t med = next_out ? inst.outer_medium : inst.inner_medium;
                        const float stored_pdf = bs.is_delta ? -bs.pdf : bs.pdf;
                        const int32_t ni = stripe_push(q->next_n, k, stripe, true);
                        next.ray_o[ni] = f4(next_o, __uint_as_float(med));
                        next.ray_d[ni] = f4(dir, MAX_FLOAT);
                        next.beta[ni] = f4(beta, __int_as_float(depth + 1));
                        next.rad[ni] = f4(path_rad, __uint_as_float(train));
                        next.bsdf_le[ni] = f4(new_beta_le, stored_pdf);
                        next.smp[ni] = hs.pack(pixel);
                    }
                    else
                        film_splat(film, pixel, path_rad);
                }
            }
        }
    }
}

// ---------------------------------------------------------------- medium

// sample_medium_scattering_kernel, wavefront/medium.cpp:28-162, in two
// kernels. For every path whose last segment runs through a medium,
// k_medium_track tracks a scattering point along the segment (delta
// tracking) and its transmittance (ratio tracking), weights beta_le by the
// transmittance and records the scattering point; k_medium_scatter then
// finishes the vertex at every scattering point (Russian roulette, light
// sample through the phase function, phase-sampled continuation). The
// shading stage only adds the emission the segment ended on.
//
// The tracking loops run for a number of steps that differs from path to
// path, and many states have no grid-medium segment at all, so
// k_medium_track is persistent like the traversal. A wavefront takes states
// in batches of 64 (a first batch by its workgroup index, then GRANULE-state
// granules dealt round robin to the chunk counters medium_fetch[k *
// COUNTER_STRIDE], as the traversal hands out rays); all 64 lanes find the
// segments of a batch together and queue the states whose segment runs
// through a grid medium in the wavefront's LDS. Every pass advances each
// busy lane's tracking by one step; once MEDIUM_REFILL_MIN lanes are idle
// they take states from the queue, topped up batch by batch.
constexpr int MEDIUM_BLOCK = 64;       // one wavefront per workgroup
constexpr int MEDIUM_REFILL_MIN = 16;  // refill once this many lanes are idle
constexpr int MEDIUM_QUEUE = 128;      // LDS queue entries (a ring; it never holds more than 127)
static_assert(GRANULE == MEDIUM_BLOCK, "a granule is one batch of the tracking kernel");

// get_medium_segment, medium.cpp:214-245: the medium of state i's last
// segment (ray origin ro, direction d) and the segment's end
__device__ __forceinline__ uint32_t medium_segment(const SceneView &sv, const HitSOA &hit, int32_t i, uint32_t flag,
                                                  float4 ro, V3 d, V3 *end)
{
    const V3 o = xyz(ro);
    if(flag & PATH_FLAG_HAS_INTERSECTION)
    {
        const float4 tuv = hit.tuv[i];
        const InstanceRec &inst = sv.instances[flag & PATH_FLAG_INSTANCE_ID_MASK];
        const V3 nor = hit_geometric_normal(sv, inst, __float_as_uint(tuv.w));
        const uint32_t inct_med = dot(d, nor) < 0.0f ? inst.outer_medium : inst.inner_medium;
        const uint32_t ray_med = __float_as_uint(ro.w);
        *end = intersection_offset(o + tuv.x * d, nor, -d);
        return inct_med < ray_med ? inct_med : ray_med;
    }
    // the volume-primitive medium up to the world's diagonal
    *end = o + normalize(d) * sv.world_diagonal;
    return (uint32_t)sv.medium_count;
}

// A state's tracking result for k_medium_scatter, two float4 per dense index:
// scattering point and phase asymmetry, phase colour
__device__ __forceinline__ void record_scatter(float4 *scatter, int32_t i, const MediumSample &ms)
{
    scatter[2 * (size_t)i] = f4(ms.position, ms.g);
    scatter[2 * (size_t)i + 1] = f4(ms.color, 0.0f);
}

template<int SAMPLER>
__global__ void __launch_bounds__(MEDIUM_BLOCK) k_medium_track(const RenderConsts *__restrict__ rcp, SceneView sv,
                                                               WaveCounters *c, PathSOA path, HitSOA hit,
                                                               float4 *scatter)
{
    // the queue: dense index, pool slot, hit flag, medium, segment start and end
    __shared__ int32_t q_i[MEDIUM_QUEUE], q_ps[MEDIUM_QUEUE];
    __shared__ uint32_t q_flag[MEDIUM_QUEUE], q_med[MEDIUM_QUEUE];
    __shared__ float q_o[3][MEDIUM_QUEUE], q_end[3][MEDIUM_QUEUE];

    const RenderConsts &rc = *rcp;
    QueueState *q = queue_state(c);
    const uint32_t n = (uint32_t)c->active;
    const int lane = (int)threadIdx.x;
    const uint64_t lanes_below = (1ull << lane) - 1ull;
    const uint32_t static_span = gridDim.x * (uint32_t)MEDIUM_BLOCK;
    const uint32_t chunks = min((uint32_t)COUNTER_STRIPES, gridDim.x);
    const uint32_t home = blockIdx.x % chunks;
    uint32_t *counter = q->medium_fetch + home * COUNTER_STRIDE;
    // wavefront-uniform: the first batch, whether the home chunk has run
    // dry, and the LDS queue's head and tail (entries head .. tail - 1)
    bool first = blockIdx.x * (uint32_t)MEDIUM_BLOCK < n;
    bool dry = static_span >= n;
    uint32_t head = 0, tail = 0;

    // the lane's state: dense index, pool slot, hit flag, film pixel, sampler,
    // the grid medium being tracked, its progress and results
    bool active = false, sampling = false, scattered = false;
    int32_t i = 0, ps = 0;
    uint32_t flag = 0, pixel = 0, med = 0;
    PathSampler<SAMPLER> hs{};
    GridTrack g{};
    float tr = 1.0f;

    for(;;)
    {
        // ---- idle lanes take queued states; the queue is topped up first
        const uint64_t idle = __ballot(!active);
        const int nidle = (int)__popcll(idle);
        if(nidle >= MEDIUM_REFILL_MIN)
        {
            while(tail - head < (uint32_t)nidle && (first || !dry))
            {
                // the next batch of up to 64 states
                uint32_t base;
                if(first)
                {
                    base = blockIdx.x * (uint32_t)MEDIUM_BLOCK;
                    first = false;
                }
                else
                {
                    uint32_t gr = 0;
                    if(lane == 0)
                        gr = atomicAdd(counter, 1u);
                    gr = __builtin_amdgcn_readfirstlane(gr);
                    const uint64_t lo = (uint64_t)static_span + ((uint64_t)gr * chunks + home) * GRANULE;
                    if(lo >= (uint64_t)n)
                    {
                        dry = true;
                        break;
                    }
                    base = (uint32_t)lo;
                }
                const uint32_t cnt = min((uint32_t)MEDIUM_BLOCK, n - base);
                // every lane finds one state's segment (get_medium_segment)
                bool grid = false;
                int32_t bj = 0, bps = 0;
                uint32_t bflag = 0, bmed = 0;
                V3 bo = v3s(0.0f), bend = v3s(0.0f);
                if((uint32_t)lane < cnt)
                {
                    bj = (int32_t)(base + (uint32_t)lane);
                    bps = pool_slot(q, (int32_t)base, bj);
                    bflag = hit.flag[bj];
                    const float4 ro = path.ray_o[bps];
                    const V3 d = xyz(path.ray_d[bps]);
                    bo = xyz(ro);
                    bmed = medium_segment(sv, hit, bj, bflag, ro, d, &bend);
                    path.ray_o[bps] = f4(bo, __uint_as_float(bmed));
                    // grid media are tracked here; segments in a homogeneous
                    // medium or the volume primitives (no long tracking
                    // loops) are left to k_medium_scatter
                    grid = bmed < (uint32_t)sv.medium_count && sv.media[bmed].type != MEDIUM_HOMOGENEOUS;
                }
                const uint64_t gm = __ballot(grid);
                if(grid)
                {
                    const uint32_t e = (tail + (uint32_t)__popcll(gm & lanes_below)) % MEDIUM_QUEUE;
                    q_i[e] = bj;
                    q_ps[e] = bps;
                    q_flag[e] = bflag;
                    q_med[e] = bmed;
                    q_o[0][e] = bo.x;
                    q_o[1][e] = bo.y;
                    q_o[2][e] = bo.z;
                    q_end[0][e] = bend.x;
                    q_end[1][e] = bend.y;
                    q_end[2][e] = bend.z;
                }
                tail += (uint32_t)__popcll(gm);
            }
            __syncthreads(); // the queue's entries are visible to every lane
            const uint32_t take = min((uint32_t)nidle, tail - head);
            if(!active)
            {
                const uint32_t rank = (uint32_t)__popcll(idle & lanes_below);
                if(rank < take)
                {
                    const uint32_t e = (head + rank) % MEDIUM_QUEUE;
                    i = q_i[e];
                    ps = q_ps[e];
                    flag = q_flag[e];
                    med = q_med[e];
                    const V3 o = v3(q_o[0][e], q_o[1][e], q_o[2][e]);
                    const V3 end = v3(q_end[0][e], q_end[1][e], q_end[2][e]);
                    const uint4 smp = path.smp[ps];
                    pixel = smp.x;
                    hs = PathSampler<SAMPLER>::load(rc, smp);
                    g = grid_track_start(sv.media[med], o, end);
                    tr = 1.0f;
                    sampling = true;
                    scattered = false;
                    active = true;
                }
            }
            head += take;
            __syncthreads(); // entries are read before later batches reuse them
        }
        if(!__any(active))
        {
            if(head == tail && !first && dry)
                break;
            continue; // the queue ran empty before the states did
        }

        // ---- one tracking step (medium_sample's loop, then medium_tr's)
        if(active)
        {
            const MediumRec &m = sv.media[med];
            V3 uvw;
            float density;
            const bool inside = grid_track_step(m, sv.voxels, g, hs, &uvw, &density);
            if(sampling)
            {
                bool done = !inside;
                if(inside && hs.get1d() < density * m.inv_max_density)
                {
                    MediumSample ms;
                    ms.scattered = true;
                    ms.position = g.a + g.t * g.world_step;
                    ms.color = grid_albedo(m, sv.voxels, uvw);
                    ms.g = m.g;
                    record_scatter(scatter, i, ms);
                    scattered = true;
                    done = true;
                }
                if(done)
                {
                    // medium_tr tracks the whole segment again from its start
                    sampling = false;
                    g.t = 0.0f;
                    g.it = 0;
                }
            }
            else if(inside)
                tr = tr * (1.0f - density * m.inv_max_density);
            else
            {
                const float4 le4 = path.bsdf_le[ps];
                path.bsdf_le[ps] = f4(xyz(le4) * v3s(tr), le4.w);
                path.smp[ps] = hs.pack(pixel);
                if(scattered)
                    hit.flag[i] = flag | PATH_FLAG_HAS_SCATTERING;
                active = false;
            }
        }
    }
}

// The scattering vertices recorded by k_medium_track (the rest of
// sample_medium_scattering_kernel): Russian roulette, the light sample
// through the phase function, the phase-sampled continuation.
template<int SAMPLER>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) k_medium_scatter(
    const RenderConsts *__restrict__ rcp, SceneView sv, WaveCounters *c, PathSOA path, HitSOA hit,
    const float4 *scatter, PathSOA next, ShadowSOA shadow, FilmView film, int parity)
{
    const RenderConsts &rc = *rcp;
    QueueState *q = queue_state(c);
    const int32_t n = c->active;
    const int32_t stripe = q->stripe[parity];
    const int32_t stride = gridDim.x * blockDim.x;
    for(int32_t base = blockIdx.x * blockDim.x; base < n; base += stride)
    {
        const int32_t i = base + (int32_t)threadIdx.x;
        const int32_t w0 = base + (int32_t)(threadIdx.x & ~63u); // this wavefront's first input
        const int32_t k = w0 / stripe;                            // its stripe
        // queue pushes where their payload is made, as in k_shade
        if(i >= n)
            continue;
        const uint32_t flag = hit.flag[i];
        const int32_t ps = pool_slot(q, w0, i); // the state's pool slot
        const float4 ro = path.ray_o[ps];
        const uint32_t med = __float_as_uint(ro.w); // the segment's medium (k_medium_track)
        const bool vol_prim = med == (uint32_t)sv.medium_count && sv.vol.count > 0;
        const bool homogeneous = med < (uint32_t)sv.medium_count && sv.media[med].type == MEDIUM_HOMOGENEOUS;
        if((flag & PATH_FLAG_HAS_SCATTERING) || vol_prim || homogeneous)
        {
            const V3 d = xyz(path.ray_d[ps]);
            const uint4 smp = path.smp[ps];
            const uint32_t pixel = smp.x;
            PathSampler<SAMPLER> hs = PathSampler<SAMPLER>::load(rc, smp);
            V3 position, color;
            float gph;
            if(vol_prim || homogeneous)
            {
                // no long tracking loop: sample and transmittance here
                V3 end;
                (void)medium_segment(sv, hit, i, flag, ro, d, &end);
                const V3 o = xyz(ro);
                MediumSample ms;
                V3 trv;
                if(vol_prim)
                {
                    ms = volume_sample(sv.vol, o, end, hs);
                    trv = volume_tr(sv.vol, o, end, hs);
                }
                else
                {
                    const MediumRec &m = sv.media[med];
                    ms = medium_sample(m, sv.voxels, o, end, hs);
                    trv = medium_tr(m, sv.voxels, o, end, hs);
                }
                const float4 le4 = path.bsdf_le[ps];
                path.bsdf_le[ps] = f4(xyz(le4) * trv, le4.w);
                if(!ms.scattered)
                {
                    path.smp[ps] = hs.pack(pixel);
                    continue;
                }
                hit.flag[i] = flag | PATH_FLAG_HAS_SCATTERING;
                position = ms.position;
                color = ms.color;
                gph = ms.g;
            }
            else
            {
                const float4 sa = scatter[2 * (size_t)i], sb = scatter[2 * (size_t)i + 1];
                position = xyz(sa);
                color = xyz(sb);
                gph = sa.w;
            }

            const float4 beta_depth = path.beta[ps];
            V3 beta = xyz(beta_depth); // medium throughput is one
            const int32_t depth = __float_as_int(beta_depth.w);
            const float4 rad4 = path.rad[ps];
            const V3 path_rad = xyz(rad4);

            bool rr_exit = false;
            if(depth >= rc.min_depth)
            {
                if(depth >= rc.max_depth)
                    rr_exit = true;
                else
                {
                    const float u = hs.get1d();
                    if(luminance(beta) < rc.rr_threshold)
                    {
                        if(u > rc.rr_cont_prob)
                            rr_exit = true;
                        else
                            beta = beta / rc.rr_cont_prob;
                    }
                }
            }
            if(rr_exit)
                film_splat(film, pixel, path_rad);
            else
            {
                float l0, l1, l2;
                hs.get3d(&l0, &l1, &l2);
                const float lsel = hs.get1d();
                ShadowSample ss = sample_light_li(sv, position, v3s(0.0f), false, l0, l1, l2, lsel);
                V3 beta_li = v3s(0.0f);
                if(ss.success)
                {
                    const float pv = hg_pdf(gph, ss.d, -d);
                    beta_li = ss.li * beta * (color * pv) / (pv + ss.light_pdf);
                }
                float s0, s1, s2;
                hs.get3d(&s0, &s1, &s2);
                const PhaseSample phs = hg_sample(gph, color, -d, s0, s1);
                if(ss.success)
                {
                    const uint32_t smed = ss.t1 > 1.0f ? (uint32_t)sv.medium_count : med;
                    const int32_t si = stripe_push(q->shadow_n[parity], k, stripe, true);
                    shadow.o[si] = f4(ss.o, __uint_as_float(pixel));
                    shadow.d[si] = f4(ss.d, ss.t1);
                    shadow.beta[si] = f4(beta_li, __uint_as_float(smed));
                    shadow.seed[si] = shadow_seed(hs.pack(pixel));
                    if(shadow.train)
                        shadow.train[si] = __float_as_uint(rad4.w);
                }
                if(!is_zero(phs.phase))
                {
                    const V3 dir = normalize(phs.dir);
                    const V3 beta_le = beta * phs.phase;
                    beta = beta_le / phs.pdf;
                    const int32_t ni = stripe_push(q->next_n, k, stripe, true);
                    next.ray_o[ni] = f4(position, __uint_as_float(med));
                    next.ray_d[ni] = f4(dir, MAX_FLOAT);
                    next.beta[ni] = f4(beta, __int_as_float(depth + 1));
                    next.rad[ni] = rad4;
                    next.bsdf_le[ni] = f4(beta_le, phs.pdf);
                    next.smp[ni] = hs.pack(pixel);
                }
                else
                    film_splat(film, pixel, path_rad);
            }
        }
    }
}

// ---------------------------------------------------------------- shadow

// shadow rays in; on a miss the light sample reaches the film, weighted by
// the transmittance of the ray's medium (__miss__shadow, shadow.cpp:43-71)
template<bool MEDIA>
struct ShadowSink
{
    ShadowSOA shadow;
    FilmView film;
    SceneView sv;
    const QueueState *q;
    int parity; // of the bounce that pushed the rays

    // the ray from its shadow queue slot, which is also its key
    __device__ __forceinline__ uint32_t load(uint32_t base, uint32_t j, V3 &o, V3 &d, float &tmin, float &tmax) const
    {
        const int32_t i = shadow_slot(q, parity, (int32_t)base, (int32_t)j);
        const float4 a = shadow.o[i];
        const float4 b = shadow.d[i];
        o = xyz(a);
        d = xyz(b);
        tmin = 0.0f;
        tmax = b.w;
        return (uint32_t)i;
    }
    __device__ __forceinline__ void store(uint32_t i, const HitRecord &h) const
    {
        if(h.gid == 0xffffffffu)
        {
            const float4 a = shadow.o[i];
            const float4 b = shadow.beta[i];
            V3 beta = xyz(b);
            if(MEDIA)
            {
                // rays to the environment (t1 > 1) cross the volume-primitive
                // medium over the world's diagonal (shadow.cpp:56-66)
                const float4 dd = shadow.d[i];
                const uint32_t med = __float_as_uint(b.w);
                const V3 o = xyz(a);
                if(dd.w > 1.0f)
                {
                    if(sv.vol.count > 0)
                    {
                        TrSampler ts = TrSampler::start(shadow.seed[i]);
                        beta = beta * volume_tr(sv.vol, o, o + normalize(xyz(dd)) * sv.world_diagonal, ts);
                    }
                }
                else if(med < (uint32_t)sv.medium_count)
                {
                    TrSampler ts = TrSampler::start(shadow.seed[i]);
                    beta = beta * medium_tr(sv.media[med], sv.voxels, o, o + xyz(dd) * dd.w, ts);
                }
                else if(med == (uint32_t)sv.medium_count && sv.vol.count > 0)
                {
                    TrSampler ts = TrSampler::start(shadow.seed[i]);
                    beta = beta * volume_tr(sv.vol, o, o + xyz(dd) * dd.w, ts);
                }
            }
            film_splat(film, __float_as_uint(a.w), beta);
            if(shadow.train)
            {
                const uint32_t tid = shadow.train[i];
                if(tid != NO_TRAIN)
                    train_add(shadow.train_acc, tid, beta);
            }
        }
    }
};

template<int SCHED, bool MEDIA>
__global__ void __launch_bounds__(TRACE_BLOCK) k_shadow(
    TraceScene ts, SceneView sv, WaveCounters *c, ShadowSOA shadow, FilmView film, TraceOverflow ovf, int parity)
{
    QueueState *q = queue_state(c);
    ShadowSink<MEDIA> sink{ shadow, film, sv, q, parity };
    trace_persistent<TRI_WATERTIGHT, true, (SCHED & SCHED_POSTPONE) != 0, (SCHED & SCHED_TWO_LEVEL) != 0>(
        ts, sink, shadow_count(q, parity), q->shadow_fetch[parity], ovf);
}

// ---------------------------------------------------------------- bookkeeping

// the states pushed this bounce become the next bounce's pool side: one
// segment per stripe, then the (empty) generated segment; the next-state
// counters restart, and so do the shadow counters of the other parity (the
// previous bounce's shadow rays were traced before this bounce shaded)
__global__ void k_end_iteration(WaveCounters *c, int parity)
{
    QueueState *q = queue_state(c);
    const int32_t stripe = q->stripe[parity];
    int32_t states = 0, shadows = 0;
    for(int k = 0; k < COUNTER_STRIPES; ++k)
    {
        q->pool_first[k] = k * stripe;
        q->pool_prefix[k] = states;
        states += q->next_n[k * COUNTER_STRIDE];
        shadows += q->shadow_n[parity][k * COUNTER_STRIDE];
        q->next_n[k * COUNTER_STRIDE] = 0;
        q->shadow_n[parity ^ 1][k * COUNTER_STRIDE] = 0;
        q->shadow_fetch[parity ^ 1][k * COUNTER_STRIDE] = 0;
    }
    q->pool_prefix[COUNTER_STRIPES] = states;
    q->pool_prefix[COUNTER_STRIPES + 1] = states;
    c->shadow_rays += (uint64_t)shadows;
    c->active = states;
    c->iterations += 1;
}

// preview resolve, wavefront/preview.cpp:28-80: value / weight (weight =
// samples per pixel); normals are normalised, or zero when shorter than 1e-3
__global__ void k_resolve(int32_t mode, const float *value, int32_t pixels, float inv_weight, float4 *out)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= pixels)
        return;
    const float *p = value + 4 * (size_t)i;
    V3 v = v3(p[0] * inv_weight, p[1] * inv_weight, p[2] * inv_weight);
    if(mode == RESOLVE_NORMAL)
        v = length(v) > 1e-3f ? normalize(v) : v3s(0.0f);
    out[i] = f4(v, 1.0f);
}

// ---------------------------------------------------------------- ray batches

template<bool ANY>
struct BatchSink
{
    const float4 *rays;
    uint32_t *out;

    __device__ __forceinline__ uint32_t load(uint32_t, uint32_t i, V3 &o, V3 &d, float &tmin, float &tmax) const
    {
        const float4 a = rays[2 * (size_t)i + 0];
        const float4 b = rays[2 * (size_t)i + 1];
        o = xyz(a);
        d = xyz(b);
        tmin = a.w;
        tmax = b.w;
        return i;
    }
    __device__ __forceinline__ void store(uint32_t i, const HitRecord &h) const
    {
        if(ANY)
            out[i] = h.gid != 0xffffffffu ? 1u : 0u;
        else
        {
            uint32_t *o = out + 5 * (size_t)i;
            const bool hitp = h.gid != 0xffffffffu;
            o[0] = __float_as_uint(h.t);
            o[1] = hitp ? h.inst : 0xffffffffu;
            o[2] = hitp ? h.prim : 0xffffffffu;
            o[3] = __float_as_uint(h.b1);
            o[4] = __float_as_uint(h.b2);
        }
    }
};

template<int TRI, bool ANY, int SCHED>
__global__ void __launch_bounds__(TRACE_BLOCK) k_trace_batch(
    TraceScene ts, const float4 *rays, int32_t n, uint32_t *out, uint32_t *fetch, TraceOverflow ovf)
{
    BatchSink<ANY> sink{ rays, out };
    trace_persistent<TRI, ANY, (SCHED & SCHED_POSTPONE) != 0, (SCHED & SCHED_TWO_LEVEL) != 0>(ts, sink, n, fetch, ovf);
}

// ---------------------------------------------------------------- test hooks

// one medium along segments a -> a + d * tmax with one PCG32 stream per
// segment: delta tracking, then ratio tracking on the same stream
__global__ void k_volume_eval(VolumeView vv, const float4 *seg, const uint2 *seed, int32_t n, float *out)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n)
        return;
    const float4 a4 = seg[2 * i], d4 = seg[2 * i + 1];
    const V3 a = xyz(a4), b = xyz(a4) + xyz(d4) * d4.w;
    TrSampler smp = TrSampler::start(seed[i]);
    const MediumSample ms = volume_sample(vv, a, b, smp);
    const V3 tr = volume_tr(vv, a, b, smp);
    float *o = out + 10 * (size_t)i;
    o[0] = ms.scattered ? 1.0f : 0.0f;
    o[1] = ms.position.x; o[2] = ms.position.y; o[3] = ms.position.z;
    o[4] = ms.color.x; o[5] = ms.color.y; o[6] = ms.color.z;
    o[7] = tr.x; o[8] = tr.y; o[9] = tr.z;
}

// world-space records of a one-level structure after a geometry's vertices
// changed: the records of every instance of `geometry` take the new
// positions through the instance's transform, as the upload computes them
__global__ void k_update_world_tris(float4 *tris, uint32_t n, const InstanceRec *insts, int32_t geometry,
                                    const float *positions)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k >= n)
        return;
    const uint32_t inst = __float_as_uint(tris[3 * (size_t)k + 1].w);
    const InstanceRec &r = insts[inst];
    if(r.geometry_id != geometry)
        return;
    const float *p = positions + 9 * (size_t)__float_as_uint(tris[3 * (size_t)k + 2].w);
    for(int v = 0; v < 3; ++v)
    {
        const V3 w = xf_point(r.xf, v3(p[3 * v], p[3 * v + 1], p[3 * v + 2]));
        tris[3 * (size_t)k + v] = make_float4(w.x, w.y, w.z, tris[3 * (size_t)k + v].w);
    }
}

// the importance table's per-tile luminance and solid angle, one thread per tile
__global__ void k_envir_table(TextureView tv, int32_t tex, int32_t w, int32_t h, int32_t n_samples, float *lum,
                              float *area)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < w * h)
        envir_tile_stats(tv, tex, w, h, i % w, i / w, n_samples, &lum[i], &area[i]);
}

__global__ void k_envir_eval(SceneView sv, const float *dirs, const float *sam, int32_t n, float *out)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n)
        return;
    const LightRec &l = sv.lights[sv.envir_light_index];
    const V3 d = v3(dirs[3 * i + 0], dirs[3 * i + 1], dirs[3 * i + 2]);
    const V3 le = envir_eval_le(sv, l, d);
    const float pdf = envir_pdf_li(sv, l, d);
    V3 sd, li;
    float spdf;
    envir_sample_li(sv, l, sam[3 * i + 0], sam[3 * i + 1], sam[3 * i + 2], &sd, &li, &spdf);
    const float r[12] = { le.x, le.y, le.z, pdf, sd.x, sd.y, sd.z, spdf, li.x, li.y, li.z, 0.0f };
    for(int k = 0; k < 12; ++k)
        out[12 * i + k] = r[k];
}

__global__ void k_texture_eval(TextureView tv, int32_t id, const float2 *uv, int32_t n, float4 *out)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
        out[i] = tex_sample(tv, id, uv[i].x, uv[i].y);
}

__global__ void k_medium_eval(SceneView sv, int32_t med, const float4 *seg, const uint2 *seed, int32_t n, float *out)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n)
        return;
    const float4 a4 = seg[2 * i], d4 = seg[2 * i + 1];
    const V3 a = xyz(a4), b = xyz(a4) + xyz(d4) * d4.w;
    TrSampler smp = TrSampler::start(seed[i]);
    const MediumRec &m = sv.media[med];
    const MediumSample ms = medium_sample(m, sv.voxels, a, b, smp);
    const V3 tr = medium_tr(m, sv.voxels, a, b, smp);
    float *o = out + 10 * (size_t)i;
    o[0] = ms.scattered ? 1.0f : 0.0f;
    o[1] = ms.position.x; o[2] = ms.position.y; o[3] = ms.position.z;
    o[4] = ms.color.x; o[5] = ms.color.y; o[6] = ms.color.z;
    o[7] = tr.x; o[8] = tr.y; o[9] = tr.z;
}

__global__ void k_independent_eval(const uint32_t *px, const uint32_t *py, const int32_t *si, int32_t n,
                                   int32_t dims, float *out)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n)
        return;
    Pcg32 r = independent_start(px[i], py[i], si[i]);
    for(int32_t k = 0; k < dims; ++k)
        out[(size_t)i * dims + k] = pcg32_uniform_float(r);
}

// STREAM-style copy that measures the device's HBM bandwidth: each thread
// moves four 16-byte vectors per pass, all four loads issued before the
// stores, with non-temporal hints (the data is not reused)
typedef float stream_f4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_stream_copy(const stream_f4 *__restrict__ a, stream_f4 *__restrict__ b,
                                                      int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * 1024;
    for(int64_t base = (int64_t)blockIdx.x * 1024 + threadIdx.x; base < n; base += stride)
    {
        stream_f4 v[4];
#pragma unroll
        for(int k = 0; k < 4; ++k)
            if(base + k * 256 < n)
                v[k] = __builtin_nontemporal_load(a + base + k * 256);
#pragma unroll
        for(int k = 0; k < 4; ++k)
            if(base + k * 256 < n)
                __builtin_nontemporal_store(v[k], b + base + k * 256);
    }
}

__global__ void k_halton_eval(HaltonConsts hc, HaltonTables tab, const uint32_t *px, const uint32_t *py,
                              const int32_t *si, int32_t n, int32_t dims, float *out)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n)
        return;
    // even samples read the first-digit tables, odd ones run the digit loop
    // alone (the camera samples and the later stages' forms)
    const HaltonTables t = (i & 1) ? HaltonTables{ nullptr, nullptr } : tab;
    HaltonState hs = halton_start(hc, px[i], py[i], si[i]);
    for(int32_t k = 0; k < dims; ++k)
        out[(size_t)i * dims + k] = halton_get1d(c_primes.p, hs, t);
}

// ACESToneMap's kernel, postprocess/tonemap.cpp:14-24,35-51: Narkowicz's
// fit of the exposure-scaled colour, saturated (NaN to 0); alpha unchanged
__device__ __forceinline__ float aces_fit(float x)
{
    const float y = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
    return !(y > 0.0f) ? 0.0f : (y > 1.0f ? 1.0f : y);
}

__global__ void k_tonemap_aces(float4 *rgba, int32_t pixels, float exposure)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= pixels)
        return;
    float4 c = rgba[i];
    c.x = aces_fit(c.x * exposure);
    c.y = aces_fit(c.y * exposure);
    c.z = aces_fit(c.z * exposure);
    rgba[i] = c;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t value)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
        p[i] = value;
}

__global__ void k_detmath_eval(int32_t fn, const float *x, const float *y, int32_t n, float *out)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n)
        return;
    const float a = x[i];
    float r;
    switch(fn)
    {
    case 0: r = btrc_detmath::sin(a); break;
    case 1: r = btrc_detmath::cos(a); break;
    case 2: r = btrc_detmath::log(a); break;
    case 3: r = btrc_detmath::exp(a); break;
    case 4: r = btrc_detmath::pow(a, y[i]); break;
    case 5: r = sqrtf(a); break;
    case 7: r = btrc_detmath::atan2(a, y[i]); break;
    case 8: r = btrc_detmath::acos(a); break;
    case 9: r = intersection_offset(v3(a, 0.0f, 0.0f), v3(y[i], 0.0f, 0.0f)).x; break;
    default: r = a / y[i]; break;
    }
    out[i] = r;
}

} // namespace

// ---------------------------------------------------------------- launchers

void launch_plan(hipStream_t s, WaveCounters *c, int32_t state_count, int parity)
{
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(1), 0, s, c, state_count, parity);
}

void launch_generate(hipStream_t s, int grid, int sampler, const RenderConsts *rc, const WaveCounters *c,
                     PathSOA path)
{
    if(sampler == SAMPLER_INDEPENDENT)
        hipLaunchKernelGGL(k_generate<SAMPLER_INDEPENDENT>, dim3(grid), dim3(256), 0, s, rc, c, path);
    else
        hipLaunchKernelGGL(k_generate<SAMPLER_HALTON>, dim3(grid), dim3(256), 0, s, rc, c, path);
}

void launch_trace_extension(hipStream_t s, int grid, int sched, const TraceScene &ts, WaveCounters *c,
                            PathSOA path, HitSOA hit, const TraceOverflow &ovf)
{
    const dim3 g(grid), b(TRACE_BLOCK);
    if(sched == SCHED_TWO_LEVEL_POSTPONE)
        hipLaunchKernelGGL(k_trace_extension<SCHED_TWO_LEVEL_POSTPONE>, g, b, 0, s, ts, c, path, hit, ovf);
    else if(sched == SCHED_TWO_LEVEL)
        hipLaunchKernelGGL(k_trace_extension<SCHED_TWO_LEVEL>, g, b, 0, s, ts, c, path, hit, ovf);
    else if(sched == SCHED_POSTPONE)
        hipLaunchKernelGGL(k_trace_extension<SCHED_POSTPONE>, g, b, 0, s, ts, c, path, hit, ovf);
    else
        hipLaunchKernelGGL(k_trace_extension<SCHED_EAGER>, g, b, 0, s, ts, c, path, hit, ovf);
}

void launch_shade(hipStream_t s, int grid, int sampler, const RenderConsts *rc, const SceneView &sv,
                  WaveCounters *c, PathSOA path, HitSOA hit, PathSOA next, ShadowSOA shadow, FilmView film,
                  const NrcView &nrc, int parity)
{
    if(sampler == SAMPLER_INDEPENDENT)
        hipLaunchKernelGGL(k_shade<SAMPLER_INDEPENDENT>, dim3(grid), dim3(256), 0, s, rc, sv, c, path, hit, next,
                           shadow, film, nrc, parity);
    else
        hipLaunchKernelGGL(k_shade<SAMPLER_HALTON>, dim3(grid), dim3(256), 0, s, rc, sv, c, path, hit, next, shadow,
                           film, nrc, parity);
}

void launch_medium(hipStream_t s, int grid, int track_grid, int sampler, const RenderConsts *rc,
                   const SceneView &sv, WaveCounters *c, PathSOA path, HitSOA hit, float4 *scatter, PathSOA next,
                   ShadowSOA shadow, FilmView film, int parity)
{
    if(sampler == SAMPLER_INDEPENDENT)
    {
        hipLaunchKernelGGL(k_medium_track<SAMPLER_INDEPENDENT>, dim3(track_grid), dim3(MEDIUM_BLOCK), 0, s, rc, sv,
                           c, path, hit, scatter);
        hipLaunchKernelGGL(k_medium_scatter<SAMPLER_INDEPENDENT>, dim3(grid), dim3(256), 0, s, rc, sv, c, path, hit,
                           scatter, next, shadow, film, parity);
    }
    else
    {
        hipLaunchKernelGGL(k_medium_track<SAMPLER_HALTON>, dim3(track_grid), dim3(MEDIUM_BLOCK), 0, s, rc, sv, c,
                           path, hit, scatter);
        hipLaunchKernelGGL(k_medium_scatter<SAMPLER_HALTON>, dim3(grid), dim3(256), 0, s, rc, sv, c, path, hit,
                           scatter, next, shadow, film, parity);
    }
}

template<int SCHED>
static void launch_shadow_t(hipStream_t s, int grid, bool media, const TraceScene &ts, const SceneView &sv,
                            WaveCounters *c, ShadowSOA shadow, FilmView film, const TraceOverflow &ovf, int parity)
{
    const dim3 g(grid), b(TRACE_BLOCK);
    if(media)
        hipLaunchKernelGGL((k_shadow<SCHED, true>), g, b, 0, s, ts, sv, c, shadow, film, ovf, parity);
    else
        hipLaunchKernelGGL((k_shadow<SCHED, false>), g, b, 0, s, ts, sv, c, shadow, film, ovf, parity);
}

void launch_shadow(hipStream_t s, int grid, int sched, bool media, const TraceScene &ts, const SceneView &sv,
                   WaveCounters *c, ShadowSOA shadow, FilmView film, const TraceOverflow &ovf, int parity)
{
    if(sched == SCHED_TWO_LEVEL_POSTPONE)
        launch_shadow_t<SCHED_TWO_LEVEL_POSTPONE>(s, grid, media, ts, sv, c, shadow, film, ovf, parity);
    else if(sched == SCHED_TWO_LEVEL)
        launch_shadow_t<SCHED_TWO_LEVEL>(s, grid, media, ts, sv, c, shadow, film, ovf, parity);
    else if(sched == SCHED_POSTPONE)
        launch_shadow_t<SCHED_POSTPONE>(s, grid, media, ts, sv, c, shadow, film, ovf, parity);
    else
        launch_shadow_t<SCHED_EAGER>(s, grid, media, ts, sv, c, shadow, film, ovf, parity);
}

void launch_end_iteration(hipStream_t s, WaveCounters *c, int parity)
{
    hipLaunchKernelGGL(k_end_iteration, dim3(1), dim3(1), 0, s, c, parity);
}

void launch_resolve(hipStream_t s, int mode, const float *value, int32_t pixels, float inv_weight, float4 *out)
{
    hipLaunchKernelGGL(k_resolve, dim3((pixels + 255) / 256), dim3(256), 0, s, mode, value, pixels, inv_weight, out);
}

template<int TRI, bool ANY>
static void launch_batch_t(hipStream_t s, int grid, int sched, const TraceScene &ts, const float4 *rays,
                           int32_t n, uint32_t *out, uint32_t *fetch, const TraceOverflow &ovf)
{
    const dim3 g(grid), b(TRACE_BLOCK);
    if(sched == SCHED_TWO_LEVEL_POSTPONE)
        hipLaunchKernelGGL((k_trace_batch<TRI, ANY, SCHED_TWO_LEVEL_POSTPONE>), g, b, 0, s, ts, rays, n, out, fetch, ovf);
    else if(sched == SCHED_TWO_LEVEL)
        hipLaunchKernelGGL((k_trace_batch<TRI, ANY, SCHED_TWO_LEVEL>), g, b, 0, s, ts, rays, n, out, fetch, ovf);
    else if(sched == SCHED_POSTPONE)
        hipLaunchKernelGGL((k_trace_batch<TRI, ANY, SCHED_POSTPONE>), g, b, 0, s, ts, rays, n, out, fetch, ovf);
    else
        hipLaunchKernelGGL((k_trace_batch<TRI, ANY, SCHED_EAGER>), g, b, 0, s, ts, rays, n, out, fetch, ovf);
}

void launch_trace_batch(hipStream_t s, int grid, int sched, const TraceScene &ts, int tri_test, bool any_hit,
                        const float4 *rays, int32_t n, uint32_t *out, uint32_t *fetch, const TraceOverflow &ovf)
{
    if(tri_test == TRI_WATERTIGHT)
    {
        if(any_hit)
            launch_batch_t<TRI_WATERTIGHT, true>(s, grid, sched, ts, rays, n, out, fetch, ovf);
        else
            launch_batch_t<TRI_WATERTIGHT, false>(s, grid, sched, ts, rays, n, out, fetch, ovf);
    }
    else
    {
        if(any_hit)
            launch_batch_t<TRI_MOLLER_TRUMBORE, true>(s, grid, sched, ts, rays, n, out, fetch, ovf);
        else
            launch_batch_t<TRI_MOLLER_TRUMBORE, false>(s, grid, sched, ts, rays, n, out, fetch, ovf);
    }
}

void launch_stream_copy(hipStream_t s, int grid, const void *src, void *dst, int64_t vectors)
{
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, s, (const stream_f4 *)src, (stream_f4 *)dst, vectors);
}

void launch_halton_eval(hipStream_t s, const HaltonConsts &hc, const HaltonTables &tab, const uint32_t *px,
                        const uint32_t *py, const int32_t *sample_index, int32_t n, int32_t dims, float *out)
{
    hipLaunchKernelGGL(k_halton_eval, dim3((n + 127) / 128), dim3(128), 0, s, hc, tab, px, py, sample_index, n, dims,
                       out);
}

void launch_volume_eval(hipStream_t s, const VolumeView &vv, const float4 *seg, const uint2 *seed, int32_t n,
                        float *out)
{
    hipLaunchKernelGGL(k_volume_eval, dim3((n + 127) / 128), dim3(128), 0, s, vv, seg, seed, n, out);
}

void launch_update_world_tris(hipStream_t s, float4 *tris, uint32_t n, const InstanceRec *insts, int32_t geometry,
                              const float *positions)
{
    if(n)
        hipLaunchKernelGGL(k_update_world_tris, dim3((n + 255) / 256), dim3(256), 0, s, tris, n, insts, geometry,
                           positions);
}

void launch_envir_table(hipStream_t s, const TextureView &tv, int32_t tex, int32_t w, int32_t h, int32_t n_samples,
                        float *lum, float *area)
{
    hipLaunchKernelGGL(k_envir_table, dim3((w * h + 63) / 64), dim3(64), 0, s, tv, tex, w, h, n_samples, lum, area);
}

void launch_envir_eval(hipStream_t s, const SceneView &sv, const float *dirs, const float *sam, int32_t n, float *out)
{
    hipLaunchKernelGGL(k_envir_eval, dim3((n + 127) / 128), dim3(128), 0, s, sv, dirs, sam, n, out);
}

void launch_texture_eval(hipStream_t s, const TextureView &tv, int32_t id, const float2 *uv, int32_t n, float4 *out)
{
    hipLaunchKernelGGL(k_texture_eval, dim3((n + 127) / 128), dim3(128), 0, s, tv, id, uv, n, out);
}

void launch_medium_eval(hipStream_t s, const SceneView &sv, int32_t med, const float4 *seg, const uint2 *seed,
                        int32_t n, float *out)
{
    hipLaunchKernelGGL(k_medium_eval, dim3((n + 127) / 128), dim3(128), 0, s, sv, med, seg, seed, n, out);
}

void launch_independent_eval(hipStream_t s, const uint32_t *px, const uint32_t *py, const int32_t *sample_index,
                             int32_t n, int32_t dims, float *out)
{
    hipLaunchKernelGGL(k_independent_eval, dim3((n + 127) / 128), dim3(128), 0, s, px, py, sample_index, n, dims,
                       out);
}

void launch_detmath_eval(hipStream_t s, int32_t fn, const float *x, const float *y, int32_t n, float *out)
{
    hipLaunchKernelGGL(k_detmath_eval, dim3((n + 255) / 256), dim3(256), 0, s, fn, x, y, n, out);
}

void launch_tonemap_aces(hipStream_t s, float4 *rgba, int32_t pixels, float exposure)
{
    if(pixels > 0)
        hipLaunchKernelGGL(k_tonemap_aces, dim3((pixels + 255) / 256), dim3(256), 0, s, rgba, pixels, exposure);
}

void launch_fill_u32(hipStream_t s, uint32_t *p, uint32_t n, uint32_t value)
{
    if(n)
        hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n, value);
}

} // namespace btrc
