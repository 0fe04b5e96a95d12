/*
 * btrc_hip.h - C ABI of the MI355X (gfx950) wavefront path tracer.
 *
 * This is the drop-in boundary for Btrc's render path. Everything behind it
 * (device memory, HIP streams, BVH, kernels) is owned by the library; the
 * caller passes plain host structs and pointers. No HIP or CUDA type appears
 * here. Every function returns BTRC_OK (0) or a negative status; the message
 * of the last failure on a context is available from btrc_hip_last_error.
 *
 * Reference interfaces replaced (paths relative to the Btrc source tree):
 *   Renderer (core/renderer.h:13-59), registered as "wfpt"
 *       (builtin/renderer/wavefront.h:62-69, wavefront.cpp:375-398)
 *   Scene::commit (core/scene.cpp:92-239) and TriangleMesh::commit
 *       (builtin/geometry/triangle_mesh.cpp:24-171)
 *   PinholeCamera::commit (builtin/camera/pinhole.cpp:55-77)
 *   TracePipeline::trace / ShadowPipeline::test
 *       (builtin/renderer/wavefront/trace.cpp:127-141, shadow.cpp:143-157)
 *   Renderer::stop_async (core/renderer.cpp:38-48)
 */
#ifndef BTRC_HIP_H
#define BTRC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct btrc_ctx btrc_ctx;

enum
{
    BTRC_OK = 0,
    BTRC_ERR_INVALID = -1, /* bad argument                                  */
    BTRC_ERR_HIP = -2,     /* HIP runtime error (message has the HIP text)   */
    BTRC_ERR_OOM = -3,     /* device or host allocation failed               */
    BTRC_ERR_STATE = -4,   /* call out of order (e.g. render before scene)   */
    BTRC_ERR_STOPPED = -5  /* render ended by btrc_hip_request_stop          */
};

enum
{
    BTRC_MATERIAL_DIFFUSE = 0,  /* builtin/material/diffuse.cpp          */
    BTRC_MATERIAL_DISNEY = 1,   /* builtin/material/disney.cpp           */
    BTRC_MATERIAL_INVISIBLE = 2 /* builtin/material/invisible.cpp        */
};

enum
{
    BTRC_MEDIUM_HOMOGENEOUS = 0,  /* builtin/medium/homogeneous.cpp     */
    BTRC_MEDIUM_HETEROGENEOUS = 1 /* builtin/medium/hetergeneous.cpp    */
};

enum
{
    BTRC_TEXTURE_RGBA8 = 0,   /* UNorm8x4 */
    BTRC_TEXTURE_RGBA32F = 1  /* F32x4    */
};

enum
{
    BTRC_ADDRESS_CLAMP = 0,
    BTRC_ADDRESS_WRAP = 1,
    BTRC_ADDRESS_MIRROR = 2,
    BTRC_ADDRESS_BORDER = 3
};

enum
{
    BTRC_TEX_FILTER_LINEAR = 0,
    BTRC_TEX_FILTER_POINT = 1
};

enum
{
    BTRC_ACCEL_FLAT = 0,      /* one tree over world-space triangles   */
    BTRC_ACCEL_TWO_LEVEL = 1, /* instance tree over per-geometry trees */
    BTRC_ACCEL_FLAT_GPU = 2   /* one tree over world-space triangles, built on the GPU */
};

enum
{
    BTRC_ENVIR_NONE = 0,
    BTRC_ENVIR_GRADIENT_SKY = 1, /* builtin/light/gradient_sky.cpp       */
    BTRC_ENVIR_IBL = 2           /* builtin/light/ibl.cpp, env_sampler.cpp */
};

enum
{
    BTRC_SAMPLER_HALTON = 0,      /* builtin/sampler/halton.cpp           */
    BTRC_SAMPLER_INDEPENDENT = 1  /* builtin/sampler/independent.cpp      */
};

enum
{
    BTRC_FILTER_BOX = 0,      /* builtin/film_filter/box.cpp              */
    BTRC_FILTER_GAUSSIAN = 1  /* builtin/film_filter/gaussian.cpp         */
};

enum
{
    BTRC_TRI_WATERTIGHT = 0,       /* renderer's triangle test             */
    BTRC_TRI_MOLLER_TRUMBORE = 1   /* lib/bvh Triangle::intersect order    */
};

/* A triangle mesh in its local space, one record per triangle corner as
 * TriangleMeshLoader produces (utils/triangle_mesh_loader.cpp:12-146). */
typedef struct btrc_geometry_desc
{
    int32_t triangle_count;
    const float *positions;  /* 9 floats per triangle: a, b, c                 */
    const float *normals;    /* 9 floats per triangle or NULL (geometric z)     */
    const float *tex_coords; /* 6 floats per triangle or NULL (all zero)        */
} btrc_geometry_desc;

/* Constant-parameter material. Fields unused by a type are ignored.
 * Disney defaults (disney.cpp:844-894): specular 1, ior 1.5, clearcoat_gloss 1,
 * transmission_roughness = roughness, the rest 0. */
typedef struct btrc_material_desc
{
    int32_t type;
    int32_t shadow_terminator; /* default 1 */
    float color[3];
    float metallic, roughness, specular, specular_tint, anisotropic;
    float sheen, sheen_tint, clearcoat, clearcoat_gloss;
    float transmission, transmission_roughness, ior;
    /* Array2D textures in place of the constants (DisneyMaterialCreator,
     * material/disney.cpp:844-894; DiffuseCreator's albedo): 1 + index into
     * btrc_scene_desc::textures, 0 for the constant. Colour reads RGB, the
     * other parameters read the first channel, at the hit's texture
     * coordinate (DisneyMaterial::create_shader, disney.cpp:813-842). The
     * reference's transmission_roughness defaults to the roughness texture
     * itself; a caller mirrors that by passing roughness_texture here too. */
    int32_t color_texture;
    int32_t metallic_texture;
    int32_t roughness_texture;
    int32_t specular_texture, specular_tint_texture, anisotropic_texture;
    int32_t sheen_texture, sheen_tint_texture, clearcoat_texture, clearcoat_gloss_texture;
    int32_t transmission_texture, transmission_roughness_texture, ior_texture;
    /* NormalMap (material/utils/normal_map.cpp:13-21), Disney and diffuse:
     * the texture's RGB as 2 rgb - 1 in the shading frame becomes the new
     * shading normal; 1 + texture index, 0 for none */
    int32_t normal_texture;
} btrc_material_desc;

/* An Array2D texture (texture2d/array2d.cpp, utils/cuda/texture.cpp:63-105,
 * texture2d/description.cpp:32-93). Texels row-major, row 0 at v = 0.
 * RGBA8 texels read as value / 255 (sRGB-decoded when srgb_to_linear);
 * linear filtering blends the four texels around (u W - 0.5, v H - 0.5) with
 * FP32 weights; address modes act on texel indices. */
typedef struct btrc_texture_desc
{
    int32_t width, height;     /* 1 .. 65536 each */
    int32_t format;            /* BTRC_TEXTURE_* */
    const void *texels;        /* RGBA8: 4 bytes per texel; RGBA32F: 4 floats */
    int32_t address_u;         /* BTRC_ADDRESS_*, default clamp */
    int32_t address_v;
    int32_t filter;            /* BTRC_TEX_FILTER_*, default linear */
    int32_t srgb_to_linear;
    float border[4];
} btrc_texture_desc;

/* Scene::Instance (core/scene.h:71-80) */
typedef struct btrc_instance_desc
{
    int32_t geometry;
    int32_t material;
    float local_to_world[16]; /* row-major Transform3D::mat */
    float world_to_local[16]; /* row-major Transform3D::inv */
    int32_t has_light;        /* mesh_light on this instance */
    float light_intensity[3];
    int32_t inner_medium;     /* index into btrc_scene_desc::media, -1: none */
    int32_t outer_medium;     /* index into btrc_scene_desc::media, -1: none */
} btrc_instance_desc;

/* A participating medium, referenced by instances' inner/outer_medium.
 * HomogeneousMedium: sigma_t, albedo, g. HetergeneousMedium: a density grid
 * (Array3D text layout, utils/cuda/array.cpp:313-385: x fastest, then y,
 * then z) sampled at 0.5 + 0.5 * uvw with linear filtering and clamping,
 * albedo from albedo_grid (3 floats per voxel) or the constant albedo, and
 * a constant g. With has_transform, uvw = world_to_local(p)
 * (TransformMedium, core/medium.cpp:98-129); else uvw = p. */
typedef struct btrc_medium_desc
{
    int32_t type;             /* BTRC_MEDIUM_* */
    float priority;           /* where media meet, the higher wins (core/scene.cpp:144-161) */
    float sigma_t;            /* homogeneous extinction */
    float albedo[3];
    float g;                  /* Henyey-Greenstein asymmetry, |g| < 1 */
    int32_t grid_size[3];     /* heterogeneous: width, height, depth */
    const float *sigma_t_grid;
    const float *albedo_grid; /* or NULL */
    int32_t has_transform;
    float world_to_local[16]; /* row-major */
} btrc_medium_desc;

/* A volume primitive (VolumePrimitive, core/volume.cpp:21-79; the scene's
 * "volumes" array, factory/scene.cpp:57-79): the parallelepiped
 * o + u x + v y + w z, u, v, w in [0, 1], with a density sigma_t(u, v, w)
 * and an albedo(u, v, w), each a grid (Array3D layout, x fastest, linear
 * filtering with clamping at texel centres (i + 0.5) / n) or a constant.
 * All volume primitives form the scene's volume-primitive medium, the
 * lowest-priority medium that camera rays start in and that fills
 * everything outside the instances' media; where volumes overlap their
 * densities add, and scattering in them is isotropic. */
typedef struct btrc_volume_desc
{
    float o[3], x[3], y[3], z[3];
    int32_t sigma_t_size[3];     /* grid width, height, depth; 0, 0, 0: constant sigma_t */
    const float *sigma_t_grid;
    float sigma_t;
    int32_t albedo_size[3];      /* 0, 0, 0: constant albedo */
    const float *albedo_grid;    /* 3 floats per voxel */
    float albedo[3];
} btrc_volume_desc;

typedef struct btrc_envir_light_desc
{
    int32_t type;   /* BTRC_ENVIR_* */
    float up[3];    /* both types: the direction that maps to the sky's top */
    float lower[3]; /* GRADIENT_SKY */
    float upper[3]; /* GRADIENT_SKY */
    /* IBL (IBLCreator, light/ibl.cpp:80-91): radiance texture (an index
     * into btrc_scene_desc.textures) looked up at (phi / 2pi, theta / pi) of
     * the direction in the frame whose z axis is `up`, and the resolution of
     * the importance table over that (u, v) square (the creator's defaults
     * are 400 x 200) */
    int32_t texture;
    int32_t lut_width, lut_height;
} btrc_envir_light_desc;

typedef struct btrc_scene_desc
{
    int32_t geometry_count;
    const btrc_geometry_desc *geometries;
    int32_t material_count;
    const btrc_material_desc *materials;
    int32_t instance_count;
    const btrc_instance_desc *instances;
    btrc_envir_light_desc envir_light;
    int32_t medium_count;
    const btrc_medium_desc *media;
    /* BTRC_ACCEL_*: FLAT traces one tree over the world-space triangles;
     * TWO_LEVEL traces an instance tree over one tree per geometry (the
     * TLAS/BLAS split of Scene::commit, core/scene.cpp:92-239), transforming
     * the ray into each instance's local space. */
    int32_t accel;
    int32_t texture_count;
    const btrc_texture_desc *textures;
    int32_t volume_count;
    const btrc_volume_desc *volumes;
} btrc_scene_desc;

/* PinholeCamera (builtin/camera/pinhole.cpp) */
typedef struct btrc_camera_desc
{
    float eye[3];
    float dst[3];
    float up[3];
    float fov_y_deg;
    float w_over_h;
} btrc_camera_desc;

/* WavefrontPathTracer::Params (builtin/renderer/wavefront.h:13-27) */
typedef struct btrc_render_params
{
    int32_t width, height;
    int32_t spp;
    int32_t min_depth, max_depth;
    float rr_threshold, rr_cont_prob;
    int32_t state_count;
    int32_t film_fp64;   /* also accumulate an FP64 film */
    int32_t time_stages; /* record per-stage kernel time with HIP events */
    int32_t tile;        /* Params::tile: path i is sample i % spp of pixel i / spp */
    int32_t sampler;     /* BTRC_SAMPLER_*; the reference's wfpt uses INDEPENDENT */
    int32_t albedo;      /* Params::albedo: first-hit albedo output */
    int32_t normal;      /* Params::normal: first-hit shading normal output */
    int32_t sample_offset; /* first sample index (progressive frames) */
    int32_t accumulate;  /* add to the previous render's film instead of clearing it */
} btrc_render_params;

/* Neural radiance cache (no counterpart in the reference; README.md:17
 * lists it as future work). With `enabled`, every render (spp 1, no media)
 * ends non-training paths at vertex depth `query_depth`, where the cache's
 * prediction times the path throughput goes to the film; paths of one pixel
 * in `train_stride` (rotating with sample_offset) are traced to the end and
 * their vertices (up to 65536) train the cache after the frame: `steps`
 * Adam steps over batches of 16384 vertices, relative L2 loss. */
typedef struct btrc_nrc_params
{
    int32_t enabled;
    int32_t train;          /* 1: train after each frame */
    int32_t query_depth;    /* default 2 */
    int32_t train_stride;   /* default 16 */
    int32_t steps;          /* optimizer steps per frame, default 4 */
    float learning_rate;    /* default 1e-2 */
    float loss_scale;       /* gradient scale of the BF16 backward pass, default 128 */
    /* 0 (default): training paths are traced to the end, so targets are
     * unbiased path estimates. k >= query_depth: training paths end at depth
     * k with a cache query whose answer joins their radiance sums, so the
     * bounce loop ends after about k bounces (the cache trains on its own
     * tails, as the radiance cache paper's self-training does); their
     * vertices at depths 0..k-1 are the training records. k = query_depth
     * adds no bounce beyond the rendered paths' own */
    int32_t train_depth;
    /* 0 (default): frames query the trained weights. d in (0, 1): they
     * query an exponential moving average e <- d e + (1 - d) w, updated after
     * every optimizer step (the radiance cache paper's averaged weights) */
    float ema_decay;
} btrc_nrc_params;

typedef struct btrc_nrc_stats
{
    int64_t queries;         /* cache queries of the last frame */
    int64_t records;         /* training vertices of the last frame */
    int64_t steps_total;     /* optimizer steps so far */
    double loss[4];          /* per step of the last frame (relative L2, FP64 sum) */
    double grad_norm[4];
    int32_t skipped;         /* steps of the last frame skipped (non-finite gradient) */
    int32_t reserved;
} btrc_nrc_stats;

/* FilmFilter (core/film_filter.h); GaussianFilter defaults radius 1.3,
 * alpha 1.6 (gaussian.cpp:68-76) */
typedef struct btrc_filter_desc
{
    int32_t type; /* BTRC_FILTER_* */
    float radius;
    float alpha;
} btrc_filter_desc;

typedef struct btrc_render_stats
{
    double seconds;          /* wall time of btrc_hip_render */
    uint64_t extension_rays; /* closest-hit rays traced */
    uint64_t shadow_rays;    /* visibility rays traced */
    uint64_t paths;          /* camera paths completed */
    uint64_t iterations;     /* wavefront bounces executed */
    double stage_ms[7];      /* generate, trace, shade, shadow, medium, cache inference,
                                cache training (time_stages) */
    uint64_t stage_launches[7];
} btrc_render_stats;

/* A ray of a traversal batch and its closest hit */
typedef struct btrc_ray
{
    float o[3];
    float tmin;
    float d[3];
    float tmax;
} btrc_ray;

typedef struct btrc_hit
{
    float t;           /* tmax on a miss */
    uint32_t instance; /* 0xffffffff on a miss */
    uint32_t prim;     /* triangle index within the instance's geometry */
    float u, v;        /* barycentric weights of vertices b and c */
} btrc_hit;

/* Scheduling knobs of the kernels; they change speed, never results.
 * Picked per device by btrc_amd/autotune.py. */
typedef struct btrc_tuning
{
    int32_t closest_postpone_leaves; /* one-level scenes, 1: leaves wait until every lane holds one (default 0) */
    int32_t any_postpone_leaves;     /* same for visibility rays (default 0) */
    int32_t trace_waves_per_cu;      /* persistent traversal grid, 1..32 (default 20) */
    int32_t launch_graph;            /* 1: the bounce loop replays a captured HIP graph (default 1) */
    int32_t two_level_postpone;      /* two-level scenes: bit 0 closest-hit, bit 1 visibility rays
                                        postpone leaves (default 0: both test leaves when reached) */
    int32_t stage_blocks_per_cu;     /* grid of the 256-lane stage kernels (generate, medium, shade): at most
                                        this many workgroups per CU, 1..64 (0: 32) */
    int32_t graph_bounces;           /* bounces per captured graph block between counter read-backs,
                                        even, 2..16 (0: 4) */
    int32_t reserved;
} btrc_tuning;

/* ---- context ---- */
int btrc_hip_create(int device, btrc_ctx **out);
void btrc_hip_destroy(btrc_ctx *ctx);
const char *btrc_hip_last_error(const btrc_ctx *ctx);
int btrc_hip_device_name(btrc_ctx *ctx, char *buf, size_t len);
int btrc_hip_set_tuning(btrc_ctx *ctx, const btrc_tuning *tuning);
int btrc_hip_get_tuning(btrc_ctx *ctx, btrc_tuning *tuning);

/* ---- scene and camera (Scene::commit, PinholeCamera::commit) ---- */
int btrc_hip_upload_scene(btrc_ctx *ctx, const btrc_scene_desc *scene);
int btrc_hip_set_camera(btrc_ctx *ctx, const btrc_camera_desc *camera);
/* statistics of the built BVH: nodes, depth, stack bound, triangles */
int btrc_hip_scene_info(btrc_ctx *ctx, int64_t *nodes, int32_t *depth, int32_t *stack_bound, int64_t *triangles);
/* wall time of the last upload's acceleration-structure build, in ms
 * (the host builder's, or the GPU builder's kernels and read-backs) */
int btrc_hip_scene_build_ms(btrc_ctx *ctx, double *ms);
/* New vertex data (same triangle count; positions, optional normals and
 * texture coordinates as at upload) for geometry `geometry` of a scene
 * uploaded with BTRC_ACCEL_FLAT or BTRC_ACCEL_FLAT_GPU: the geometry's
 * shading and light-sampling records are rebuilt and the tree is refitted
 * in place (same topology, every box the union of what it holds), so the
 * next render sees the moved triangles without a rebuild. *ms (may be NULL)
 * gets the update's wall time. No counterpart in the reference, which builds
 * its structures once per commit (core/scene.cpp:92-239). */
int btrc_hip_update_geometry(btrc_ctx *ctx, int32_t geometry, const btrc_geometry_desc *desc, double *ms);

/* WavefrontPathTracer::set_film_filter (wavefront.cpp:72-75); box until set */
int btrc_hip_set_filter(btrc_ctx *ctx, const btrc_filter_desc *filter);

/* ---- rendering (WavefrontPathTracer::render) ---- */
int btrc_hip_render(btrc_ctx *ctx, const btrc_render_params *params, btrc_render_stats *stats);
/* Several views in one wavefront (no counterpart in the reference, which
 * renders one camera per Renderer): the views' paths share the pool and
 * every stage launch, so small films still fill the GPU. View v gets the
 * samples btrc_hip_render would give it with its camera and size;
 * params->width and height are ignored. 1 .. 64 views. */
typedef struct btrc_view_desc
{
    btrc_camera_desc camera;
    int32_t width, height;
} btrc_view_desc;
int btrc_hip_render_views(btrc_ctx *ctx, const btrc_render_params *params, int32_t view_count,
                          const btrc_view_desc *views, btrc_render_stats *stats);
/* value / weight per pixel as RGBA (alpha 1), row 0 at the top, all views of
 * the last render one after another; rgb64 (3 doubles per pixel, same
 * normalisation) needs film_fp64, may be NULL */
int btrc_hip_film_readback(btrc_ctx *ctx, float *rgba, double *rgb64);
/* RGBA of one view of the last render */
int btrc_hip_view_readback(btrc_ctx *ctx, int32_t view, float *rgba);
/* RenderResult::albedo / normal (wavefront.cpp:320-350, preview.cpp:28-80):
 * RGBA per pixel, alpha 1; either pointer may be NULL; each needs the
 * matching flag in the last render's params */
int btrc_hip_aov_readback(btrc_ctx *ctx, float *albedo_rgba, float *normal_rgba);
/* The "aces" post processor (ACESToneMap::process, postprocess/tonemap.cpp:
 * 14-24,35-51,86-97) on the device: the resolved colour film (as
 * btrc_hip_film_readback returns it) through Narkowicz's ACES fit of
 * exposure * value per channel, saturated to [0, 1]; alpha unchanged. */
int btrc_hip_tonemap_readback(btrc_ctx *ctx, float exposure, float *rgba);
/* thread-safe; the running render returns BTRC_ERR_STOPPED */
int btrc_hip_request_stop(btrc_ctx *ctx);

/* Reporter (core/reporter.h:7-36; set by Renderer::set_reporter,
 * core/renderer.h): callbacks made from the thread that calls a render
 * entry point, during that call. new_stage when a render starts,
 * progress(percentage of the frame's paths finished) at every read-back of
 * the device counters (once per block of captured bounces, as the
 * reference reports once per bounce, wavefront.cpp:287), complete_stage
 * when the bounce loop ends, also when it stops early. Any callback may be
 * NULL. The reference's preview images (need_preview / new_preview) are a
 * GUI feature and are not called. */
typedef struct btrc_reporter
{
    void *user;
    void (*new_stage)(void *user, const char *name);
    void (*complete_stage)(void *user);
    void (*progress)(void *user, float percentage);
} btrc_reporter;
/* copies *reporter; NULL removes it */
int btrc_hip_set_reporter(btrc_ctx *ctx, const btrc_reporter *reporter);

/* ---- neural radiance cache ---- */
int btrc_hip_nrc_configure(btrc_ctx *ctx, const btrc_nrc_params *params);
/* fresh weights (uniform He initialisation from PCG32 stream `seed`), zero
 * Adam state */
int btrc_hip_nrc_reset(btrc_ctx *ctx, uint64_t seed);
int btrc_hip_nrc_stats(btrc_ctx *ctx, btrc_nrc_stats *stats);
/* the 22528 stored FP32 weights: six layers [out][in], the output layer as
 * 32 rows of which 3 are used */
int btrc_hip_nrc_get_weights(btrc_ctx *ctx, float *weights);
int btrc_hip_nrc_set_weights(btrc_ctx *ctx, const float *weights);
/* the moving average of the weights (btrc_nrc_params.ema_decay), same layout;
 * it restarts at the weights on every reset and set */
int btrc_hip_nrc_get_ema_weights(btrc_ctx *ctx, float *weights);
/* network outputs for n vertex records (16 floats each: position,
 * direction to the previous vertex, shading normal, roughness, diffuse,
 * specular); out: 3 floats each */
int btrc_hip_nrc_eval(btrc_ctx *ctx, const float *vertices, int32_t n, float *out);
/* gradient of the mean relative L2 loss over n samples (n <= 65536) with
 * respect to the 22528 weights, and the loss */
int btrc_hip_nrc_gradients(btrc_ctx *ctx, const float *vertices, const float *targets, int32_t n, float *grad,
                           double *loss);
/* `steps` Adam steps on one batch (n <= 16384); per-step loss out (may be NULL) */
int btrc_hip_nrc_fit(btrc_ctx *ctx, const float *vertices, const float *targets, int32_t n, int32_t steps,
                     double *loss);

/* ---- traversal on ray batches in host memory ---- */
int btrc_hip_trace_closest(btrc_ctx *ctx, const btrc_ray *rays, int64_t n, int32_t tri_test, btrc_hit *hits);
int btrc_hip_trace_any(btrc_ctx *ctx, const btrc_ray *rays, int64_t n, int32_t tri_test, uint8_t *occluded);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): btrc_hip_trace_closest on rays and hits that already
 * live in device memory (`rays` and `hits` are device or managed
 * allocations holding n entries each). It allocates nothing: the
 * traversal's ray counter and its per-lane stack entries beyond the LDS part
 * live in a static 8 MiB device array per GPU, which caps the persistent
 * grid and fails with BTRC_ERR_INVALID when the scene's tree needs more
 * stack than the array holds. Calls that overlap on one GPU race, since they
 * share that array. Returns once the hits are written. */
int btrc_hip_trace_closest_noalloc(btrc_ctx *ctx, const btrc_ray *rays, int64_t n, int32_t tri_test, btrc_hit *hits);

/* ---- device-resident ray batch (throughput measurement) ---- */
int btrc_hip_batch_upload(btrc_ctx *ctx, const btrc_ray *rays, int64_t n);
/* traces the resident batch `repeats` times; *kernel_ms = mean kernel time */
int btrc_hip_batch_trace(btrc_ctx *ctx, int32_t tri_test, int32_t any_hit, int32_t repeats, double *kernel_ms);
int btrc_hip_batch_download(btrc_ctx *ctx, btrc_hit *hits);
/* No counterpart in the reference (a measurement aid): the device's copy
 * bandwidth, from a STREAM-style copy kernel moving `bytes` (a positive
 * multiple of 16) between two device buffers allocated for the call;
 * *gb_per_s = bytes read + bytes written per second (GB/s), the median of
 * `repeats` copies timed with HIP events after one untimed copy */
int btrc_hip_measure_copy(btrc_ctx *ctx, int64_t bytes, int32_t repeats, double *gb_per_s);

/* ---- rays of a render, for offline comparison ----
 * Renders like btrc_hip_render and copies the extension rays of the first
 * `bounces` bounces of the first pool wave (at most max_rays). */
int btrc_hip_capture_rays(btrc_ctx *ctx, const btrc_render_params *params, int32_t bounces,
                          btrc_ray *rays, int64_t max_rays, int64_t *count);

/* lib/bvh's Bvh<float>::Node (lib/bvh/include/bvh/bvh.hpp:25-28): bounds
 * min x, max x, min y, max y, min z, max z; a leaf holds primitive_count
 * primitives from first_child_or_primitive in the primitive index list, an
 * inner node (primitive_count 0) its children at first_child_or_primitive
 * and the next index */
typedef struct btrc_bvh_node
{
    float bounds[6];
    uint32_t primitive_count;
    uint32_t first_child_or_primitive;
} btrc_bvh_node;

/* ---- GPU BVH build on host-given boxes ----
 * A binary tree over n boxes (6 floats each: lo xyz, hi xyz) built on the
 * GPU the way lib/bvh's LocallyOrderedClusteringBuilder<Bvh<float>,
 * uint64_t> builds it (locally_ordered_clustering_builder.hpp; box centres
 * 0.5 (lo + hi), 63-bit Morton codes over the union box, search radius 14),
 * then LeafCollapser (leaf_collapser.hpp) unless max_leaf_size < 0
 * (0: no leaf size limit, as lib/bvh; k > 0: leaves of at most k). With
 * refit_boxes, HierarchyRefitter (hierarchy_refitter.hpp) then moves the
 * tree onto those boxes. nodes needs room for 2n - 1, primitive_indices
 * for n; *node_count gets the nodes written, *build_ms the build's wall
 * time (may be NULL). Replaces the CPU builders of core/volume/bvh.cpp:40-63
 * and, for triangles, optixAccelBuild (utils/optix/context.cpp:158-295). */
int btrc_hip_build_bvh(btrc_ctx *ctx, const float *boxes, int32_t n, int32_t max_leaf_size, const float *refit_boxes,
                       btrc_bvh_node *nodes, uint32_t *primitive_indices, int32_t *node_count, double *build_ms);

/* ---- test hooks: evaluate device building blocks on host-given inputs ---- */
int btrc_hip_eval_halton(btrc_ctx *ctx, int32_t width, int32_t height, const uint32_t *px, const uint32_t *py,
                         const int32_t *sample_index, int32_t n, int32_t dims, float *out);
/* the independent sampler's first `dims` numbers for each (px, py, sample) */
int btrc_hip_eval_independent(btrc_ctx *ctx, const uint32_t *px, const uint32_t *py, const int32_t *sample_index,
                              int32_t n, int32_t dims, float *out);
/* the volume-primitive medium along segments o -> o + d * tmax, one PCG32
 * stream per segment as in btrc_hip_eval_medium: out holds 10 floats per
 * segment (scattered, position xyz, albedo xyz, tr xyz) */
int btrc_hip_eval_volumes(btrc_ctx *ctx, const btrc_ray *segments, const uint64_t *seeds, int32_t n, float *out);
/* texture `texture` of the current scene sampled at n (u, v) pairs; out
 * gets 4 floats per pair */
int btrc_hip_eval_texture(btrc_ctx *ctx, int32_t texture, const float *uv, int32_t n, float *out);
/* the current scene's environment light at n world directions (3 floats
 * each) and n sample triples (3 floats each); out gets 12 floats per query:
 * eval_le rgb, pdf_li, then sample_li's direction xyz, pdf and radiance rgb,
 * then 0 (EnvirLight::eval_le / pdf_li / sample_li, core/light.h:90-132) */
int btrc_hip_eval_envir(btrc_ctx *ctx, const float *dirs, const float *samples, int32_t n, float *out);
/* the IBL's importance table (tile probabilities, row-major lut_height x
 * lut_width; EnvirLightSampler::preprocess, env_sampler.cpp:72-122); *count
 * gets the table size, at most `capacity` floats are written */
int btrc_hip_envir_table(btrc_ctx *ctx, float *probs, int32_t capacity, int32_t *count);
/* medium id (rank by descending priority) along segments o -> o + d * tmax,
 * one PCG32 stream per segment (seeded like a shadow ray's): out holds 10
 * floats per segment: scattered (0/1), position xyz, albedo xyz, tr xyz */
int btrc_hip_eval_medium(btrc_ctx *ctx, int32_t medium, const btrc_ray *segments, const uint64_t *seeds, int32_t n,
                         float *out);
/* fn: 0 sin, 1 cos, 2 log, 3 exp, 4 pow(x, y), 5 sqrt, 6 x / y, 7 atan2(x, y), 8 acos,
 * 9 one axis of the self-intersection offset (x = position, y = normal coordinate;
 *   intersection_offset, utils/intersection.cpp:5-31, acts on each axis alone) */
int btrc_hip_eval_detmath(btrc_ctx *ctx, int32_t fn, const float *x, const float *y, int32_t n, float *out);

#ifdef __cplusplus
}
#endif

#endif /* BTRC_HIP_H */
