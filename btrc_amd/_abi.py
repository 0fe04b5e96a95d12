"""ctypes mirror of include/btrc_hip.h (plain C structs, no HIP types)."""
import ctypes as C

BTRC_OK = 0
BTRC_ERR_INVALID = -1
BTRC_ERR_HIP = -2
BTRC_ERR_OOM = -3
BTRC_ERR_STATE = -4
BTRC_ERR_STOPPED = -5

MATERIAL_DIFFUSE = 0
MATERIAL_DISNEY = 1
MATERIAL_INVISIBLE = 2
MEDIUM_HOMOGENEOUS = 0
MEDIUM_HETEROGENEOUS = 1
TEXTURE_RGBA8 = 0
TEXTURE_RGBA32F = 1
ADDRESS_CLAMP = 0
ADDRESS_WRAP = 1
ADDRESS_MIRROR = 2
ADDRESS_BORDER = 3
TEX_FILTER_LINEAR = 0
TEX_FILTER_POINT = 1
ACCEL_FLAT = 0
ACCEL_TWO_LEVEL = 1
ACCEL_FLAT_GPU = 2
ENVIR_NONE = 0
ENVIR_GRADIENT_SKY = 1
ENVIR_IBL = 2
SAMPLER_HALTON = 0
SAMPLER_INDEPENDENT = 1
FILTER_BOX = 0
FILTER_GAUSSIAN = 1
TRI_WATERTIGHT = 0
TRI_MOLLER_TRUMBORE = 1

F3 = C.c_float * 3
F16 = C.c_float * 16


class GeometryDesc(C.Structure):
    _fields_ = [("triangle_count", C.c_int32),
                ("positions", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)),
                ("tex_coords", C.POINTER(C.c_float))]


class MaterialDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("shadow_terminator", C.c_int32), ("color", F3),
                ("metallic", C.c_float), ("roughness", C.c_float), ("specular", C.c_float),
                ("specular_tint", C.c_float), ("anisotropic", C.c_float),
                ("sheen", C.c_float), ("sheen_tint", C.c_float), ("clearcoat", C.c_float),
                ("clearcoat_gloss", C.c_float), ("transmission", C.c_float),
                ("transmission_roughness", C.c_float), ("ior", C.c_float),
                ("color_texture", C.c_int32), ("metallic_texture", C.c_int32), ("roughness_texture", C.c_int32),
                ("specular_texture", C.c_int32), ("specular_tint_texture", C.c_int32),
                ("anisotropic_texture", C.c_int32), ("sheen_texture", C.c_int32), ("sheen_tint_texture", C.c_int32),
                ("clearcoat_texture", C.c_int32), ("clearcoat_gloss_texture", C.c_int32),
                ("transmission_texture", C.c_int32), ("transmission_roughness_texture", C.c_int32),
                ("ior_texture", C.c_int32), ("normal_texture", C.c_int32)]


# texture slots of btrc_material_desc, in declaration order (Disney parameters, then the normal map)
MATERIAL_TEXTURE_SLOTS = ("color", "metallic", "roughness", "specular", "specular_tint", "anisotropic", "sheen",
                          "sheen_tint", "clearcoat", "clearcoat_gloss", "transmission", "transmission_roughness",
                          "ior", "normal")


class TextureDesc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("format", C.c_int32), ("texels", C.c_void_p),
                ("address_u", C.c_int32), ("address_v", C.c_int32), ("filter", C.c_int32),
                ("srgb_to_linear", C.c_int32), ("border", C.c_float * 4)]


class InstanceDesc(C.Structure):
    _fields_ = [("geometry", C.c_int32), ("material", C.c_int32),
                ("local_to_world", F16), ("world_to_local", F16),
                ("has_light", C.c_int32), ("light_intensity", F3),
                ("inner_medium", C.c_int32), ("outer_medium", C.c_int32)]


class MediumDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("priority", C.c_float), ("sigma_t", C.c_float), ("albedo", F3),
                ("g", C.c_float), ("grid_size", C.c_int32 * 3), ("sigma_t_grid", C.POINTER(C.c_float)),
                ("albedo_grid", C.POINTER(C.c_float)), ("has_transform", C.c_int32), ("world_to_local", F16)]


class EnvirLightDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("up", F3), ("lower", F3), ("upper", F3), ("texture", C.c_int32),
                ("lut_width", C.c_int32), ("lut_height", C.c_int32)]


class VolumeDesc(C.Structure):
    _fields_ = [("o", F3), ("x", F3), ("y", F3), ("z", F3), ("sigma_t_size", C.c_int32 * 3),
                ("sigma_t_grid", C.POINTER(C.c_float)), ("sigma_t", C.c_float), ("albedo_size", C.c_int32 * 3),
                ("albedo_grid", C.POINTER(C.c_float)), ("albedo", F3)]


class SceneDesc(C.Structure):
    _fields_ = [("geometry_count", C.c_int32), ("geometries", C.POINTER(GeometryDesc)),
                ("material_count", C.c_int32), ("materials", C.POINTER(MaterialDesc)),
                ("instance_count", C.c_int32), ("instances", C.POINTER(InstanceDesc)),
                ("envir_light", EnvirLightDesc), ("medium_count", C.c_int32), ("media", C.POINTER(MediumDesc)),
                ("accel", C.c_int32), ("texture_count", C.c_int32), ("textures", C.POINTER(TextureDesc)),
                ("volume_count", C.c_int32), ("volumes", C.POINTER(VolumeDesc))]


class CameraDesc(C.Structure):
    _fields_ = [("eye", F3), ("dst", F3), ("up", F3), ("fov_y_deg", C.c_float), ("w_over_h", C.c_float)]


class ViewDesc(C.Structure):
    _fields_ = [("camera", CameraDesc), ("width", C.c_int32), ("height", C.c_int32)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32),
                ("min_depth", C.c_int32), ("max_depth", C.c_int32),
                ("rr_threshold", C.c_float), ("rr_cont_prob", C.c_float),
                ("state_count", C.c_int32), ("film_fp64", C.c_int32), ("time_stages", C.c_int32),
                ("tile", C.c_int32), ("sampler", C.c_int32), ("albedo", C.c_int32), ("normal", C.c_int32),
                ("sample_offset", C.c_int32), ("accumulate", C.c_int32)]


class NrcParams(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("train", C.c_int32), ("query_depth", C.c_int32),
                ("train_stride", C.c_int32), ("steps", C.c_int32), ("learning_rate", C.c_float),
                ("loss_scale", C.c_float), ("train_depth", C.c_int32), ("ema_decay", C.c_float)]


class NrcStats(C.Structure):
    _fields_ = [("queries", C.c_int64), ("records", C.c_int64), ("steps_total", C.c_int64),
                ("loss", C.c_double * 4), ("grad_norm", C.c_double * 4), ("skipped", C.c_int32),
                ("reserved", C.c_int32)]


NRC_PARAMS = 22528
NRC_VERTEX_FLOATS = 16


class FilterDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("radius", C.c_float), ("alpha", C.c_float)]


class RenderStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("paths", C.c_uint64), ("iterations", C.c_uint64),
                ("stage_ms", C.c_double * 7), ("stage_launches", C.c_uint64 * 7)]


class Tuning(C.Structure):
    _fields_ = [("closest_postpone_leaves", C.c_int32), ("any_postpone_leaves", C.c_int32),
                ("trace_waves_per_cu", C.c_int32), ("launch_graph", C.c_int32),
                ("two_level_postpone", C.c_int32), ("stage_blocks_per_cu", C.c_int32),
                ("graph_bounces", C.c_int32), ("reserved", C.c_int32)]


class BvhNode(C.Structure):
    _fields_ = [("bounds", C.c_float * 6), ("primitive_count", C.c_uint32), ("first_child_or_primitive", C.c_uint32)]


def bvh_node_dtype():
    import numpy as np
    return np.dtype([("bounds", np.float32, 6), ("primitive_count", np.uint32), ("first", np.uint32)])


class Ray(C.Structure):
    _fields_ = [("o", F3), ("tmin", C.c_float), ("d", F3), ("tmax", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("instance", C.c_uint32), ("prim", C.c_uint32),
                ("u", C.c_float), ("v", C.c_float)]


# numpy dtypes with the same layout
def ray_dtype():
    import numpy as np
    return np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])


def hit_dtype():
    import numpy as np
    return np.dtype([("t", "<f4"), ("instance", "<u4"), ("prim", "<u4"), ("u", "<f4"), ("v", "<f4")])


REPORTER_STAGE = C.CFUNCTYPE(None, C.c_void_p, C.c_char_p)
REPORTER_DONE = C.CFUNCTYPE(None, C.c_void_p)
REPORTER_PROGRESS = C.CFUNCTYPE(None, C.c_void_p, C.c_float)


class Reporter(C.Structure):
    _fields_ = [("user", C.c_void_p), ("new_stage", REPORTER_STAGE), ("complete_stage", REPORTER_DONE),
                ("progress", REPORTER_PROGRESS)]


# every symbol include/btrc_hip.h declares, in declaration order
EXPORTS = [
    "btrc_hip_create", "btrc_hip_destroy", "btrc_hip_last_error", "btrc_hip_device_name",
    "btrc_hip_set_tuning", "btrc_hip_get_tuning",
    "btrc_hip_upload_scene", "btrc_hip_set_camera", "btrc_hip_scene_info", "btrc_hip_scene_build_ms",
    "btrc_hip_update_geometry",
    "btrc_hip_set_filter",
    "btrc_hip_render", "btrc_hip_render_views", "btrc_hip_film_readback", "btrc_hip_view_readback",
    "btrc_hip_aov_readback", "btrc_hip_tonemap_readback", "btrc_hip_request_stop",
    "btrc_hip_set_reporter",
    "btrc_hip_nrc_configure", "btrc_hip_nrc_reset", "btrc_hip_nrc_stats", "btrc_hip_nrc_get_weights",
    "btrc_hip_nrc_set_weights", "btrc_hip_nrc_get_ema_weights", "btrc_hip_nrc_eval", "btrc_hip_nrc_gradients", "btrc_hip_nrc_fit",
    "btrc_hip_trace_closest", "btrc_hip_trace_any", "btrc_hip_trace_closest_noalloc",
    "btrc_hip_batch_upload", "btrc_hip_batch_trace", "btrc_hip_batch_download", "btrc_hip_measure_copy",
    "btrc_hip_capture_rays", "btrc_hip_build_bvh", "btrc_hip_eval_halton", "btrc_hip_eval_independent", "btrc_hip_eval_volumes", "btrc_hip_eval_texture",
    "btrc_hip_eval_envir", "btrc_hip_envir_table", "btrc_hip_eval_medium",
    "btrc_hip_eval_detmath",
]


def declare(lib):
    """Attach argtypes/restype to the library's exported functions."""
    P = C.POINTER
    ctx = C.c_void_p
    sig = {
        "btrc_hip_create": (C.c_int, [C.c_int, P(C.c_void_p)]),
        "btrc_hip_destroy": (None, [ctx]),
        "btrc_hip_last_error": (C.c_char_p, [ctx]),
        "btrc_hip_device_name": (C.c_int, [ctx, C.c_char_p, C.c_size_t]),
        "btrc_hip_set_tuning": (C.c_int, [ctx, P(Tuning)]),
        "btrc_hip_get_tuning": (C.c_int, [ctx, P(Tuning)]),
        "btrc_hip_upload_scene": (C.c_int, [ctx, P(SceneDesc)]),
        "btrc_hip_set_camera": (C.c_int, [ctx, P(CameraDesc)]),
        "btrc_hip_scene_info": (C.c_int, [ctx, P(C.c_int64), P(C.c_int32), P(C.c_int32), P(C.c_int64)]),
        "btrc_hip_scene_build_ms": (C.c_int, [ctx, P(C.c_double)]),
        "btrc_hip_update_geometry": (C.c_int, [ctx, C.c_int32, C.c_void_p, P(C.c_double)]),
        "btrc_hip_build_bvh": (C.c_int, [ctx, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                         P(C.c_int32), P(C.c_double)]),
        "btrc_hip_set_filter": (C.c_int, [ctx, P(FilterDesc)]),
        "btrc_hip_render": (C.c_int, [ctx, P(RenderParams), P(RenderStats)]),
        "btrc_hip_render_views": (C.c_int, [ctx, P(RenderParams), C.c_int32, P(ViewDesc), P(RenderStats)]),
        "btrc_hip_view_readback": (C.c_int, [ctx, C.c_int32, C.c_void_p]),
        "btrc_hip_aov_readback": (C.c_int, [ctx, C.c_void_p, C.c_void_p]),
        "btrc_hip_tonemap_readback": (C.c_int, [ctx, C.c_float, C.c_void_p]),
        "btrc_hip_film_readback": (C.c_int, [ctx, C.c_void_p, C.c_void_p]),
        "btrc_hip_request_stop": (C.c_int, [ctx]),
        "btrc_hip_set_reporter": (C.c_int, [ctx, C.POINTER(Reporter)]),
        "btrc_hip_nrc_configure": (C.c_int, [ctx, P(NrcParams)]),
        "btrc_hip_nrc_reset": (C.c_int, [ctx, C.c_uint64]),
        "btrc_hip_nrc_stats": (C.c_int, [ctx, P(NrcStats)]),
        "btrc_hip_nrc_get_weights": (C.c_int, [ctx, C.c_void_p]),
        "btrc_hip_nrc_get_ema_weights": (C.c_int, [ctx, C.c_void_p]),
        "btrc_hip_nrc_set_weights": (C.c_int, [ctx, C.c_void_p]),
        "btrc_hip_nrc_eval": (C.c_int, [ctx, C.c_void_p, C.c_int32, C.c_void_p]),
        "btrc_hip_nrc_gradients": (C.c_int, [ctx, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, P(C.c_double)]),
        "btrc_hip_nrc_fit": (C.c_int, [ctx, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
        "btrc_hip_trace_closest": (C.c_int, [ctx, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p]),
        "btrc_hip_trace_any": (C.c_int, [ctx, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p]),
        "btrc_hip_trace_closest_noalloc": (C.c_int, [ctx, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p]),
        "btrc_hip_batch_upload": (C.c_int, [ctx, C.c_void_p, C.c_int64]),
        "btrc_hip_batch_trace": (C.c_int, [ctx, C.c_int32, C.c_int32, C.c_int32, P(C.c_double)]),
        "btrc_hip_batch_download": (C.c_int, [ctx, C.c_void_p]),
        "btrc_hip_measure_copy": (C.c_int, [ctx, C.c_int64, C.c_int32, P(C.c_double)]),
        "btrc_hip_capture_rays": (C.c_int, [ctx, P(RenderParams), C.c_int32, C.c_void_p, C.c_int64, P(C.c_int64)]),
        "btrc_hip_eval_halton": (C.c_int, [ctx, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_int32, C.c_int32, C.c_void_p]),
        "btrc_hip_eval_independent": (C.c_int, [ctx, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                                C.c_void_p]),
        "btrc_hip_eval_volumes": (C.c_int, [ctx, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
        "btrc_hip_eval_texture": (C.c_int, [ctx, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p]),
        "btrc_hip_eval_envir": (C.c_int, [ctx, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
        "btrc_hip_envir_table": (C.c_int, [ctx, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
        "btrc_hip_eval_medium": (C.c_int, [ctx, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
        "btrc_hip_eval_detmath": (C.c_int, [ctx, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
