"""MI355X (gfx950) wavefront path tracer with Btrc's render-path interface.

The compute lives in the native library libbtrc_hip.so (HIP kernels for
gfx950 behind the C ABI of include/btrc_hip.h). This module loads it with
ctypes and mirrors the reference's renderer surface for Python callers:

    WavefrontPathTracer  - the "wfpt" renderer (builtin/renderer/wavefront.h:9-60),
                           same parameters and defaults (wavefront.h:13-27)
    Scene, Camera        - the scene factory's output (factory/scene.cpp:5-94)
                           and PinholeCamera (builtin/camera/pinhole.cpp)

Errors surface as BtrcException, as in the reference (utils/exception.h).
There is no fallback: without the built library every entry point raises.
"""
import ctypes as C
import os
import threading

import numpy as np

from . import _abi
from .scenes import Camera, Scene  # noqa: F401

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libbtrc_hip.so")
_lib = None
_lib_lock = threading.Lock()


class BtrcException(RuntimeError):
    """Failure reported by the native library (status code + message)."""

    def __init__(self, status, message):
        super().__init__(f"btrc_hip error {status}: {message}")
        self.status = status


def library_path():
    return _LIB_PATH


def load_library():
    """Load libbtrc_hip.so; raises if it has not been built."""
    global _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(_LIB_PATH):
                raise BtrcException(_abi.BTRC_ERR_STATE,
                                    f"{_LIB_PATH} is missing: build it with __graft_entry__.build()")
            _lib = _abi.declare(C.CDLL(_LIB_PATH))
        return _lib


class Context:
    """One device context: a HIP stream, device memory and the uploaded scene."""

    def __init__(self, device=0):
        self.lib = load_library()
        h = C.c_void_p()
        st = self.lib.btrc_hip_create(device, C.byref(h))
        if st != _abi.BTRC_OK:
            raise BtrcException(st, f"cannot create a context on device {device}")
        self.h = h
        self._keep = None

    def close(self):
        if self.h:
            self.lib.btrc_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != _abi.BTRC_OK:
            raise BtrcException(st, self.lib.btrc_hip_last_error(self.h).decode(errors="replace"))

    def device_name(self):
        buf = C.create_string_buffer(128)
        self._check(self.lib.btrc_hip_device_name(self.h, buf, 128))
        return buf.value.decode()

    def get_tuning(self):
        t = _abi.Tuning()
        self._check(self.lib.btrc_hip_get_tuning(self.h, C.byref(t)))
        return dict(closest_postpone_leaves=t.closest_postpone_leaves, any_postpone_leaves=t.any_postpone_leaves,
                    trace_waves_per_cu=t.trace_waves_per_cu, launch_graph=t.launch_graph,
                    two_level_postpone=t.two_level_postpone, stage_blocks_per_cu=t.stage_blocks_per_cu,
                    graph_bounces=t.graph_bounces)

    def set_tuning(self, **kw):
        cur = self.get_tuning()
        cur.update(kw)
        t = _abi.Tuning()
        for k, v in cur.items():
            setattr(t, k, int(v))
        self._check(self.lib.btrc_hip_set_tuning(self.h, C.byref(t)))

    def upload_scene(self, scene):
        desc, holder = scene.to_desc()
        self._check(self.lib.btrc_hip_upload_scene(self.h, C.byref(desc)))
        self._keep = holder
        self._box = scene.world_box()

    def scene_info(self):
        n, d, s, t = C.c_int64(), C.c_int32(), C.c_int32(), C.c_int64()
        self._check(self.lib.btrc_hip_scene_info(self.h, C.byref(n), C.byref(d), C.byref(s), C.byref(t)))
        return dict(nodes=n.value, depth=d.value, stack=s.value, triangles=t.value)

    def set_camera(self, camera, width, height):
        self._check(self.lib.btrc_hip_set_camera(self.h, C.byref(camera.to_desc(width, height))))

    def render(self, params):
        stats = _abi.RenderStats()
        self._check(self.lib.btrc_hip_render(self.h, C.byref(params), C.byref(stats)))
        return stats

    def render_views(self, params, views):
        """views: list of (Camera, width, height); one wavefront over all"""
        vd = (_abi.ViewDesc * len(views))()
        for k, (cam, w, h) in enumerate(views):
            vd[k].camera = cam.to_desc(w, h)
            vd[k].width, vd[k].height = w, h
        stats = _abi.RenderStats()
        self._check(self.lib.btrc_hip_render_views(self.h, C.byref(params), len(views), vd, C.byref(stats)))
        return stats

    def view_film(self, view, width, height):
        rgba = np.empty((height, width, 4), dtype=np.float32)
        self._check(self.lib.btrc_hip_view_readback(self.h, view, rgba.ctypes.data))
        return rgba

    def film(self, width, height, fp64=False):
        rgba = np.empty((height, width, 4), dtype=np.float32)
        rgb64 = np.empty((height, width, 3), dtype=np.float64) if fp64 else None
        self._check(self.lib.btrc_hip_film_readback(
            self.h, rgba.ctypes.data, rgb64.ctypes.data if fp64 else None))
        return rgba, rgb64

    def set_reporter(self, reporter):
        """progress callbacks of later renders (Reporter, core/reporter.h):
        `reporter` has new_stage(name), complete_stage() and
        progress(percentage) methods (any may be missing); None removes it.
        The library calls them from the rendering thread during a render."""
        if reporter is None:
            self._reporter = None
            self._check(self.lib.btrc_hip_set_reporter(self.h, None))
            return

        def method(name, cb_type, wrap):
            f = getattr(reporter, name, None)
            return cb_type(wrap(f)) if f is not None else cb_type()
        r = _abi.Reporter()
        r.user = None
        r.new_stage = method("new_stage", _abi.REPORTER_STAGE, lambda f: lambda _u, name: f(name.decode()))
        r.complete_stage = method("complete_stage", _abi.REPORTER_DONE, lambda f: lambda _u: f())
        r.progress = method("progress", _abi.REPORTER_PROGRESS, lambda f: lambda _u, pct: f(pct))
        self._reporter = r   # the callbacks must outlive the library's copy
        self._check(self.lib.btrc_hip_set_reporter(self.h, C.byref(r)))

    def tonemapped(self, width, height, exposure=1.0):
        """the resolved colour film through the ACES tone map, on the device
        (ACESToneMap, postprocess/tonemap.cpp)"""
        rgba = np.empty((height, width, 4), dtype=np.float32)
        self._check(self.lib.btrc_hip_tonemap_readback(self.h, float(exposure), rgba.ctypes.data))
        return rgba

    def aov(self, width, height, albedo=True, normal=True):
        """first-hit albedo and shading-normal images (RGBA), or None"""
        a = np.empty((height, width, 4), dtype=np.float32) if albedo else None
        n = np.empty((height, width, 4), dtype=np.float32) if normal else None
        self._check(self.lib.btrc_hip_aov_readback(self.h, a.ctypes.data if albedo else None,
                                                   n.ctypes.data if normal else None))
        return a, n

    def set_filter(self, kind="box", radius=1.3, alpha=1.6):
        f = _abi.FilterDesc()
        if kind == "box":
            f.type = _abi.FILTER_BOX
        elif kind == "gaussian":
            f.type = _abi.FILTER_GAUSSIAN
        else:
            raise BtrcException(_abi.BTRC_ERR_INVALID, f"unknown film filter {kind!r}")
        f.radius, f.alpha = float(radius), float(alpha)
        self._check(self.lib.btrc_hip_set_filter(self.h, C.byref(f)))

    def request_stop(self):
        self._check(self.lib.btrc_hip_request_stop(self.h))

    # ---- neural radiance cache
    NRC_DEFAULTS = dict(enabled=0, train=1, query_depth=2, train_stride=16, steps=4, learning_rate=1e-2,
                        loss_scale=128.0, train_depth=0, ema_decay=0.0)

    def nrc_configure(self, **kw):
        unknown = set(kw) - set(self.NRC_DEFAULTS)
        if unknown:
            raise BtrcException(_abi.BTRC_ERR_INVALID, f"unknown cache parameters: {sorted(unknown)}")
        cfg = dict(self.NRC_DEFAULTS)
        cfg.update(getattr(self, "_nrc_cfg", {}))
        cfg.update(kw)
        p = _abi.NrcParams()
        for k, v in cfg.items():
            setattr(p, k, v)
        self._check(self.lib.btrc_hip_nrc_configure(self.h, C.byref(p)))
        self._nrc_cfg = cfg

    def nrc_reset(self, seed=1):
        self._check(self.lib.btrc_hip_nrc_reset(self.h, seed))

    def nrc_stats(self):
        st = _abi.NrcStats()
        self._check(self.lib.btrc_hip_nrc_stats(self.h, C.byref(st)))
        return st

    def nrc_weights(self):
        w = np.empty(_abi.NRC_PARAMS, dtype=np.float32)
        self._check(self.lib.btrc_hip_nrc_get_weights(self.h, w.ctypes.data))
        return w

    def nrc_ema_weights(self):
        w = np.empty(_abi.NRC_PARAMS, dtype=np.float32)
        self._check(self.lib.btrc_hip_nrc_get_ema_weights(self.h, w.ctypes.data))
        return w

    def nrc_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == _abi.NRC_PARAMS
        self._check(self.lib.btrc_hip_nrc_set_weights(self.h, w.ctypes.data))

    def nrc_eval(self, vertices):
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, _abi.NRC_VERTEX_FLOATS)
        out = np.empty((len(v), 3), dtype=np.float32)
        self._check(self.lib.btrc_hip_nrc_eval(self.h, v.ctypes.data, len(v), out.ctypes.data))
        return out

    def nrc_gradients(self, vertices, targets):
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, _abi.NRC_VERTEX_FLOATS)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, 3)
        g = np.empty(_abi.NRC_PARAMS, dtype=np.float32)
        loss = C.c_double()
        self._check(self.lib.btrc_hip_nrc_gradients(self.h, v.ctypes.data, t.ctypes.data, len(v), g.ctypes.data,
                                                    C.byref(loss)))
        return g, loss.value

    def nrc_fit(self, vertices, targets, steps):
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, _abi.NRC_VERTEX_FLOATS)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, 3)
        loss = np.empty(max(1, steps), dtype=np.float64)
        self._check(self.lib.btrc_hip_nrc_fit(self.h, v.ctypes.data, t.ctypes.data, len(v), steps, loss.ctypes.data))
        return loss[:steps]

    def scene_box(self):
        """(lo, 1 / extent) of the scene's triangles: the cache's position frame"""
        return self._box

    # ---- traversal on ray batches
    def trace_closest(self, rays, tri_test=_abi.TRI_WATERTIGHT):
        rays = np.ascontiguousarray(rays, dtype=_abi.ray_dtype())
        hits = np.empty(len(rays), dtype=_abi.hit_dtype())
        self._check(self.lib.btrc_hip_trace_closest(self.h, rays.ctypes.data, len(rays), tri_test,
                                                    hits.ctypes.data))
        return hits

    def trace_closest_noalloc(self, rays_ptr, n, hits_ptr, tri_test=_abi.TRI_WATERTIGHT):
        """closest hits of n rays already in device memory (n btrc_ray at
        device address rays_ptr) into device memory (n btrc_hit at hits_ptr);
        allocates nothing"""
        self._check(self.lib.btrc_hip_trace_closest_noalloc(self.h, rays_ptr, n, tri_test, hits_ptr))

    def trace_any(self, rays, tri_test=_abi.TRI_WATERTIGHT):
        rays = np.ascontiguousarray(rays, dtype=_abi.ray_dtype())
        occ = np.empty(len(rays), dtype=np.uint8)
        self._check(self.lib.btrc_hip_trace_any(self.h, rays.ctypes.data, len(rays), tri_test, occ.ctypes.data))
        return occ.astype(bool)

    def batch_upload(self, rays):
        rays = np.ascontiguousarray(rays, dtype=_abi.ray_dtype())
        self._check(self.lib.btrc_hip_batch_upload(self.h, rays.ctypes.data, len(rays)))
        self._batch_n = len(rays)

    def batch_trace(self, tri_test=_abi.TRI_WATERTIGHT, any_hit=False, repeats=1):
        ms = C.c_double()
        self._check(self.lib.btrc_hip_batch_trace(self.h, tri_test, 1 if any_hit else 0, repeats, C.byref(ms)))
        return ms.value

    def batch_download(self):
        hits = np.empty(self._batch_n, dtype=_abi.hit_dtype())
        self._check(self.lib.btrc_hip_batch_download(self.h, hits.ctypes.data))
        return hits

    def measure_copy(self, mib=4096, repeats=8):
        """the device's copy bandwidth in GB/s (bytes read + written), from a
        STREAM-style copy kernel over `mib` MiB, median of `repeats` copies"""
        gbs = C.c_double()
        self._check(self.lib.btrc_hip_measure_copy(self.h, int(mib) << 20, repeats, C.byref(gbs)))
        return gbs.value

    def capture_rays(self, params, bounces, max_rays):
        rays = np.empty(max_rays, dtype=_abi.ray_dtype())
        n = C.c_int64()
        self._check(self.lib.btrc_hip_capture_rays(self.h, C.byref(params), bounces, rays.ctypes.data,
                                                   max_rays, C.byref(n)))
        return rays[:n.value]

    # ---- device building blocks, for parity tests
    def eval_halton(self, width, height, px, py, sample_index, dims):
        px = np.ascontiguousarray(px, dtype=np.uint32)
        py = np.ascontiguousarray(py, dtype=np.uint32)
        si = np.ascontiguousarray(sample_index, dtype=np.int32)
        out = np.empty((len(px), dims), dtype=np.float32)
        self._check(self.lib.btrc_hip_eval_halton(self.h, width, height, px.ctypes.data, py.ctypes.data,
                                                  si.ctypes.data, len(px), dims, out.ctypes.data))
        return out

    def eval_independent(self, px, py, sample_index, dims):
        px = np.ascontiguousarray(px, dtype=np.uint32)
        py = np.ascontiguousarray(py, dtype=np.uint32)
        si = np.ascontiguousarray(sample_index, dtype=np.int32)
        out = np.empty((len(px), dims), dtype=np.float32)
        self._check(self.lib.btrc_hip_eval_independent(self.h, px.ctypes.data, py.ctypes.data, si.ctypes.data,
                                                       len(px), dims, out.ctypes.data))
        return out

    def eval_volumes(self, segments, seeds):
        """(n, 10) float32 of the volume-primitive medium: scattered, position, albedo, transmittance"""
        seg = np.ascontiguousarray(segments, dtype=_abi.ray_dtype())
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        out = np.empty((len(seg), 10), dtype=np.float32)
        self._check(self.lib.btrc_hip_eval_volumes(self.h, seg.ctypes.data, seeds.ctypes.data, len(seg),
                                                   out.ctypes.data))
        return out

    def eval_texture(self, texture, uv):
        """(n, 4) float32: texture `texture` of the current scene at (n, 2) uv"""
        uv = np.ascontiguousarray(uv, dtype=np.float32).reshape(-1, 2)
        out = np.empty((len(uv), 4), dtype=np.float32)
        self._check(self.lib.btrc_hip_eval_texture(self.h, texture, uv.ctypes.data, len(uv), out.ctypes.data))
        return out

    def scene_build_ms(self):
        """wall time of the last upload's acceleration-structure build"""
        ms = C.c_double(0)
        self._check(self.lib.btrc_hip_scene_build_ms(self.h, C.byref(ms)))
        return ms.value

    def update_geometry(self, scene, geometry):
        """re-send geometry `geometry` of `scene` (its arrays changed in
        place, same triangle count) to the uploaded one-level scene: records
        rebuilt, tree refitted; returns the update's wall time in ms"""
        g = scene.geometries[geometry]
        fp = C.POINTER(C.c_float)
        d = _abi.GeometryDesc()
        d.triangle_count = len(g["positions"])
        d.positions = g["positions"].ctypes.data_as(fp)
        d.normals = g["normals"].ctypes.data_as(fp) if g["normals"] is not None else None
        d.tex_coords = g["uvs"].ctypes.data_as(fp) if g["uvs"] is not None else None
        ms = C.c_double(0)
        self._check(self.lib.btrc_hip_update_geometry(self.h, geometry, C.byref(d), C.byref(ms)))
        return ms.value

    def build_bvh(self, boxes, max_leaf_size=0, refit_boxes=None):
        """GPU binary BVH over (n, 6) boxes (lo xyz, hi xyz): (nodes as
        _abi.bvh_node_dtype(), primitive indices, build ms)"""
        boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
        n = len(boxes)
        nodes = np.zeros(max(1, 2 * n - 1), dtype=_abi.bvh_node_dtype())
        prims = np.zeros(max(1, n), np.uint32)
        rb = None if refit_boxes is None else np.ascontiguousarray(refit_boxes, dtype=np.float32).reshape(-1, 6)
        if rb is not None and rb.shape != boxes.shape:
            raise BtrcException("build_bvh: refit boxes differ in shape")
        count = C.c_int32(0)
        ms = C.c_double(0)
        self._check(self.lib.btrc_hip_build_bvh(self.h, boxes.ctypes.data, n, max_leaf_size,
                                                None if rb is None else rb.ctypes.data, nodes.ctypes.data,
                                                prims.ctypes.data, C.byref(count), C.byref(ms)))
        return nodes[:count.value], prims[:n], ms.value

    def eval_envir(self, dirs, samples):
        """(n, 12) float32: the environment light's eval_le rgb, pdf_li at
        world directions `dirs` (n, 3), then sample_li's direction, pdf and
        radiance for the sample triples `samples` (n, 3), then 0"""
        d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        sm = np.ascontiguousarray(samples, dtype=np.float32).reshape(-1, 3)
        if len(d) != len(sm):
            raise BtrcException("eval_envir: dirs and samples differ in length")
        out = np.empty((len(d), 12), dtype=np.float32)
        self._check(self.lib.btrc_hip_eval_envir(self.h, d.ctypes.data, sm.ctypes.data, len(d), out.ctypes.data))
        return out

    def envir_table(self):
        """the IBL importance table's tile probabilities (empty without an IBL)"""
        n = C.c_int32(0)
        self._check(self.lib.btrc_hip_envir_table(self.h, None, 0, C.byref(n)))
        out = np.empty(n.value, dtype=np.float32)
        self._check(self.lib.btrc_hip_envir_table(self.h, out.ctypes.data, n.value, C.byref(n)))
        return out

    def eval_medium(self, medium, segments, seeds):
        """(n, 10) float32: scattered, position, albedo, transmittance"""
        seg = np.ascontiguousarray(segments, dtype=_abi.ray_dtype())
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        out = np.empty((len(seg), 10), dtype=np.float32)
        self._check(self.lib.btrc_hip_eval_medium(self.h, medium, seg.ctypes.data, seeds.ctypes.data, len(seg),
                                                  out.ctypes.data))
        return out

    def eval_detmath(self, fn, x, y=None):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(np.ones_like(x) if y is None else y, dtype=np.float32)
        out = np.empty_like(x)
        self._check(self.lib.btrc_hip_eval_detmath(self.h, fn, x.ctypes.data, y.ctypes.data, len(x),
                                                   out.ctypes.data))
        return out


class RenderResult:
    """Renderer::RenderResult (core/renderer.h:19-24): the resolved colour,
    albedo and normal images as host memory (albedo / normal None unless
    requested), plus the run's statistics."""

    def __init__(self, color, color64, stats, albedo=None, normal=None):
        self.color = color
        self.color64 = color64
        self.stats = stats
        self.albedo = albedo
        self.normal = normal


class WavefrontPathTracer:
    """The "wfpt" renderer. Parameters and defaults follow
    WavefrontPathTracer::Params (builtin/renderer/wavefront.h:13-27) and its
    creator (wavefront.cpp:375-398)."""

    DEFAULTS = dict(tile=False, spp=128, min_depth=5, max_depth=10, rr_threshold=0.1, rr_cont_prob=0.5,
                    state_count=1000000, albedo=False, normal=False)
    SAMPLERS = {"halton": _abi.SAMPLER_HALTON, "independent": _abi.SAMPLER_INDEPENDENT}

    def __init__(self, context=None, **params):
        unknown = set(params) - set(self.DEFAULTS) - {"film_fp64", "time_stages", "sampler"}
        if unknown:
            raise BtrcException(_abi.BTRC_ERR_INVALID, f"unknown wfpt parameters: {sorted(unknown)}")
        self.ctx = context or Context()
        self.params = dict(self.DEFAULTS)
        self.params.update(params)
        self.scene = None
        self.camera = None
        if self.params.get("sampler", "halton") not in self.SAMPLERS:
            raise BtrcException(_abi.BTRC_ERR_INVALID, f"unknown sampler {self.params['sampler']!r}")
        self.width, self.height = 512, 512
        self.filter = ("box", 1.3, 1.6)
        self._committed = False

    def set_scene(self, scene):
        self.scene = scene
        self._committed = False

    def set_camera(self, camera):
        self.camera = camera
        self._committed = False

    def set_film_filter(self, kind="box", radius=1.3, alpha=1.6):
        """WavefrontPathTracer::set_film_filter (wavefront.cpp:72-75): "box"
        or "gaussian" (GaussianFilterCreator defaults, gaussian.cpp:68-76)"""
        self.filter = (kind, radius, alpha)
        self._committed = False

    def set_film(self, width, height):
        self.width, self.height = int(width), int(height)
        self._committed = False

    def set_reporter(self, reporter):
        """Renderer::set_reporter (core/renderer.h): an object with
        new_stage / complete_stage / progress methods, or None"""
        self.ctx.set_reporter(reporter)

    def commit(self):
        if self.scene is None or self.camera is None:
            raise BtrcException(_abi.BTRC_ERR_STATE, "commit needs a scene and a camera")
        self.ctx.upload_scene(self.scene)
        self.ctx.set_camera(self.camera, self.width, self.height)
        self.ctx.set_filter(*self.filter)
        self._committed = True

    def render_params(self):
        p = _abi.RenderParams()
        p.width, p.height = self.width, self.height
        p.spp = int(self.params["spp"])
        p.min_depth = int(self.params["min_depth"])
        p.max_depth = int(self.params["max_depth"])
        p.rr_threshold = float(self.params["rr_threshold"])
        p.rr_cont_prob = float(self.params["rr_cont_prob"])
        p.state_count = int(self.params["state_count"])
        p.film_fp64 = 1 if self.params.get("film_fp64") else 0
        p.time_stages = 1 if self.params.get("time_stages") else 0
        p.tile = 1 if self.params["tile"] else 0
        p.sampler = self.SAMPLERS[self.params.get("sampler", "halton")]
        p.albedo = 1 if self.params["albedo"] else 0
        p.normal = 1 if self.params["normal"] else 0
        return p

    def render(self):
        if not self._committed:
            self.commit()
        stats = self.ctx.render(self.render_params())
        color, color64 = self.ctx.film(self.width, self.height, fp64=bool(self.params.get("film_fp64")))
        albedo = normal = None
        if self.params["albedo"] or self.params["normal"]:
            albedo, normal = self.ctx.aov(self.width, self.height, albedo=bool(self.params["albedo"]),
                                          normal=bool(self.params["normal"]))
        return RenderResult(color, color64, stats, albedo, normal)

    def stop_async(self):
        """Renderer::stop_async (core/renderer.cpp:38-48)"""
        self.ctx.request_stop()
