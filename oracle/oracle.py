"""Loader of the CPU oracle (oracle/oracle.cpp -> oracle/liboracle.so) and of
the reference's own CPU traversal built from lib/bvh (oracle/_ref).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os

import numpy as np

from btrc_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libbvhref.so")

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: build it with __graft_entry__.build()")
        L = C.CDLL(ORACLE_SO)
        vp, P = C.c_void_p, C.POINTER
        L.oracle_scene_create.restype = vp
        L.oracle_scene_create.argtypes = [P(_abi.SceneDesc)]
        L.oracle_scene_destroy.argtypes = [vp]
        L.oracle_scene_bvh_nodes.restype = C.c_int64
        L.oracle_scene_bvh_nodes.argtypes = [vp]
        L.oracle_trace.argtypes = [vp, vp, C.c_int64, C.c_int32, C.c_int32, C.c_int32, vp, vp, C.c_int32]
        L.oracle_render.argtypes = [vp, P(_abi.CameraDesc), P(_abi.RenderParams), P(_abi.FilterDesc), C.c_uint64,
                                    C.c_uint64, vp, vp, C.c_int32, P(C.c_uint64), P(C.c_uint64)]
        L.oracle_independent.argtypes = [vp, vp, vp, C.c_int32, C.c_int32, vp]
        L.oracle_gaussian_filter.argtypes = [C.c_float, C.c_float, vp, C.c_int32, vp]
        L.oracle_medium_eval.argtypes = [vp, C.c_int32, vp, vp, C.c_int32, vp]
        L.oracle_texture_eval.argtypes = [vp, C.c_int32, vp, C.c_int32, vp]
        L.oracle_volume_eval.argtypes = [vp, vp, vp, C.c_int32, vp]
        L.oracle_envir_eval.argtypes = [vp, vp, vp, C.c_int32, vp]
        L.oracle_envir_table.restype = C.c_int32
        L.oracle_envir_table.argtypes = [vp, vp, C.c_int32]
        L.oracle_hg.argtypes = [C.c_float, vp, C.c_int32, vp]
        L.oracle_tex3.restype = C.c_float
        L.oracle_tex3.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_float]
        L.oracle_nrc_forward.argtypes = [vp, vp, vp, C.c_int32, C.c_int32, vp]
        L.oracle_nrc_gradients.argtypes = [vp, vp, vp, vp, C.c_int32, C.c_int32, C.c_float, vp, P(C.c_double)]
        L.oracle_murmur64a.restype = C.c_uint64
        L.oracle_murmur64a.argtypes = [vp, C.c_uint64, C.c_uint64]
        L.oracle_pcg32.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int32, vp]
        L.oracle_halton.argtypes = [C.c_int32, C.c_int32, vp, vp, vp, C.c_int32, C.c_int32, vp]
        L.oracle_detmath.argtypes = [C.c_int32, vp, vp, C.c_int32, vp]
        L.oracle_bsdf_sample.argtypes = [P(_abi.MaterialDesc), vp, vp, C.c_int32, vp, vp, vp]
        L.oracle_bsdf_eval.argtypes = [P(_abi.MaterialDesc), vp, vp, C.c_int32, vp, vp]
        L.oracle_kat_rays.argtypes = [C.c_int32, C.c_int32, vp, vp, vp, C.c_float, vp]
        L.oracle_kat_pixels.argtypes = [C.c_int32, C.c_int32, vp, vp, vp]
        L.oracle_aces.argtypes = [vp, C.c_int32, C.c_float, vp]
        L.oracle_tracking_stats.argtypes = [P(C.c_uint64), P(C.c_uint64)]
        L.oracle_halton_index.restype = C.c_uint64
        L.oracle_halton_index.argtypes = [C.c_int32, C.c_int32, C.c_uint32, C.c_uint32, C.c_int32]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    """lib/bvh (the reference's CPU traversal) compiled by oracle/build_ref.sh"""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"{REF_SO} missing: run oracle/build_ref.sh where the reference is present")
        L = C.CDLL(REF_SO)
        vp = C.c_void_p
        L.bvhref_create.restype = vp
        L.bvhref_create.argtypes = [vp, C.c_int64, C.c_int32]
        L.bvhref_destroy.argtypes = [vp]
        L.bvhref_trace.argtypes = [vp, vp, C.c_int64, C.c_int32, vp, vp, vp, vp, vp, vp, C.c_int32]
        L.bvhref_ploc.restype = C.c_int64
        L.bvhref_ploc.argtypes = [vp, C.c_int64, C.c_int32, vp, vp, vp]
        _ref = L
    return _ref


class OracleScene:
    def __init__(self, scene):
        self.desc, self.holder = scene.to_desc()
        self.h = lib().oracle_scene_create(C.byref(self.desc))

    def __del__(self):
        try:
            lib().oracle_scene_destroy(self.h)
        except Exception:
            pass

    def volume_eval(self, segments, seeds):
        """(n, 10) float32 of the volume-primitive medium: scattered, position, albedo, transmittance"""
        seg = np.ascontiguousarray(segments, dtype=_abi.ray_dtype())
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        out = np.empty((len(seg), 10), dtype=np.float32)
        lib().oracle_volume_eval(self.h, seg.ctypes.data, seeds.ctypes.data, len(seg), out.ctypes.data)
        return out

    def texture_eval(self, texture, uv):
        """(n, 4) float32: the scene's texture `texture` at (n, 2) uv"""
        uv = np.ascontiguousarray(uv, dtype=np.float32).reshape(-1, 2)
        out = np.empty((len(uv), 4), dtype=np.float32)
        lib().oracle_texture_eval(self.h, texture, uv.ctypes.data, len(uv), out.ctypes.data)
        return out

    def envir_eval(self, dirs, samples):
        """(n, 12) float32 like Context.eval_envir"""
        d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        sm = np.ascontiguousarray(samples, dtype=np.float32).reshape(-1, 3)
        out = np.empty((len(d), 12), dtype=np.float32)
        lib().oracle_envir_eval(self.h, d.ctypes.data, sm.ctypes.data, len(d), out.ctypes.data)
        return out

    def envir_table(self):
        """the IBL's tile probabilities"""
        n = lib().oracle_envir_table(self.h, None, 0)
        out = np.empty(n, dtype=np.float32)
        lib().oracle_envir_table(self.h, out.ctypes.data, n)
        return out

    def medium_eval(self, medium, segments, seeds):
        """(n, 10) float32 like Context.eval_medium"""
        seg = np.ascontiguousarray(segments, dtype=_abi.ray_dtype())
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        out = np.empty((len(seg), 10), dtype=np.float32)
        lib().oracle_medium_eval(self.h, medium, seg.ctypes.data, seeds.ctypes.data, len(seg), out.ctypes.data)
        return out

    def trace(self, rays, tri_test=_abi.TRI_WATERTIGHT, any_hit=False, use_bvh=True, threads=8):
        rays = np.ascontiguousarray(rays, dtype=_abi.ray_dtype())
        n = len(rays)
        if any_hit:
            occ = np.empty(n, dtype=np.uint8)
            lib().oracle_trace(self.h, rays.ctypes.data, n, tri_test, 1, 1 if use_bvh else 0, None,
                               occ.ctypes.data, threads)
            return occ.astype(bool)
        hits = np.empty(n, dtype=_abi.hit_dtype())
        lib().oracle_trace(self.h, rays.ctypes.data, n, tri_test, 0, 1 if use_bvh else 0, hits.ctypes.data,
                           None, threads)
        return hits

    def render(self, camera, params, first=0, count=None, threads=8, filt=None, aov=False):
        """Radiance sums (H, W, 3) float64 of paths [first, first + count) and
        the (extension, shadow) ray counts; with aov=True also the first-hit
        (albedo, normal) sums, each (H, W, 3). filt: None (box) or
        ("gaussian", radius, alpha)."""
        w, h = params.width, params.height
        if count is None:
            count = w * h * params.spp
        film = np.zeros((h, w, 3), dtype=np.float64)
        aovs = np.zeros((h, w, 6), dtype=np.float64) if aov else None
        cam = camera.to_desc(w, h)
        fd = None
        if filt is not None and filt[0] == "gaussian":
            fd = _abi.FilterDesc(_abi.FILTER_GAUSSIAN, float(filt[1]), float(filt[2]))
        e, s = C.c_uint64(), C.c_uint64()
        lib().oracle_render(self.h, C.byref(cam), C.byref(params), C.byref(fd) if fd is not None else None, first,
                            count, film.ctypes.data, aovs.ctypes.data if aov else None, threads,
                            C.byref(e), C.byref(s))
        if aov:
            return film, e.value, s.value, aovs[..., :3], aovs[..., 3:]
        return film, e.value, s.value


def hg_sample(g, u):
    """Henyey-Greenstein samples for wo = +z: (n, 5) dir xyz, pdf, pdf
    re-evaluated from the direction"""
    u = np.ascontiguousarray(u, dtype=np.float32)
    out = np.empty((len(u), 5), dtype=np.float32)
    lib().oracle_hg(g, u.ctypes.data, len(u), out.ctypes.data)
    return out


def tex3(grid, uvw):
    """linear lookup of a [z, y, x] grid at normalised coordinates"""
    grid = np.ascontiguousarray(grid, dtype=np.float32)
    nz, ny, nx = grid.shape
    return np.array([lib().oracle_tex3(grid.ctypes.data, nx, ny, nz, *map(float, p)) for p in uvw], dtype=np.float32)


def independent(px, py, sample_index, dims):
    """first `dims` numbers of the independent sampler per (px, py, sample)"""
    px = np.ascontiguousarray(px, dtype=np.uint32)
    py = np.ascontiguousarray(py, dtype=np.uint32)
    si = np.ascontiguousarray(sample_index, dtype=np.int32)
    out = np.empty((len(px), dims), dtype=np.float32)
    lib().oracle_independent(px.ctypes.data, py.ctypes.data, si.ctypes.data, len(px), dims, out.ctypes.data)
    return out


def gaussian_filter(radius, alpha, u):
    """GaussianFilter::sample offsets (n, 2) for uniforms u (n, 3)"""
    u = np.ascontiguousarray(u, dtype=np.float32)
    out = np.empty((len(u), 2), dtype=np.float32)
    lib().oracle_gaussian_filter(radius, alpha, u.ctypes.data, len(u), out.ctypes.data)
    return out


NRC_PARAMS = 22528


def nrc_forward(weights, box, vertices, bf16=True):
    """cache network outputs (n, 3) float64; box = (lo xyz, 1 / extent xyz)"""
    w = np.ascontiguousarray(weights, dtype=np.float32)
    b = np.ascontiguousarray(box, dtype=np.float32)
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 16)
    out = np.empty((len(v), 3), dtype=np.float64)
    lib().oracle_nrc_forward(w.ctypes.data, b.ctypes.data, v.ctypes.data, len(v), 1 if bf16 else 0, out.ctypes.data)
    return out


def nrc_gradients(weights, box, vertices, targets, bf16=True, loss_scale=128.0):
    """(gradient (22528,) float64, loss) of the mean relative L2 loss"""
    w = np.ascontiguousarray(weights, dtype=np.float32)
    b = np.ascontiguousarray(box, dtype=np.float32)
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 16)
    t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, 3)
    g = np.empty(NRC_PARAMS, dtype=np.float64)
    loss = C.c_double()
    lib().oracle_nrc_gradients(w.ctypes.data, b.ctypes.data, v.ctypes.data, t.ctypes.data, len(v), 1 if bf16 else 0,
                               loss_scale, g.ctypes.data, C.byref(loss))
    return g, loss.value


def murmur64a(key, seed):
    key = np.frombuffer(bytes(key), dtype=np.uint8)
    buf = np.ascontiguousarray(key) if len(key) else np.zeros(1, np.uint8)
    return int(lib().oracle_murmur64a(buf.ctypes.data, len(key), seed))


def pcg32(seq, offset, n, skip=0):
    out = np.empty(n, dtype=np.uint32)
    lib().oracle_pcg32(seq, offset, skip, n, out.ctypes.data)
    return out


def halton(width, height, px, py, sample_index, dims):
    px = np.ascontiguousarray(px, dtype=np.uint32)
    py = np.ascontiguousarray(py, dtype=np.uint32)
    si = np.ascontiguousarray(sample_index, dtype=np.int32)
    out = np.empty((len(px), dims), dtype=np.float32)
    lib().oracle_halton(width, height, px.ctypes.data, py.ctypes.data, si.ctypes.data, len(px), dims,
                        out.ctypes.data)
    return out


def tracking_stats():
    """(delta-tracking segments, density lookups) of the oracle's medium
    stage since the last call"""
    a, b = C.c_uint64(), C.c_uint64()
    lib().oracle_tracking_stats(C.byref(a), C.byref(b))
    return a.value, b.value


def aces(rgba, exposure=1.0):
    """ACES tone map (Narkowicz fit) of (..., 4) float32 RGBA"""
    a = np.ascontiguousarray(rgba, dtype=np.float32)
    out = np.empty_like(a)
    lib().oracle_aces(a.ctypes.data, a.size // 4, exposure, out.ctypes.data)
    return out


def halton_index(width, height, px, py, sample_index):
    return lib().oracle_halton_index(width, height, px, py, sample_index)


def detmath(fn, x, y=None):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(np.ones_like(x) if y is None else y, dtype=np.float32)
    out = np.empty_like(x)
    lib().oracle_detmath(fn, x.ctypes.data, y.ctypes.data, len(x), out.ctypes.data)
    return out


def _material(**kw):
    m = _abi.MaterialDesc()
    m.type = kw.get("type", _abi.MATERIAL_DISNEY)
    m.shadow_terminator = kw.get("shadow_terminator", 1)
    m.color = _abi.F3(*kw.get("color", (0.5, 0.5, 0.5)))
    defaults = dict(metallic=0.0, roughness=0.5, specular=1.0, specular_tint=0.0, anisotropic=0.0, sheen=0.0,
                    sheen_tint=0.0, clearcoat=0.0, clearcoat_gloss=1.0, transmission=0.0,
                    transmission_roughness=None, ior=1.5)
    for k, v in defaults.items():
        v = kw.get(k, v)
        if k == "transmission_roughness" and v is None:
            v = kw.get("roughness", 0.5)
        setattr(m, k, float(v))
    return m


def bsdf_sample(material, wo, u):
    """BSDF samples at a surface with identity frames: (f, dir, pdf)"""
    m = _material(**material)
    wo = np.ascontiguousarray(wo, dtype=np.float32)
    u = np.ascontiguousarray(u, dtype=np.float32)
    n = len(wo)
    f = np.empty((n, 3), np.float32)
    d = np.empty((n, 3), np.float32)
    p = np.empty(n, np.float32)
    lib().oracle_bsdf_sample(C.byref(m), wo.ctypes.data, u.ctypes.data, n, f.ctypes.data, d.ctypes.data,
                             p.ctypes.data)
    return f, d, p


def bsdf_eval(material, wi, wo):
    m = _material(**material)
    wi = np.ascontiguousarray(wi, dtype=np.float32)
    wo = np.ascontiguousarray(wo, dtype=np.float32)
    n = len(wi)
    f = np.empty((n, 3), np.float32)
    p = np.empty(n, np.float32)
    lib().oracle_bsdf_eval(C.byref(m), wi.ctypes.data, wo.ctypes.data, n, f.ctypes.data, p.ctypes.data)
    return f, p


class RefBvh:
    """lib/bvh BinnedSahBuilder<16> BVH + SingleRayTraverser over flattened
    world triangles (9 floats each)."""

    def __init__(self, tris, builder=0):
        self.tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
        self.h = ref().bvhref_create(self.tris.ctypes.data, len(self.tris), builder)

    def __del__(self):
        try:
            ref().bvhref_destroy(self.h)
        except Exception:
            pass

    def trace(self, rays, any_hit=False, threads=8):
        """returns prim (uint32, 0xffffffff miss), t, u, v, per-ray traversal
        steps and primitive intersections"""
        rays = np.ascontiguousarray(rays, dtype=_abi.ray_dtype())
        n = len(rays)
        prim = np.empty(n, np.uint32)
        t = np.empty(n, np.float32)
        u = np.empty(n, np.float32)
        v = np.empty(n, np.float32)
        steps = np.empty(n, np.uint32)
        isect = np.empty(n, np.uint32)
        ref().bvhref_trace(self.h, rays.ctypes.data, n, 1 if any_hit else 0, prim.ctypes.data, t.ctypes.data,
                           u.ctypes.data, v.ctypes.data, steps.ctypes.data, isect.ctypes.data, threads)
        return prim, t, u, v, steps, isect


def ref_ploc(boxes, collapse=True, refit_boxes=None):
    """lib/bvh's LocallyOrderedClusteringBuilder (+ LeafCollapser, + a
    HierarchyRefitter onto refit_boxes) over (n, 6) boxes: (nodes as
    _abi.bvh_node_dtype(), primitive indices)"""
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    n = len(boxes)
    nodes = np.zeros(2 * n - 1, dtype=_abi.bvh_node_dtype())
    prims = np.zeros(n, np.uint32)
    rb = None if refit_boxes is None else np.ascontiguousarray(refit_boxes, dtype=np.float32).reshape(-1, 6)
    count = ref().bvhref_ploc(boxes.ctypes.data, n, 1 if collapse else 0, None if rb is None else rb.ctypes.data,
                              nodes.ctypes.data, prims.ctypes.data)
    return nodes[:count], prims


def kat_rays(width, height, eye, direction, up, fov):
    """camera rays of lib/bvh's benchmark renderer (ray index = width * row + column)"""
    out = np.empty(width * height, dtype=_abi.ray_dtype())
    e = np.asarray(eye, np.float32)
    d = np.asarray(direction, np.float32)
    u = np.asarray(up, np.float32)
    lib().oracle_kat_rays(width, height, e.ctypes.data, d.ctypes.data, u.ctypes.data, fov, out.ctypes.data)
    return out


def kat_pixels(width, height, tris, hits):
    """8-bit |normal| image (height, width, 3), top row first"""
    tris = np.ascontiguousarray(tris, dtype=np.float32)
    hits = np.ascontiguousarray(hits, dtype=_abi.hit_dtype())
    out = np.empty((height, width, 3), dtype=np.uint8)
    lib().oracle_kat_pixels(width, height, tris.ctypes.data, hits.ctypes.data, out.ctypes.data)
    return out
