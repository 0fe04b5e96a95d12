"""btrc_hip_trace_closest_noalloc: the closest-hit batch traversal on device
buffers with the library's static scratch instead of allocations. Its hits
must equal btrc_hip_trace_closest's bit for bit, on two calls in a row (the
second smaller, over the scratch the first left behind)."""
import ctypes as C

import numpy as np
import pytest

import btrc_amd
from btrc_amd import _abi, scenes

from util import assert_hits_equal, random_rays

H2D, D2H = 1, 2


class DeviceBuffer:
    """device memory from the HIP runtime the library itself links
    (libamdhip64.so.7), so that no second runtime enters the process"""

    def __init__(self, nbytes):
        btrc_amd.load_library()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), nbytes) == 0
        self.nbytes = nbytes

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


@pytest.mark.gpu
def test_noalloc_trace_matches_allocating_entry(ctx):
    scene, _ = scenes.sphere_field(instances=512, lights=8)
    ctx.upload_scene(scene)
    info = ctx.scene_info()
    assert info["stack"] > 16   # the overflow part of the stack is exercised
    box = scene.world_box()
    lo, hi = box[:3], box[:3] + 1.0 / box[3:]
    for n, seed in ((200_000, 11), (37_001, 12)):
        rays = random_rays(n, lo, hi, seed)
        want = ctx.trace_closest(rays)
        d_rays, d_hits = DeviceBuffer(rays.nbytes), DeviceBuffer(n * 20)
        d_rays.upload(rays)
        d_hits.upload(np.full(n * 20, 0xAB, np.uint8))
        ctx.trace_closest_noalloc(d_rays.ptr.value, n, d_hits.ptr.value)
        got = d_hits.download(_abi.hit_dtype(), n)
        d_rays.free()
        d_hits.free()
        assert_hits_equal(got, want)
        assert (got["instance"] != 0xffffffff).sum() > n // 10


@pytest.mark.gpu
def test_noalloc_trace_rejects_bad_buffers(ctx):
    ctx.upload_scene(scenes.cornell()[0])
    rays = random_rays(1000, -1.0, 1.0, 3)
    d_rays, d_small = DeviceBuffer(rays.nbytes), DeviceBuffer(999 * 20)
    d_rays.upload(rays)
    try:
        with pytest.raises(btrc_amd.BtrcException, match="smaller than n"):
            ctx.trace_closest_noalloc(d_rays.ptr.value, 1000, d_small.ptr.value)
        host_hits = np.empty(1000, dtype=_abi.hit_dtype())
        with pytest.raises(btrc_amd.BtrcException, match="not device memory"):
            ctx.trace_closest_noalloc(d_rays.ptr.value, 1000, host_hits.ctypes.data)
    finally:
        d_rays.free()
        d_small.free()
