"""Reporter callbacks during a render (Reporter, core/reporter.h:7-36, as
WavefrontPathTracer::render calls it, wavefront.cpp:221,287,314): one new
stage, progress rising to 100 % of the frame's paths, one completed stage,
also when the render is stopped from inside a progress callback."""
import pytest

import btrc_amd
from btrc_amd import _abi, scenes

from util import render_params

pytestmark = pytest.mark.gpu


class Recorder:
    def __init__(self, stop_ctx=None):
        self.calls = []
        self.stop_ctx = stop_ctx

    def new_stage(self, name):
        self.calls.append(("new_stage", name))

    def complete_stage(self):
        self.calls.append(("complete_stage",))

    def progress(self, pct):
        self.calls.append(("progress", pct))
        if self.stop_ctx is not None:
            self.stop_ctx.lib.btrc_hip_request_stop(self.stop_ctx.h)


@pytest.mark.parametrize("graph", [1, 0])
def test_reporter_sees_the_whole_frame(ctx, graph):
    scene, cam = scenes.cornell()
    ctx.upload_scene(scene)
    ctx.set_camera(cam, 64, 64)
    ctx.set_tuning(launch_graph=graph)
    rec = Recorder()
    ctx.set_reporter(rec)
    try:
        ctx.render(render_params(64, 64, 16, state_count=1 << 14))
    finally:
        ctx.set_reporter(None)
    assert rec.calls[0] == ("new_stage", "") and rec.calls[-1] == ("complete_stage",)
    pct = [c[1] for c in rec.calls if c[0] == "progress"]
    assert len(pct) >= 2 and len(rec.calls) == len(pct) + 2
    assert all(0.0 <= a <= b for a, b in zip(pct, pct[1:]))
    assert pct[-1] == 100.0
    # removed: later renders report nothing
    ctx.render(render_params(64, 64, 1, state_count=1 << 14))
    assert rec.calls[-1] == ("complete_stage",) and len(rec.calls) == len(pct) + 2


def test_reporter_stage_completes_when_stopped(ctx):
    scene, cam = scenes.cornell()
    ctx.upload_scene(scene)
    ctx.set_camera(cam, 64, 64)
    rec = Recorder(stop_ctx=ctx)
    ctx.set_reporter(rec)
    try:
        with pytest.raises(btrc_amd.BtrcException) as e:
            ctx.render(render_params(64, 64, 64, state_count=1 << 12))
        assert e.value.status == _abi.BTRC_ERR_STOPPED
    finally:
        ctx.set_reporter(None)
    assert rec.calls[0][0] == "new_stage" and rec.calls[-1] == ("complete_stage",)
    assert sum(1 for c in rec.calls if c[0] == "progress") == 1
