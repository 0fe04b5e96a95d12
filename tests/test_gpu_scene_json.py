"""A scene file in Btrc's JSON format rendered through the loader and the
C ABI, against the oracle per pixel, and BtrcCLI's run() end to end."""
import os
import shutil

import numpy as np
import pytest

import btrc_amd
from btrc_amd import scene_json as sj
from oracle import oracle

from btrc_amd import scenes
from test_scene_json import ibl_scene_file
from util import render_params

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "golden", "scene_json")


def test_scene_file_per_pixel_parity(ctx):
    check_scene_file(ctx, sj.load(os.path.join(SRC, "box.json")))


def test_ibl_scene_file_per_pixel_parity(ctx, tmp_path):
    sj.save_image(str(tmp_path / "sky.hdr"), scenes.sky_texture(32, 16, seed=2))
    check_scene_file(ctx, sj.load(ibl_scene_file(tmp_path, '{ "type": "array", "filename": "' +
                                                 str(tmp_path / "sky.hdr") + '" }')))


def check_scene_file(ctx, sf):
    r = btrc_amd.WavefrontPathTracer(ctx, **sf.renderer)
    r.set_scene(sf.scene)
    r.set_camera(sf.camera)
    r.set_film(sf.width, sf.height)
    r.set_film_filter(*sf.filter)
    res = r.render()
    p = render_params(sf.width, sf.height, sf.renderer["spp"], min_depth=sf.renderer["min_depth"],
                      max_depth=sf.renderer["max_depth"], state_count=sf.renderer["state_count"])
    ref, e, s = oracle.OracleScene(sf.scene).render(sf.camera, p, filt=sf.filter)
    ref /= p.spp
    assert res.stats.extension_rays == e and res.stats.shadow_rays == s
    rel = np.abs(res.color[..., :3] - ref) / np.maximum(np.abs(ref), 1e-2)
    assert rel.max() <= 1e-4, f"max rel err {rel.max()}"
    assert ref.sum() > 0


def test_cli_run_writes_the_image(ctx, tmp_path):
    for name in ("box.json", "materials.json", "cube.obj"):
        shutil.copy(os.path.join(SRC, name), tmp_path / name)
    shutil.copy(os.path.join(HERE, "golden", "cornell_box.obj"), tmp_path / "cornell_box.obj")
    src = (tmp_path / "box.json").read_text().replace("$env{scene_directory}/../cornell_box.obj",
                                                      "$env{scene_directory}/cornell_box.obj")
    (tmp_path / "box.json").write_text(src)
    os.makedirs(tmp_path / "out")
    res = sj.render_file(str(tmp_path / "box.json"), ctx=ctx, log=lambda *_: None)
    from PIL import Image
    png = np.asarray(Image.open(tmp_path / "out" / "color.png"))
    want = np.clip(np.floor(np.maximum(res.color[..., :3], 0) ** (1 / 2.2) * 256), 0, 255).astype(np.uint8)
    assert png.shape == (36, 48, 3) and np.array_equal(png, want)
    # tone mapped on the device: the oracle's ACES fit of the film, bit for bit
    film = ctx.film(48, 36)[0]
    assert np.array_equal(res.color.view(np.uint32), oracle.aces(film, 1.0).view(np.uint32))
    assert (res.color[..., :3] <= 1.0).all()
