"""Btrc's JSON scene format (factory/node/parser.cpp, node.cpp,
factory/scene.cpp, cli/main.cpp) read into this package's scene: text
includes, references, comments, typed values and transforms with the
reference's conventions, the creators, and the post processors' outputs."""
import math
import os
import struct

import numpy as np
import pytest

from btrc_amd import _abi, scenes
from btrc_amd import scene_json as sj
from oracle import oracle

from util import render_params

HERE = os.path.dirname(os.path.abspath(__file__))
BOX = os.path.join(HERE, "golden", "scene_json", "box.json")


def test_front_end_includes_comments_references():
    root = sj.parse_text('''{
        // a comment
        "a": { "b": [1, 2, "$reference{c}"], "c": 3 },
        /* block */ "d": "$reference{$root/a/b}",
        "e": "$reference{a/c}"
    }''')
    assert root["a"]["b"] == [1, 2, 3]
    assert root["d"] is root["a"]["b"]
    assert root["e"] == 3
    with pytest.raises(sj.SceneFileError):
        sj.parse_text('{"a": "$reference{nothing/here}"}')


def test_transforms_follow_the_reference():
    def pt(m, p):
        return (m @ np.array([*p, 1.0]))[:3]
    m, i = sj.transform3d({"rotate_x": {"deg": 90}})
    assert np.allclose(pt(m, (0, 1, 0)), (0, 0, 1))            # Mat4::rotate_x, counter-clockwise
    assert np.allclose(m @ i, np.eye(4))
    m2, i2 = sj.transform3d({"rotate": {"axis": [1, 0, 0], "angle": {"rad": math.pi / 2}}})
    assert np.allclose(pt(m2, (0, 1, 0)), (0, 0, -1))          # Mat4::rotate stores the transpose
    assert np.allclose(m2 @ i2, np.eye(4))
    mz, iz = sj.transform3d({"rotate_z": {"deg": 90}})
    assert np.allclose(pt(mz, (1, 0, 0)), (0, -1, 0)) and np.allclose(mz, iz)   # the reference's pair
    m3, i3 = sj.transform3d([{"scale": 2}, {"translate": [1, 0, 0]}])
    assert np.allclose(pt(m3, (1, 1, 1)), (3, 2, 2))           # applied first to last
    assert np.allclose(m3 @ i3, np.eye(4))
    assert sj.vec3(0.5) == [0.5] * 3 and sj.vec3([2]) == [2.0] * 3
    assert sj.degree({"rad": math.pi}) == pytest.approx(180.0)


def test_obj_loader_fans_quads_and_unit_cube():
    pos, nor, uvs = sj.load_obj(os.path.join(HERE, "golden", "scene_json", "cube.obj"), unit_cube=True)
    assert pos.shape == (12, 9)
    p = pos.reshape(-1, 3)
    assert p.min() == -0.5 and p.max() == 0.5
    # per-corner normals of the first face (z = -1)
    assert np.allclose(nor[0].reshape(3, 3), [[0, 0, -1]] * 3)
    # the quad (v1, v4, v3, v2) becomes (v1, v4, v3), (v1, v3, v2)
    assert np.allclose(pos[0].reshape(3, 3), [[-.5, -.5, -.5], [-.5, .5, -.5], [.5, .5, -.5]])
    assert np.allclose(pos[1].reshape(3, 3), [[-.5, -.5, -.5], [.5, .5, -.5], [.5, -.5, -.5]])
    a, b, c = pos.reshape(-1, 3, 3).transpose(1, 0, 2)
    n = np.cross(b - a, c - a)
    assert (np.einsum("ij,ij->i", n, a + b + c) > 0).all()     # every face wound outward


def test_load_box_scene():
    sf = sj.load(BOX)
    s = sf.scene
    assert (sf.width, sf.height) == (48, 36)
    assert len(s.instances) == 4 and len(s.geometries) == 4
    assert [m["type"] for m in s.materials] == [0, 1, 2, 0]      # diffuse, disney, invisible, black
    assert s.materials[1]["roughness"] == pytest.approx(0.35) and s.materials[1]["clearcoat"] == 0.4
    assert s.materials[3]["color"] == [0.0, 0.0, 0.0]
    assert s.instances[3]["light"] == (12.0, 11.0, 9.0)
    assert len(s.media) == 1 and s.instances[2]["inner"] == 0
    assert len(s.volumes) == 1
    assert sf.renderer["spp"] == 4 and sf.renderer["state_count"] == 4000
    assert sf.filter == ("gaussian", 1.3, 1.6)
    assert [p["type"] for p in sf.post_processors] == ["aces_tonemap", "save_to_image"]
    # the oracle renders it: lights, media and the volume all contribute
    p = render_params(sf.width, sf.height, 2, min_depth=3, max_depth=8)
    img, e, sh = oracle.OracleScene(s).render(sf.camera, p)
    assert np.isfinite(img).all() and img.sum() > 0 and e > 0 and sh > 0


def ibl_scene_file(tmp_path, tex):
    """the box scene under an image-based environment light"""
    src = open(BOX).read().replace("$env{scene_directory}", os.path.dirname(BOX))
    src = src.replace('"volumes": [', '"envir_light": { "type": "ibl", "tex": ' + tex +
                      ', "up": [ 0, 1, 0 ], "lut_width": 32, "lut_height": 16 },\n    "volumes": [')
    (tmp_path / "materials.json").write_text(open(os.path.join(os.path.dirname(BOX), "materials.json")).read())
    path = tmp_path / "ibl.json"
    path.write_text(src)
    return str(path)


def test_ibl_light(tmp_path):
    sky = scenes.sky_texture(32, 16, seed=2)
    sj.save_image(str(tmp_path / "sky.hdr"), sky)
    sf = sj.load(ibl_scene_file(tmp_path, '{ "type": "array", "filename": "' + str(tmp_path / "sky.hdr") + '" }'))
    kind, tex, up, lw, lh = sf.scene.envir
    assert kind == "ibl" and up == (0.0, 1.0, 0.0) and (lw, lh) == (32, 16)
    t = sf.scene.textures[tex]["texels"]
    assert t.shape == (16, 32, 4) and t.dtype == np.float32
    # RGBE keeps 8 bits of mantissa under the texel's largest channel
    assert (np.abs(t[..., :3] - sky[..., :3]) <= sky[..., :3].max(axis=2, keepdims=True) / 128).all()
    p = render_params(sf.width, sf.height, 2, min_depth=3, max_depth=8)
    img, e, sh = oracle.OracleScene(sf.scene).render(sf.camera, p)
    assert np.isfinite(img).all() and img.sum() > 0
    # a constant texture becomes a one-texel texture; the defaults are up = +z and a 400 x 200 table
    src = open(ibl_scene_file(tmp_path, '[ 0.5, 0.25, 1 ]')).read()
    src = src.replace(', "up": [ 0, 1, 0 ], "lut_width": 32, "lut_height": 16', "")
    (tmp_path / "const.json").write_text(src)
    sf = sj.load(str(tmp_path / "const.json"))
    kind, tex, up, lw, lh = sf.scene.envir
    assert up == (0.0, 0.0, 1.0) and (lw, lh) == (400, 200)
    np.testing.assert_array_equal(sf.scene.textures[tex]["texels"], [[[0.5, 0.25, 1.0, 1.0]]])


@pytest.mark.parametrize("snippet,name", [
    ('{"type": "glass"}', "glass"), ('{"type": "metal"}', "metal"), ('{"type": "mirror"}', "mirror")])
def test_unsupported_types_are_named(tmp_path, snippet, name):
    src = open(BOX).read().replace('"material": { "type": "invisible" }', f'"material": {snippet}')
    path = tmp_path / "scene.json"
    path.write_text(src.replace("$env{scene_directory}", os.path.dirname(BOX)))
    (tmp_path / "materials.json").write_text(open(os.path.join(os.path.dirname(BOX), "materials.json")).read())
    with pytest.raises(sj.SceneFileError, match=name):
        sj.load(str(path))


def test_image_writers(tmp_path):
    from PIL import Image
    img = np.array([[[0.0, 0.5, 0.999], [1.0, -0.2, 0.00390625]]], np.float32)
    sj.save_image(str(tmp_path / "a.png"), np.concatenate([img, np.ones((1, 2, 1), np.float32)], axis=2))
    got = np.asarray(Image.open(tmp_path / "a.png"))
    assert got.tolist() == [[[0, 128, 255], [255, 0, 1]]]       # min(int(256 x), 255)
    rgb = np.random.default_rng(1).uniform(0, 4, (3, 5, 3)).astype(np.float32)
    sj.save_image(str(tmp_path / "a.exr"), rgb)
    data = open(tmp_path / "a.exr", "rb").read()
    assert struct.unpack("<i", data[:4])[0] == 20000630
    # the last scanline holds B, G, R of the bottom row
    last = data[-(4 * 5 * 3):]
    assert np.array_equal(np.frombuffer(last, "<f4").reshape(3, 5), rgb[2].T[::-1])


def test_textured_parameters_and_normal_maps(tmp_path):
    """a Disney material with textures on sheen and ior and a normal map, and
    a diffuse material with a normal map (DisneyMaterialCreator,
    disney.cpp:844-894; NormalMap::load, normal_map.cpp:5-11)"""
    sky = scenes.sky_texture(8, 4, seed=1)
    sj.save_image(str(tmp_path / "t.hdr"), sky)
    tex = '{ "type": "array", "filename": "' + str(tmp_path / "t.hdr") + '" }'
    mats = open(os.path.join(os.path.dirname(BOX), "materials.json")).read()
    mats = mats.replace('"specular": 0.5, "clearcoat": 0.4 }',
                        f'"specular": 0.5, "clearcoat": 0.4, "sheen": {tex}, "ior": {tex}, "normal": {tex} }}')
    mats = mats.replace('"white": { "type": "diffuse", "color": [ 0.73, 0.73, 0.73 ] }',
                        '"white": { "type": "diffuse", "color": [ 0.73, 0.73, 0.73 ], "normal": [ 0.5, 0.5, 1 ] }')
    (tmp_path / "materials.json").write_text(mats)
    src = open(BOX).read().replace("$env{scene_directory}", os.path.dirname(BOX))
    (tmp_path / "scene.json").write_text(src)
    sf = sj.load(str(tmp_path / "scene.json"))
    glossy = [m for m in sf.scene.materials if m["type"] == _abi.MATERIAL_DISNEY and m.get("sheen_texture") is not None]
    assert glossy, "the Disney material lost its sheen texture"
    g = glossy[0]
    for slot in ("sheen", "ior", "normal"):      # one texture object per node, as the factory makes them
        assert sf.scene.textures[g[slot + "_texture"]]["texels"].shape == (4, 8, 4)
    assert g["specular"] == 0.5 and g.get("specular_texture") is None
    white = [m for m in sf.scene.materials if m["type"] == _abi.MATERIAL_DIFFUSE and m.get("normal_texture") is not None]
    assert white and sf.scene.textures[white[0]["normal_texture"]]["texels"].shape == (1, 1, 4)
    p = render_params(sf.width, sf.height, 1, min_depth=3, max_depth=4)
    img = oracle.OracleScene(sf.scene).render(sf.camera, p)[0]
    assert np.isfinite(img).all() and img.sum() > 0
