"""The C-ABI library: it loads, exports every symbol include/btrc_hip.h
declares, and its struct layouts match the Python mirror. No GPU calls."""
import ctypes as C
import subprocess
import os
import re

import pytest

import btrc_amd
from btrc_amd import _abi

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "btrc_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(btrc_hip_\w+)\s*\(", text, re.M)


def test_header_declarations_match_mirror():
    assert declared_functions() == _abi.EXPORTS


def test_library_exports_every_declared_symbol():
    lib = btrc_amd.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_struct_sizes():
    # sizes fixed by include/btrc_hip.h (all fields 4-byte, doubles/uint64 8-aligned)
    assert C.sizeof(_abi.Ray) == 32
    assert C.sizeof(_abi.Hit) == 20
    assert C.sizeof(_abi.CameraDesc) == 44
    assert C.sizeof(_abi.RenderParams) == 64
    assert C.sizeof(_abi.MaterialDesc) == 80 + 11 * 4
    assert C.sizeof(_abi.TextureDesc) == 4 * 4 + 8 + 4 * 4 + 16
    assert C.sizeof(_abi.InstanceDesc) == 4 * (2 + 32 + 1 + 3 + 2)
    assert C.sizeof(_abi.RenderStats) == 8 * (1 + 4 + 7 + 7)


STRUCTS = {
    "btrc_geometry_desc": "GeometryDesc", "btrc_material_desc": "MaterialDesc",
    "btrc_instance_desc": "InstanceDesc", "btrc_medium_desc": "MediumDesc",
    "btrc_envir_light_desc": "EnvirLightDesc", "btrc_scene_desc": "SceneDesc",
    "btrc_camera_desc": "CameraDesc", "btrc_render_params": "RenderParams",
    "btrc_filter_desc": "FilterDesc", "btrc_render_stats": "RenderStats", "btrc_ray": "Ray",
    "btrc_hit": "Hit", "btrc_tuning": "Tuning", "btrc_nrc_params": "NrcParams", "btrc_nrc_stats": "NrcStats",
    "btrc_texture_desc": "TextureDesc", "btrc_view_desc": "ViewDesc", "btrc_volume_desc": "VolumeDesc",
    "btrc_bvh_node": "BvhNode",
}


def test_mirror_layout_matches_header(tmp_path):
    """every field of every ctypes mirror sits at the offset the C compiler
    gives the header's struct"""
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{os.path.join(ROOT, "include", "btrc_hip.h")}"',
             "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        lines.append(f'printf("{pyname} size %zu\\n", sizeof({cname}));')
        for field, _ in getattr(_abi, pyname)._fields_:
            lines.append(f'printf("{pyname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        pyname, what, value = line.split()
        cls = getattr(_abi, pyname)
        want = C.sizeof(cls) if what == "size" else getattr(cls, what).offset
        assert int(value) == want, f"{pyname}.{what}: header {value}, mirror {want}"
    assert _abi.ray_dtype().itemsize == 32 and _abi.hit_dtype().itemsize == 20


def test_null_context_and_bad_arguments_fail_cleanly():
    lib = btrc_amd.load_library()
    assert lib.btrc_hip_create(0, None) == _abi.BTRC_ERR_INVALID
    assert lib.btrc_hip_upload_scene(None, None) == _abi.BTRC_ERR_INVALID
    assert lib.btrc_hip_render(None, None, None) == _abi.BTRC_ERR_INVALID
    assert lib.btrc_hip_request_stop(None) == _abi.BTRC_ERR_INVALID
    assert lib.btrc_hip_last_error(None) == b"null context"


def test_unknown_renderer_parameter_is_rejected():
    with pytest.raises(btrc_amd.BtrcException):
        btrc_amd.WavefrontPathTracer(context=object(), sppp=4)
