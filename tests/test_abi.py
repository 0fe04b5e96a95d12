"""The C-ABI library: it loads, exports every symbol include/btrc_hip.h
declares, and its struct layouts match the Python mirror. No GPU calls."""
import ctypes as C
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
    assert C.sizeof(_abi.MaterialDesc) == 68
    assert C.sizeof(_abi.InstanceDesc) == 4 * (2 + 32 + 1 + 3 + 2)
    assert C.sizeof(_abi.RenderStats) == 8 * (1 + 4 + 4 + 4)
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
