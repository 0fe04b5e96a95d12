"""The C++ host side (include/btrc_hip.hpp): a program written against its
Renderer-shaped WavefrontPathTracer compiles with g++ against the C ABI,
reports library errors as BtrcException, and on the GPU renders the same
image as the oracle for the same scene."""
import os
import subprocess

import numpy as np
import pytest

from btrc_amd import _abi, scenes
from oracle import oracle

from conftest import ROOT
from util import render_params

SRC = os.path.join(ROOT, "tests", "cpp", "render_quads.cpp")
LIBDIR = os.path.join(ROOT, "btrc_amd")


def compile_program(tmp_path):
    exe = str(tmp_path / "render_quads")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    SRC, "-L" + LIBDIR, "-lbtrc_hip", "-Wl,-rpath," + LIBDIR, "-o", exe], check=True)
    return exe


def test_program_compiles_and_reports_errors(tmp_path):
    exe = compile_program(tmp_path)
    r = subprocess.run([exe, str(tmp_path / "film.f32")], capture_output=True, text=True, timeout=120)
    # without a GPU the context cannot be created: BtrcException, status 2
    assert r.returncode in (0, 2), r.stderr
    if r.returncode == 2:
        assert "BtrcException" in r.stderr and "btrc_hip_create" in r.stderr


def quads_scene():
    """the scene tests/cpp/render_quads.cpp builds, through the Python mirror"""
    s = scenes.Scene()
    floor = np.array([[-1, 0, -1, -1, 0, 1, 1, 0, 1], [-1, 0, -1, 1, 0, 1, 1, 0, -1]], np.float32)
    light = np.array([[-0.3, 1.5, -0.3, 0.3, 1.5, 0.3, -0.3, 1.5, 0.3],
                      [-0.3, 1.5, -0.3, 0.3, 1.5, -0.3, 0.3, 1.5, 0.3]], np.float32)
    s.add_instance(s.add_geometry(floor), s.add_diffuse([0.6, 0.5, 0.4]))
    s.add_instance(s.add_geometry(light), s.add_diffuse([0.0, 0.0, 0.0]), light=(4.0, 4.0, 4.0))
    s.set_gradient_sky((0.0, 1.0, 0.0), (0.1, 0.1, 0.1), (0.3, 0.4, 0.6))
    cam = scenes.Camera((0.0, 1.0, 3.0), (0.0, 0.3, 0.0), (0.0, 1.0, 0.0), 45.0)
    return s, cam


@pytest.mark.gpu
def test_program_renders_like_the_oracle(tmp_path):
    exe = compile_program(tmp_path)
    out = tmp_path / "film.f32"
    r = subprocess.run([exe, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    img = np.fromfile(out, np.float32).reshape(32, 48, 4)[..., :3]
    scene, cam = quads_scene()
    p = render_params(48, 32, 4, min_depth=3, max_depth=5, state_count=1 << 14)
    ref, e, sh = oracle.OracleScene(scene).render(cam, p)
    ref /= p.spp
    rel = np.abs(img - ref) / np.maximum(np.abs(ref), 1e-2)
    assert rel.max() <= 1e-4, f"max rel err {rel.max()}"
    assert f"{e} extension rays, {sh} shadow rays" in r.stdout
    assert "reporter: 1 stage, 1 completed," in r.stdout and "last 100.0;" in r.stdout
    toned_max = float(r.stdout.split("tone mapped max ")[1].split()[0])
    assert 0.0 < toned_max <= 1.0
    assert ref.sum() > 0
