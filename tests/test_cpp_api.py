"""The C++ scene API (include/mpt/mpt.hpp) as a reference program uses it.

tests/cpp/api_render.cpp is written the way a program of the reference is
(load_scene or objects, then camera::render(world, lights, save_every,
on_pass), camera.h:79-105). On the CPU it must compile and link against the
library; on the GPU its render of a .scene file must equal, bit for bit, the
C ABI's render of the same file at the same seed, and on_pass must follow
camera.h:98-101 (every save_every passes, not after the last)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, scene_path

LIB_DIR = os.path.join(ROOT, "mypbr_amd")
SRC = os.path.join(ROOT, "tests", "cpp", "api_render.cpp")


def _build(tmp_path):
    exe = str(tmp_path / "api_render")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
           "-L", LIB_DIR, "-lmpt", "-Wl,-rpath," + LIB_DIR, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_cpp_program_builds_against_the_library(tmp_path):
    exe = _build(tmp_path)
    assert os.access(exe, os.X_OK)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
def test_cpp_render_of_scene_file_equals_c_abi_render(tmp_path):
    import mypbr_amd
    from mypbr_amd.abi import camera_copy
    exe = _build(tmp_path)
    out = str(tmp_path / "cornell.bin")
    r = subprocess.run([exe, "scene", scene_path("cornell.scene"), "40", "3", "5", out],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    h = int(r.stdout.split()[1])
    cpp = np.fromfile(out, dtype=np.float64).reshape(h, 40, 3)
    s = mypbr_amd.Scene.load(scene_path("cornell.scene"))
    cam = camera_copy(s.camera)
    cam.image_width = 40
    cam.samples_per_pixel = 3
    py = s.render(cam, seed=5).radiance
    assert py.shape == cpp.shape
    assert np.array_equal(py, cpp)
    assert py.mean() > 0


@pytest.mark.gpu
def test_cpp_objects_render_and_pass_callbacks(tmp_path):
    exe = _build(tmp_path)
    out = str(tmp_path / "objects.bin")
    r = subprocess.run([exe, "objects", out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    f = r.stdout.split()
    h, calls, last = int(f[1]), int(f[3]), int(f[5])
    assert (calls, last) == (2, 4)  # 6 spp, save_every 2: after passes 2 and 4
    data = np.fromfile(out, dtype=np.float64).reshape(2, h, 48, 3)
    final, partial = data
    assert np.all(np.isfinite(final)) and np.all(final >= 0)
    assert 0.02 < final.mean() < 5
    # the callback frame is the average of passes 0..3; the final frame adds
    # passes 4 and 5, whose sum cannot be negative
    assert np.all(6 * final - 4 * partial >= -1e-9 * np.maximum(1, 6 * final))
    assert not np.array_equal(final, partial)
    r2 = subprocess.run([exe, "objects", out + "2"], capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0 and np.array_equal(np.fromfile(out + "2", dtype=np.float64), data.ravel())


@pytest.mark.gpu
def test_cpp_denoiser_equals_c_abi_denoise(tmp_path):
    import mypbr_amd
    exe = _build(tmp_path)
    out = str(tmp_path / "dn.bin")
    r = subprocess.run([exe, "denoise", scene_path("cornell.scene"), out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    h = int(r.stdout.split()[1])
    data = np.fromfile(out, dtype=np.float64)
    n = h * 40
    rad, alb, nrm, dn = (data[k * 3 * n:(k + 1) * 3 * n].reshape(h, 40, 3) for k in range(4))
    var = data[12 * n:].reshape(h, 40)
    want = mypbr_amd.denoise({"radiance": rad, "albedo": alb, "normal": nrm, "variance": var})
    assert np.array_equal(dn, want)
