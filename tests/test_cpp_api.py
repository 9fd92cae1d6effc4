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


def _light_lines(out):
    box = None
    mirror, loaded, power, picks, core = [], [], [], [], []
    for line in out.splitlines():
        f = line.split()
        if f[0] == "box":
            box = [float(x) for x in f[1:]]
        elif f[0] == "light":
            mirror.append(float(f[2]))
            loaded.append(float(f[3]))
            power.append(float(f[4]))
        elif f[0] == "pick":
            picks.append((float(f[1]), int(f[2])))
        elif f[0] == "core":
            core.append(float(f[2]))
    return box, mirror, loaded, power, picks, core


@pytest.mark.parametrize("name", ["cornell.scene", "test_lights.scene"])
def test_cpp_light_list_matches_reference_probabilities(tmp_path, name):
    """light_list::build(world.bounding_box()), probability(k) and pick(xi)
    (light_list.h:45-127), and hittable::bounding_box / power, against the
    probabilities the compiled reference's light_list gives for the same file
    (tests/golden/light_probs.json) bit for bit. The GPU part (the uploaded
    scene's probabilities) runs where a GPU is present."""
    import json
    import mypbr_amd as m
    exe = _build(tmp_path)
    r = subprocess.run([exe, "lights", scene_path(name)], capture_output=True, text=True, timeout=120)
    box, mirror, loaded, power, picks, core = _light_lines(r.stdout)
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "light_probs.json")))[name]
    assert mirror == gold and loaded == gold  # the loader's build and a program's build agree
    # hittable::power of the emitters; the environment has none without a radius (environment.h:115)
    assert [p > 0 for p in power] == {"cornell.scene": [True], "test_lights.scene": [True, True, True, False]}[name]
    # the world box is the one the core builds its BVH from
    d = m.describe(scene_path(name))
    import ctypes as Cc
    b6 = (Cc.c_double * 6)()
    m._check(m.lib.mpt_desc_bounds(Cc.byref(d.desc), b6))
    assert box == list(b6)
    cdf = np.concatenate([[0.0], np.cumsum(gold)])
    cdf[-1] = 1.0
    for xi, k in picks:
        assert k == min(max(int(np.searchsorted(cdf, xi, side="right")), 1), len(gold)) - 1
    if r.returncode == 0:  # the GPU part ran
        assert core == gold


def test_desc_light_entry_points_match_reference_probabilities():
    import ctypes as Cc
    import json
    import mypbr_amd as m
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "light_probs.json")))
    for name, probs in gold.items():
        d = m.describe(scene_path(name))
        n = len(probs)
        out = (Cc.c_double * max(n, 1))()
        m._check(m.lib.mpt_desc_light_probabilities(Cc.byref(d.desc), None, out))
        assert list(out)[:n] == probs, name
        pw = (Cc.c_double * max(n, 1))()
        m._check(m.lib.mpt_desc_light_powers(Cc.byref(d.desc), None, pw))
        assert all(x > 0 for x in list(pw)[:n]), name
    # bounds given explicitly: a bigger scene radius raises the environment's share (light_list.h:52-58)
    d = m.describe(scene_path("test_lights.scene"))
    b = (Cc.c_double * 6)()
    m._check(m.lib.mpt_desc_bounds(Cc.byref(d.desc), b))
    big = (Cc.c_double * 6)(*[x * 10 for x in b])
    out = (Cc.c_double * 4)()
    m._check(m.lib.mpt_desc_light_probabilities(Cc.byref(d.desc), big, out))
    assert out[3] > gold["test_lights.scene"][3] and abs(sum(out) - 1) < 1e-12


def test_cpp_atmosphere_bake_equals_the_loaders(tmp_path):
    """mpt::atmosphere (atmosphere.h:36-165) for in-memory scenes: its bake and
    sun_transmittance equal, bit for bit, what the .scene loader makes of
    scenes/sky_atmosphere.scene (the environment's pixels and the dimmed sun),
    which the sun tests check against the compiled reference."""
    import mypbr_amd as m
    exe = _build(tmp_path)
    out = str(tmp_path / "air.bin")
    r = subprocess.run([exe, "atmosphere", out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, dtype=np.float64)
    keep = m.describe(scene_path("sky_atmosphere.scene"))  # owns the arrays d points to
    d = keep.desc
    assert (d.env_width, d.env_height) == (256, 128)
    env = np.ctypeslib.as_array(d.env_rgb, (256 * 128 * 3,))
    assert np.array_equal(got[:-3].view(np.uint64), env.view(np.uint64))
    assert np.array_equal(got[-3:].view(np.uint64), np.array(list(d.sun_irradiance)).view(np.uint64))
