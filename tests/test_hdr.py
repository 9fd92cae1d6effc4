"""Radiance .hdr environment maps (environment_light's loader,
environment.h:30-41, through stb_image).

CPU: the loader's pixels equal stb_image's floats bit for bit (the compiled
reference's own stb_image, oracle/_ref) on a run-length encoded map of this
repository, on a flat map written by mpt_write_image and, in the container
that holds the reference sources, on the reference's own HDRI; malformed
files are refused. scenes/hdr_env.scene (the reference's hdri.scene layout
on this repository's map) against tests/golden/hdr_env.npz, written by the
compiled reference: hits bitwise, converged means. GPU: the same, and the
render against the oracle at equal seed."""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_lib as ol
from conftest import ROOT, scene_path

GOLD = os.path.join(ROOT, "tests", "golden", "hdr_env.npz")
SCENE = "hdr_env.scene"
REF_HDRI = "/root/reference/textures/quarry_01_1k.hdr"


def _golden():
    g = np.load(GOLD)
    return g, json.loads(str(g["digest"]))


def loaded_env(path, tmp_path):
    """The environment pixels the scene loader reads from path."""
    import mypbr_amd as m
    s = tmp_path / "e.scene"
    s.write_text(f"image_width 8\nenvironment {path}\nsphere 0 0 -1 0.5\n")
    d = m.describe(str(s))
    w, h = d.desc.env_width, d.desc.env_height
    return np.ctypeslib.as_array(d.desc.env_rgb, shape=(w * h * 3,)).reshape(h, w, 3).copy()


def stb_floats(path):
    R = ol.reference()
    R.ref_load_image_f.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_float), C.c_int64]
    R.ref_load_image_f.restype = C.c_int
    w, h = C.c_int(), C.c_int()
    buf = np.zeros(4096 * 2048 * 3, np.float32)
    assert R.ref_load_image_f(os.fsencode(path), C.byref(w), C.byref(h), buf.ctypes.data_as(C.POINTER(C.c_float)),
                              buf.size) == 0
    return buf[: w.value * h.value * 3].reshape(h.value, w.value, 3)


@pytest.mark.skipif(not ol.reference_available(), reason="the compiled reference (oracle/_ref) is not built")
def test_hdr_pixels_equal_stb_image(tmp_path):
    import mypbr_amd as m
    paths = [os.path.join(ROOT, "scenes", "models", "sky_rle.hdr")]
    flat = str(tmp_path / "flat.hdr")
    m.write_image(flat, np.random.default_rng(1).lognormal(0, 2, (20, 30, 3)))
    paths.append(flat)
    if os.path.exists(REF_HDRI):
        paths.append(REF_HDRI)
    for p in paths:
        assert np.array_equal(loaded_env(p, tmp_path), stb_floats(p).astype(np.float64)), p


def test_hdr_values_known(tmp_path):
    """RGBE decoding by hand: m * 2^(e - 136), 0 for e = 0, from a flat file."""
    import mypbr_amd as m
    rows = [bytes([128, 64, 32, 129, 0, 0, 0, 0, 255, 1, 0, 120])]
    p = tmp_path / "k.hdr"
    p.write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 1 +X 3\n" + rows[0])
    got = loaded_env(str(p), tmp_path)
    expect = np.array([[[128 * 2.0 ** -7, 64 * 2.0 ** -7, 32 * 2.0 ** -7], [0, 0, 0],
                        [255 * 2.0 ** -16, 1 * 2.0 ** -16, 0]]])
    assert np.array_equal(got, expect)
    del m


@pytest.mark.parametrize("data", [
    b"#?NOTRADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 1 +X 1\n\x80\x80\x80\x80",
    b"#?RADIANCE\nFORMAT=32-bit_rle_xyze\n\n-Y 1 +X 1\n\x80\x80\x80\x80",
    b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n+Y 1 +X 1\n\x80\x80\x80\x80",
    b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 1\n\x80\x80\x80\x80",                # truncated
    b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 1 +X 8\n\x02\x02\x00\x08\x85\x01",        # truncated run data
])
def test_malformed_hdr_is_refused(tmp_path, data):
    import mypbr_amd as m
    p = tmp_path / "bad.hdr"
    p.write_bytes(data)
    with pytest.raises(m.MptError):
        loaded_env(str(p), tmp_path)


def test_oracle_hits_match_reference():
    import mypbr_amd as m
    g, dg = _golden()
    d = m.describe(scene_path(SCENE))
    o = ol.OracleScene(d.desc)
    rays, tt = o.ray_batch(d.camera, dg["seed"], dg["n"])
    assert np.array_equal(rays[g["index"]], g["rays"]), "the seeded batch generator changed"
    ids, t, f = o.trace(rays, tt)
    assert dg["ties"] == 0 and dg["unidentified"] == 0
    assert hashlib.sha256(ids.astype("<i8").tobytes()).hexdigest() == dg["sha256_id"]
    assert hashlib.sha256(t.astype("<f8").tobytes()).hexdigest() == dg["sha256_t"]
    assert hashlib.sha256(f.astype("u1").tobytes()).hexdigest() == dg["sha256_front"]


def test_oracle_render_converges_to_reference():
    """Environment radiance (bilinear lookups of the decoded map), its
    importance sampling and the rotation all enter this image."""
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_oracle_golden import render_agrees
    g, _ = _golden()
    d = m.describe(scene_path(SCENE))
    cam = camera_copy(d.camera, image_width=int(g["width"]), samples_per_pixel=int(g["spp"]) // 8,
                      max_depth=int(g["max_depth"]))
    o = ol.OracleScene(d.desc)
    runs = np.stack([o.render(cam, seed=100 + k)["radiance"] for k in range(8)])
    render_agrees(runs.mean(0), runs.std(0, ddof=1) / np.sqrt(8), g, SCENE)


def test_scene_display_settings():
    import mypbr_amd as m
    d = m.scene_display(scene_path(SCENE))
    assert (d.exposure, d.tonemap) == (-1.0, 2)


# ---------------------------------------------------------------- GPU

@pytest.mark.gpu
def test_gpu_hits_match_reference():
    import mypbr_amd as m
    g, dg = _golden()
    d = m.describe(scene_path(SCENE))
    rays, tt = ol.OracleScene(d.desc).ray_batch(d.camera, dg["seed"], dg["n"])
    ids, t, f = m.Scene.from_desc(d.desc, d.camera).trace_closest(rays, tt)
    assert hashlib.sha256(ids.astype("<i8").tobytes()).hexdigest() == dg["sha256_id"]
    assert hashlib.sha256(t.astype("<f8").tobytes()).hexdigest() == dg["sha256_t"]
    assert hashlib.sha256(f.astype("u1").tobytes()).hexdigest() == dg["sha256_front"]


@pytest.mark.gpu
def test_gpu_render_matches_oracle_and_writes_the_display_image(tmp_path):
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_gpu_render import _compare
    d = m.describe(scene_path(SCENE))
    cam = camera_copy(d.camera, image_width=64, samples_per_pixel=8, max_depth=16)
    cpu = ol.OracleScene(d.desc).render(cam, seed=7)["radiance"]
    f = m.Scene.from_desc(d.desc, d.camera).render(cam, seed=7)
    _compare(f.radiance, cpu, 0.97, 1e-3)
    disp = m.scene_display(scene_path(SCENE))
    out = tmp_path / "hdr_env.png"
    m.write_image(str(out), f.radiance, disp)
    assert out.stat().st_size > 100


@pytest.mark.gpu
def test_gpu_render_converges_to_reference():
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_oracle_golden import render_agrees
    g, _ = _golden()
    d = m.describe(scene_path(SCENE))
    cam = camera_copy(d.camera, image_width=int(g["width"]), samples_per_pixel=int(g["spp"]) // 8,
                      max_depth=int(g["max_depth"]))
    s = m.Scene.from_desc(d.desc, d.camera)
    runs = np.stack([s.render(cam, seed=300 + k).radiance for k in range(8)])
    ratio, relmse = render_agrees(runs.mean(0), runs.std(0, ddof=1) / np.sqrt(8), g, SCENE)
    print(f"{SCENE}: total ratio {ratio:.5f}, relMSE {relmse:.2e}", flush=True)
