"""The sun (sun_light, sun.h:23-80; the `sun` line, scene.h:543-562,
1217-1235): a distant light of small angular size, never hit, added to every
escaping ray (camera.h:515-529) and to camera rays only while its disk is
visible, sampled uniformly in its cone by next event estimation, weighed by
light_list by pi R^2 times its irradiance's luminance (light_list.h:66-67).

CPU: the scene line; the oracle's light probabilities against the
reference's light_list (tests/golden/sun.npz, written by the compiled
reference); the oracle's converged render of scenes/sun.scene against the
reference's; and, independent of the reference, the closed form of a sunlit
Lambertian ground: albedo / pi * E * cos(zenith) * (1 + cos(half angle)) / 2.
GPU: hit digests, the oracle at equal seed (disk hidden and visible), and
converged means against the same golden."""
import hashlib
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as ol
from conftest import ROOT, scene_path

GOLD = os.path.join(ROOT, "tests", "golden", "sun.npz")
SCENE = "sun.scene"


def _golden():
    g = np.load(GOLD)
    return g, json.loads(str(g["digest"]))


def test_sun_line():
    import mypbr_amd as m
    d = m.describe(scene_path(SCENE))
    el, az = 12 * math.pi / 180.0, 8 * math.pi / 180.0  # degrees_to_radians (rtweekend.h)
    want = [math.cos(el) * math.sin(az), math.sin(el), -math.cos(el) * math.cos(az)]
    assert d.desc.has_sun == 1 and d.camera.sun_disk_hidden == 1
    assert list(d.desc.sun_direction) == want
    assert d.desc.sun_angle_degrees == 5 and list(d.desc.sun_irradiance) == [2.5, 2.5 * 0.9, 2.5 * 0.75]
    lights = [d.desc.light_object[k] for k in range(d.desc.n_lights)]
    assert lights[-1] == -2 and lights.count(-2) == 1  # created at the end of the file, after the quad


@pytest.mark.parametrize("line,ok,on", [
    ("sun elevation 30 azimuth 10", True, 1), ("sun", True, 1), ("sun off", True, 0),
    ("sun elevation 30 off", True, 0), ("sun elevation 91", False, 0), ("sun angle 0", False, 0),
    ("sun angle 90", False, 0), ("sun intensity -1", False, 0), ("sun disk maybe", False, 0),
    ("sun color 1 1", False, 0), ("sun nosuch 1", False, 0),
])
def test_sun_lines(tmp_path, line, ok, on):
    import mypbr_amd as m
    s = tmp_path / "s.scene"
    s.write_text(f"image_width 8\n{line}\nsphere 0 0 -1 0.5\n")
    if ok:
        d = m.describe(str(s))
        assert d.desc.has_sun == on
        if on:  # the reference's defaults (scene.h:424-431): irradiance 3, angle 4, disk on
            assert d.camera.sun_disk_hidden == 0 and d.desc.sun_angle_degrees == (4 if line == "sun" else 4)
            assert list(d.desc.sun_irradiance) == [3, 3, 3]
    else:
        with pytest.raises(m.MptError):
            m.describe(str(s))


def test_oracle_light_probabilities_match_reference():
    import mypbr_amd as m
    g, _ = _golden()
    d = m.describe(scene_path(SCENE))
    got = ol.OracleScene(d.desc).light_probabilities(d.desc.n_lights)
    assert np.array_equal(got, g["light_probabilities"])


def test_oracle_sunlit_ground_closed_form(tmp_path):
    """An unshadowed Lambertian ground under the sun alone (black sky):
    L = albedo / pi * integral of E / omega * cos over the cone, which is
    albedo / pi * E * cos(zenith) * (1 + cos(half angle)) / 2 for a cone
    wholly above the horizon. A lost 1 / omega, radiance taken for
    irradiance, a degree/radian or diameter/half-angle slip, or a cone
    sample that is not uniform moves the mean by far more than 0.5%."""
    import mypbr_amd as m
    el, angle, e, albedo = 50.0, 8.0, 2.0, 0.5
    s = tmp_path / "sunlit.scene"
    s.write_text(f"image_width 24\naspect_ratio 1\nsamples_per_pixel 256\nmax_depth 3\nbackground 0 0 0\n"
                 f"camera lookfrom 0 6 4 lookat 0 0 0 vfov 20\n"
                 f"sun elevation {el} azimuth 70 intensity {e} angle {angle}\n"
                 f"material white lambertian {albedo} {albedo} {albedo}\nground white\n"
                 f"sphere 50 -30 50 0.1 white\n")
    d = m.describe(str(s))
    img = ol.OracleScene(d.desc).render(d.camera, seed=3)["radiance"]
    half = math.radians(angle) / 2
    want = albedo / math.pi * e * math.sin(math.radians(el)) * (1 + math.cos(half)) / 2
    assert abs(img.mean() / want - 1) < 0.005, (img.mean(), want)
    # the sun's disk is outside this view: a camera ray never reaches it
    assert np.all(np.isfinite(img)) and img.max() < 10 * want


def test_oracle_render_converges_to_reference():
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
@pytest.mark.parametrize("hidden", [1, 0])
def test_gpu_render_matches_oracle(hidden):
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_gpu_render import _compare
    d = m.describe(scene_path(SCENE))
    cam = camera_copy(d.camera, image_width=64, samples_per_pixel=8, max_depth=12, collect_features=1,
                      sun_disk_hidden=hidden)
    cpu = ol.OracleScene(d.desc).render(cam, seed=7)
    f = m.Scene.from_desc(d.desc, d.camera).render(cam, seed=7)
    _compare(f.radiance, cpu["radiance"], 0.97, 1e-3)
    _compare(f.albedo, cpu["albedo"], 0.97, 1e-3)
    if not hidden:  # the disk is in view: camera rays that reach it carry E / omega
        assert f.radiance.max() > 10


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
