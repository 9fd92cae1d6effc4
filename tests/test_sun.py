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
            assert d.camera.sun_disk_hidden == 0 and d.desc.sun_angle_degrees == 4
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
    _compare(f.radiance, cpu["radiance"], 0.995, 1e-3)
    _compare(f.albedo, cpu["albedo"], 0.995, 1e-3)
    if not hidden:
        # the disk is in view: camera rays that reach it carry E / omega
        # (about 418 here), and a camera ray's draws before it involve no
        # transcendental that could round apart, so every pixel where the
        # oracle shows the disk must be equal
        disk = cpu["radiance"].max(axis=-1) > 20
        assert disk.sum() >= 5
        assert np.array_equal(f.radiance[disk], cpu["radiance"][disk])
        assert f.radiance.max() > 10


DISK = "sun_disk.scene"
DISK_GOLD = os.path.join(ROOT, "tests", "golden", "sun_disk.npz")


def _disk_runs(render):
    """8 seeded low-resolution renders of scenes/sun_disk.scene at the
    golden's size; their mean and standard error against the reference's."""
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    g = np.load(DISK_GOLD)
    d = m.describe(scene_path(DISK))
    assert d.camera.sun_disk_hidden == 0
    cam = camera_copy(d.camera, image_width=int(g["width"]), samples_per_pixel=int(g["spp"]) // 8,
                      max_depth=int(g["max_depth"]))
    runs = np.stack([render(d, cam, k) for k in range(8)])
    assert runs.max() > 20, "the disk is not in view"
    return runs.mean(0), runs.std(0, ddof=1) / np.sqrt(8), g


def test_oracle_disk_converges_to_reference():
    from test_oracle_golden import render_agrees
    mean, se, g = _disk_runs(lambda d, cam, k: ol.OracleScene(d.desc).render(cam, seed=500 + k)["radiance"])
    render_agrees(mean, se, g, DISK)


@pytest.mark.gpu
def test_gpu_disk_converges_to_reference():
    import mypbr_amd as m
    from test_oracle_golden import render_agrees
    mean, se, g = _disk_runs(lambda d, cam, k: m.Scene.from_desc(d.desc, d.camera).render(cam, seed=700 + k).radiance)
    ratio, relmse = render_agrees(mean, se, g, DISK)
    print(f"{DISK}: total ratio {ratio:.5f}, relMSE {relmse:.2e}", flush=True)


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


# ---------------------------------------------------------------- sky atmosphere
# The `sky atmosphere` line (scene.h:574-588, 1219-1231): a single-scattering
# sky (atmosphere.h) baked on the host into a 256x128 environment lit by the
# sun's settings, and the sun's irradiance dimmed by the air's transmittance.
# The GPU path renders the bake with its environment kernels unchanged.

SKY = "sky_atmosphere.scene"
SKY_GOLD = os.path.join(ROOT, "tests", "golden", "sky_atmosphere.npz")


def _env_lookup(px, dirs):
    """environment_light::radiance (environment.h:66-79), bilinear with the
    azimuth wrapping, rotation 0, intensity 1."""
    h, w = px.shape[:2]
    d = dirs / np.linalg.norm(dirs, axis=1, keepdims=True)
    theta = np.arccos(np.clip(d[:, 1], -1, 1))
    u = 0.5 + np.arctan2(d[:, 0], -d[:, 2]) / (2 * np.pi)
    u -= np.floor(u)
    x = u * w - 0.5
    y = np.clip(theta / np.pi * h - 0.5, 0, h - 1)
    x0, y0 = np.floor(x), np.floor(y)
    fx, fy = (x - x0)[:, None], (y - y0)[:, None]
    i0 = (x0.astype(int) + w) % w
    i1 = (i0 + 1) % w
    j0 = y0.astype(int)
    j1 = np.minimum(j0 + 1, h - 1)
    top = (1 - fx) * px[j0, i0] + fx * px[j0, i1]
    bottom = (1 - fx) * px[j1, i0] + fx * px[j1, i1]
    return (1 - fy) * top + fy * bottom


def _baked(desc):
    w, h = desc.env_width, desc.env_height
    return np.ctypeslib.as_array(desc.env_rgb, (w * h * 3,)).reshape(h, w, 3).copy()


def test_sky_atmosphere_bake_matches_reference():
    import mypbr_amd as m
    g = np.load(SKY_GOLD)
    d = m.describe(scene_path(SKY))
    assert (d.desc.env_width, d.desc.env_height) == (256, 128) and d.desc.has_sun == 1
    assert [d.desc.light_object[k] for k in range(d.desc.n_lights)] == [-1, -2]
    got = _env_lookup(_baked(d.desc), g["sky_dirs"])
    want = g["sky_radiance"]
    assert np.all(want >= 0) and want.max() > 0
    assert np.allclose(got, want, rtol=1e-12, atol=1e-300)


def test_sky_atmosphere_light_probabilities_match_reference():
    """Power-based selection sums the whole bake (luminance times sin theta)
    and the dimmed sun's luminance: equal bits mean equal checksums."""
    import mypbr_amd as m
    g = np.load(SKY_GOLD)
    d = m.describe(scene_path(SKY))
    assert np.array_equal(ol.OracleScene(d.desc).light_probabilities(d.desc.n_lights), g["light_probabilities"])


def test_sky_atmosphere_closed_forms(tmp_path):
    """Independent of the reference: without air or haze the sky is black and
    the sun keeps its irradiance; with ozone alone (which absorbs and does not
    scatter) the sky stays black and an overhead sun is dimmed by
    exp(-beta_ozone * ozone * 15 km), the tent's vertical integral. The
    march takes 16 midpoint steps of 6.2 km, which cut the tent's kinks and
    fall about 3% short of the exponent; a lost factor or unit is far off."""
    import mypbr_amd as m
    base = "image_width 8\nsun elevation 90 azimuth 0 intensity 2\nsphere 0 0 -1 0.5\n"
    s = tmp_path / "clear.scene"
    s.write_text(base + "sky atmosphere rayleigh 0 mie 0 ozone 0\n")
    d = m.describe(str(s))
    assert np.all(_baked(d.desc) == 0) and list(d.desc.sun_irradiance) == [2, 2, 2]
    s.write_text(base + "sky atmosphere rayleigh 0 mie 0 ozone 60\n")
    d = m.describe(str(s))
    assert np.all(_baked(d.desc) == 0)
    tau = np.array([0.650e-6, 1.881e-6, 0.085e-6]) * 60 * 15000
    got = -np.log(np.array(d.desc.sun_irradiance) / 2)
    assert np.allclose(got, tau, rtol=0.05), (got, tau)


@pytest.mark.parametrize("line,ok", [
    ("sky atmosphere", True), ("sky atmosphere rayleigh 2 mie 0.5 mie_g 0.7 ozone 0 intensity 1", True),
    ("sky atmosphere mie_g 1", False), ("sky atmosphere rayleigh -1", False), ("sky atmosphere fog 1", False),
    ("sky nosuch", False),
])
def test_sky_lines(tmp_path, line, ok):
    import mypbr_amd as m
    s = tmp_path / "s.scene"
    s.write_text(f"image_width 8\n{line}\nsphere 0 0 -1 0.5\n")
    if ok:
        d = m.describe(str(s))
        assert (d.desc.env_width, d.desc.env_height) == (256, 128) and d.desc.has_sun == 0
    else:
        with pytest.raises(m.MptError):
            m.describe(str(s))


def test_oracle_sky_render_converges_to_reference():
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_oracle_golden import render_agrees
    g = np.load(SKY_GOLD)
    d = m.describe(scene_path(SKY))
    cam = camera_copy(d.camera, image_width=int(g["width"]), samples_per_pixel=int(g["spp"]) // 8,
                      max_depth=int(g["max_depth"]))
    o = ol.OracleScene(d.desc)
    runs = np.stack([o.render(cam, seed=100 + k)["radiance"] for k in range(8)])
    render_agrees(runs.mean(0), runs.std(0, ddof=1) / np.sqrt(8), g, SKY)


@pytest.mark.gpu
def test_gpu_sky_render_matches_oracle():
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_gpu_render import _compare
    d = m.describe(scene_path(SKY))
    cam = camera_copy(d.camera, image_width=64, samples_per_pixel=8, max_depth=12)
    cpu = ol.OracleScene(d.desc).render(cam, seed=7)
    f = m.Scene.from_desc(d.desc, d.camera).render(cam, seed=7)
    _compare(f.radiance, cpu["radiance"], 0.995, 1e-3)


@pytest.mark.gpu
def test_gpu_sky_render_converges_to_reference():
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_oracle_golden import render_agrees
    g = np.load(SKY_GOLD)
    d = m.describe(scene_path(SKY))
    cam = camera_copy(d.camera, image_width=int(g["width"]), samples_per_pixel=int(g["spp"]) // 8,
                      max_depth=int(g["max_depth"]))
    s = m.Scene.from_desc(d.desc, d.camera)
    runs = np.stack([s.render(cam, seed=300 + k).radiance for k in range(8)])
    ratio, relmse = render_agrees(runs.mean(0), runs.std(0, ddof=1) / np.sqrt(8), g, SKY)
    print(f"{SKY}: total ratio {ratio:.5f}, relMSE {relmse:.2e}", flush=True)
