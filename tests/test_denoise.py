"""The classical a-trous denoiser (denoiser::run, denoiser.h:45-193).

CPU: the oracle's restatement against the reference's own outputs
(tests/golden/denoise.npz, written by the compiled reference), and against
the known answers of the reference's denoiser test (denoise_test.cpp:50-125)
and closed forms. GPU: the HIP kernels against the oracle, to 1e-12
relative (the device exp() may differ from the host's in the last bit), on
the same frames, on a rendered frame, and on the known answers."""
import os

import numpy as np
import pytest

import denoise_frames as dfr
import oracle_lib as ol
from conftest import ROOT, scene_path

GOLDEN = os.path.join(ROOT, "tests", "golden", "denoise.npz")


def _golden():
    g = np.load(GOLDEN)
    frames = {}
    for name in dfr.CASES:
        frames[name] = {k: g[f"{name}/{k}"] for k in ("radiance", "albedo", "normal", "variance")}
        frames[name]["out"] = [g[f"{name}/out{i}"] for i in range(len(dfr.SETTINGS))]
    return frames, [tuple(x) for x in g["settings"]]


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name", list(dfr.CASES))
def test_oracle_matches_reference_outputs(name):
    frames, settings = _golden()
    f = frames[name]
    for si, st in enumerate(settings):
        st = (int(st[0]), st[1], st[2], st[3])
        got = ol.oracle_denoise(f, st)
        assert np.array_equal(got, f["out"][si]), (name, st, _rel(got, f["out"][si]))


def _known_answers(run):
    """denoise_test.cpp:50-125, with `run(frame) -> radiance`."""
    f = dfr.flat_wall()
    out = run(f)
    before, after = dfr.rms(f["radiance"], f["truth"]), dfr.rms(out, f["truth"])
    assert after < before / 5, (before, after)
    assert abs(dfr.lum(out).mean() - 0.5) < 0.01
    f, clean = dfr.albedo_edge(), dfr.make(0.0, 2, lambda i, j: ((0.2,) * 3 if i < 32 else (0.8,) * 3, (0, 0, 1),
                                                                 (0.2,) * 3 if i < 32 else (0.8,) * 3))
    out = run(f)
    worst = np.abs(dfr.lum(out[:, 31:33]) - dfr.lum(clean["radiance"][:, 31:33])).max()
    assert worst < 0.06, worst
    out = run(dfr.normal_edge())
    assert abs(dfr.lum(out[dfr.H // 2, dfr.W // 2]) - 0.1) < 0.03
    out = run(dfr.firefly())
    assert dfr.lum(out[dfr.FIREFLY]) < 1
    assert dfr.lum(out[dfr.HIGHLIGHT]) > 25


def test_oracle_known_answers():
    _known_answers(ol.oracle_denoise)


def test_oracle_closed_forms():
    # a noise-free uniform frame is a fixed point: every tap has the same
    # lighting, so the weighted mean is that lighting (to rounding)
    f = dfr.make(0.0, 1, lambda i, j: ((0.3, 0.6, 0.9), (0, 1, 0), (0.2, 0.4, 0.8)))
    out = ol.oracle_denoise(f)
    assert np.allclose(out, f["radiance"], rtol=1e-14, atol=0)
    # zero iterations: only the albedo division and multiplication (and no firefly here)
    g = dfr.mixed()
    g["variance"] = np.zeros_like(g["variance"])
    out = ol.oracle_denoise(g, (0, 4.0, 0.1, 0.1))
    assert np.allclose(out, g["radiance"], rtol=1e-14, atol=1e-300)
    # a planted mistake is caught: a dropped albedo guide changes the output
    bad = dict(g, albedo=np.full_like(g["albedo"], 0.5))
    frames, _ = _golden()
    assert not np.allclose(ol.oracle_denoise(dict(frames["mixed"], albedo=bad["albedo"][:40, :48])),
                           frames["mixed"]["out"][0])


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(dfr.CASES))
def test_gpu_matches_oracle(name):
    import mypbr_amd as m
    frames, settings = _golden()
    f = frames[name]
    for si, st in enumerate(settings):
        it, sl, sn, sa = int(st[0]), st[1], st[2], st[3]
        got = m.denoise(f, iterations=it, sigma_luminance=sl, sigma_normal=sn, sigma_albedo=sa)
        rel = _rel(got, f["out"][si])
        print(name, st, f"max rel {rel:.2e}", flush=True)
        assert rel < 1e-12


@pytest.mark.gpu
def test_gpu_known_answers():
    import mypbr_amd as m
    _known_answers(lambda f: m.denoise(f))


@pytest.mark.gpu
def test_gpu_denoises_a_render_like_the_oracle():
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    d = m.describe(scene_path("weekend_cover.scene"))
    cam = camera_copy(d.camera, image_width=192, samples_per_pixel=8, max_depth=20, collect_features=1)
    fr = m.Scene.from_desc(d.desc, d.camera).render(cam, seed=5)
    frame = {"radiance": fr.radiance, "albedo": fr.albedo, "normal": fr.normal, "variance": fr.variance}
    got = m.denoise(frame)
    want = ol.oracle_denoise(frame)
    assert _rel(got, want) < 1e-12
    # and it removes noise: the filtered image is smoother than the input
    def rough(x):
        return np.abs(np.diff(x, axis=1)).mean()
    assert rough(got) < 0.7 * rough(fr.radiance)
