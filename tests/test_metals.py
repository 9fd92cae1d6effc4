"""Measured metals (metal with a named conductor, principled's conductor:
fresnel_conductor and metal_fresnel::measured, microfacet.h:57-140;
material.h:117-121, 381-385, 537-551) and thin glass (thin_dielectric,
material.h:683-707).

CPU: the oracle's measured-metal BRDFs against the reference's (bsdf.npz
holds gold, aluminium and a principled layer over chromium, checked in
test_oracle_golden); scenes/metals.scene's hits and converged means
against tests/golden/metals.npz, written by the compiled reference. GPU: the
same hits and means, and the oracle at equal seed. The device computes the
complex Fresnel with its own complex arithmetic (hypot inside the square
root), which can differ from the C++ library's in the last bits; the image
comparisons carry the tolerance the other scenes do."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_lib as ol
from conftest import ROOT, scene_path

GOLD = os.path.join(ROOT, "tests", "golden", "metals.npz")
SCENE = "metals.scene"


def _golden():
    g = np.load(GOLD)
    return g, json.loads(str(g["digest"]))


def test_conductor_table():
    import mypbr_amd as m
    from mypbr_amd import lib
    import ctypes as C
    eta, k = (C.c_double * 3)(), (C.c_double * 3)()
    for name in ("Al", "aluminium", "aluminum", "Ti", "Cu", "copper", "Au", "gold", "Ag", "silver", "Cr", "chrome"):
        assert lib.mpt_find_conductor(name.encode(), eta, k) == 0, name
        assert all(0 < x < 10 for x in list(eta) + list(k))
    assert lib.mpt_find_conductor(b"Pt", eta, k) != 0
    del m


def test_measured_metal_parameters_reach_the_description():
    import mypbr_amd as m
    d = m.describe(scene_path(SCENE))
    mats = [d.desc.materials[i] for i in range(d.desc.n_materials)]
    measured = [x for x in mats if x.measured]
    assert len(measured) == 4 and sum(x.thin for x in mats) == 2
    gold = [x for x in measured if x.type == 1 and x.roughness == 0][0]
    assert abs(gold.eta[0] - 0.143036) < 1e-15 and abs(gold.k[2] - 1.60336) < 1e-15


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


@pytest.mark.parametrize("line", [
    "material a metal Pt 0.2",              # unknown metal
    "material a metal Au",                  # no roughness
    "material a principled conductor Xx",
    "material a thin_glass 0",
    "material a thin_glass 1.5 color 2 1 1",
    "material a thin_glass 1.5 tint 1 1 1",
])
def test_bad_lines(tmp_path, line):
    import mypbr_amd as m
    s = tmp_path / "s.scene"
    s.write_text(f"image_width 8\n{line}\nsphere 0 0 -1 0.5\n")
    with pytest.raises(m.MptError):
        m.describe(str(s))


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


@pytest.mark.gpu
def test_gpu_render_matches_oracle():
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_gpu_render import _compare
    d = m.describe(scene_path(SCENE))
    cam = camera_copy(d.camera, image_width=64, samples_per_pixel=8, max_depth=16, collect_features=1)
    cpu = ol.OracleScene(d.desc).render(cam, seed=7)
    f = m.Scene.from_desc(d.desc, d.camera).render(cam, seed=7)
    _compare(f.radiance, cpu["radiance"], 0.995, 1e-3)
    _compare(f.albedo, cpu["albedo"], 0.995, 1e-3)


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
