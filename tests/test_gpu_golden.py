"""The HIP path against the reference's own golden vectors and closed forms.

 - ray queries: the seeded 1M-ray batches' IDs, t and front_face through the
   GPU BVH hash to the SHA-256 the compiled reference produced;
 - materials: eval and scatter-pdf values at fixed directions agree with the
   reference to 1e-12 relative;
 - renders: low-resolution converged means match the reference's
   (validate.py's rules, as in the oracle tests);
 - furnace: white furnace, bounded energy, Kulla-Conty energy and
   Beer-Lambert slabs (furnace_test.cpp) on the GPU.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import scene_path
from furnace_cases import BOUNDED, ENERGY, FURNACE, SLABS, bounded_check, object_energy, run_furnace, slab_worst

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TRACE_SCENES = ["cornell.scene", "weekend_cover.scene", "test_lights.scene"]
RENDER_SCENES = ["cornell.scene", "weekend_cover.scene", "test_lights.scene", "test_media.scene"]


@pytest.fixture(scope="module")
def mpt():
    import mypbr_amd
    return mypbr_amd


def gpu_render(b, cam):
    import mypbr_amd
    desc, keep = b.build()
    s = mypbr_amd.Scene.from_desc(desc, cam)
    img = s.render(cam, seed=1).radiance
    s.close()
    del keep
    return img


@pytest.mark.parametrize("name", TRACE_SCENES)
def test_gpu_trace_1m_digest_matches_reference(mpt, name):
    from oracle_lib import OracleScene
    with open(os.path.join(GOLD, "trace_digests.json")) as f:
        dg = json.load(f)[name]
    d = mpt.describe(scene_path(name))
    rays, tt = OracleScene(d.desc).ray_batch(d.camera, dg["seed"], dg["n"])  # the seeded input batch
    s = mpt.Scene.from_desc(d.desc, d.camera)
    ids, t, f = s.trace_closest(rays, tt)
    print(f"{name}: misses {(ids < 0).sum()}", flush=True)
    assert hashlib.sha256(ids.astype("<i8").tobytes()).hexdigest() == dg["sha256_id"]
    assert hashlib.sha256(t.astype("<f8").tobytes()).hexdigest() == dg["sha256_t"]
    assert hashlib.sha256(f.astype("u1").tobytes()).hexdigest() == dg["sha256_front"]


def test_gpu_bsdf_matches_reference(mpt):
    from golden_materials import desc_for_specs
    g = np.load(os.path.join(GOLD, "bsdf.npz"))
    desc, keep, mats = desc_for_specs(g["spec"])
    s = mpt.Scene.from_desc(desc)
    ev, pdf = s.bsdf_eval(mats, g["q"], g["front"])
    ok_e = np.abs(ev - g["eval"]) <= 1e-12 * np.maximum(np.abs(g["eval"]), 1e-300)
    ok_p = np.abs(pdf - g["pdf"]) <= 1e-12 * np.maximum(np.abs(g["pdf"]), 1e-300)
    assert ok_e.all(), f"{(~ok_e).sum()} eval values differ"
    assert ok_p.all(), f"{(~ok_p).sum()} pdf values differ"


@pytest.mark.parametrize("name", RENDER_SCENES)
def test_gpu_render_converges_to_reference(mpt, name):
    from mypbr_amd.abi import camera_copy
    from test_oracle_golden import render_agrees
    g = np.load(os.path.join(GOLD, "render_" + name.replace(".scene", ".npz")))
    d = mpt.describe(scene_path(name))
    spp = int(g["spp"]) // 8
    cam = camera_copy(d.camera, image_width=int(g["width"]), samples_per_pixel=spp, max_depth=int(g["max_depth"]))
    s = mpt.Scene.from_desc(d.desc, d.camera)
    runs = np.stack([s.render(cam, seed=300 + k).radiance for k in range(8)])
    img = runs.mean(0)
    ratio, relmse = render_agrees(img, runs.std(0, ddof=1) / np.sqrt(8), g, name)
    print(f"{name}: total ratio {ratio:.5f}, relMSE {relmse:.2e}", flush=True)


@pytest.mark.parametrize("name,build,check", FURNACE, ids=[c[0] for c in FURNACE])
def test_gpu_furnace(name, build, check):
    ok, msg = check(run_furnace(gpu_render, build))
    assert ok, f"{name}: {msg}"


@pytest.mark.parametrize("name,build", BOUNDED, ids=[c[0] for c in BOUNDED])
def test_gpu_bounded(name, build):
    ok, msg = bounded_check(run_furnace(gpu_render, build))
    assert ok, f"{name}: {msg}"


@pytest.mark.parametrize("name,make", ENERGY, ids=[c[0] for c in ENERGY])
def test_gpu_energy_conserving(name, make):
    e = object_energy(gpu_render, make)
    assert abs(e - 1) < 0.02, f"{name}: sphere pixels {e}"


@pytest.mark.parametrize("name,build,expect", SLABS, ids=[c[0] for c in SLABS])
def test_gpu_slab(name, build, expect):
    assert slab_worst(gpu_render, build, expect) < 1e-5
