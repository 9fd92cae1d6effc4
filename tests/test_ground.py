"""The infinite ground plane (ground_plane, plane.h:10-42; the `ground`
line, scene.h:596-611, 1237-1238): tested beside the BVH after it with the
closest t so far (main.cpp:70-74), (u, v) = (x, -z) with constant tangents.

CPU: the oracle against tests/golden/ground.npz, written by the compiled
reference: the surface record at each hit of the subset (the ground's (u, v),
tangents, checker value and normal-mapped normal) bit for bit, 1M-ray hit
digests (the ground's canonical ID is one past the world's), the BVH-plus-
ground against a linear scan, converged means. GPU: hits and converged means
against the same golden, and the oracle at equal seed."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_lib as ol
from conftest import ROOT, scene_path

GOLD = os.path.join(ROOT, "tests", "golden", "ground.npz")
SCENE = "ground.scene"


def _golden():
    g = np.load(GOLD)
    return g, json.loads(str(g["digest"]))


def test_ground_line():
    import mypbr_amd as m
    d = m.describe(scene_path(SCENE))
    assert d.desc.has_ground == 1 and d.desc.ground_height == 0.2
    assert d.desc.materials[d.desc.ground_material].normal_map >= 0


@pytest.mark.parametrize("line,ok", [
    ("ground", True), ("ground height -1", True), ("ground off", True), ("ground height", False),
    ("ground nosuch", False),
])
def test_ground_lines(tmp_path, line, ok):
    import mypbr_amd as m
    s = tmp_path / "s.scene"
    s.write_text(f"image_width 8\n{line}\nsphere 0 0 -1 0.5\n")
    if ok:
        d = m.describe(str(s))
        assert d.desc.has_ground == (0 if "off" in line else 1)
    else:
        with pytest.raises(m.MptError):
            m.describe(str(s))


def test_oracle_surface_matches_reference():
    import mypbr_amd as m
    g, _ = _golden()
    d = m.describe(scene_path(SCENE))
    got = ol.OracleScene(d.desc).surface(g["rays"], g["tt"])
    want = g["surface"]
    assert np.array_equal(np.isnan(got), np.isnan(want))
    hit = ~np.isnan(want[:, 0])
    assert np.array_equal(got[hit].view(np.int64), want[hit].view(np.int64))


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
    ground = int(max(d.desc.object_id_base[k] + d.desc.object_count[k] for k in range(d.desc.n_objects)))
    assert (ids == ground).sum() > len(ids) // 10  # the ground takes a good part of the batch
    lid, lt = o.trace_linear(rays[:20000], tt[:20000])
    assert np.array_equal(lid, ids[:20000]) and np.array_equal(lt, t[:20000])


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
def test_gpu_render_matches_oracle():
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_gpu_render import _compare
    d = m.describe(scene_path(SCENE))
    cam = camera_copy(d.camera, image_width=64, samples_per_pixel=8, max_depth=12, collect_features=1)
    cpu = ol.OracleScene(d.desc).render(cam, seed=7)
    f = m.Scene.from_desc(d.desc, d.camera).render(cam, seed=7)
    _compare(f.radiance, cpu["radiance"], 0.995, 1e-3)
    _compare(f.albedo, cpu["albedo"], 0.995, 1e-3)
    _compare(f.normal, cpu["normal"], 0.995, 1e-3)


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
