"""Grazing ties and interval bounds against the compiled reference.

tests/golden/trace_ties.npz holds, per scene, rays aimed exactly at points
where primitives meet or end (quad corners and edges, a sphere's lowest
point on its floor, instanced box corners, from dyadic origins so that the
axis-aligned planes give bitwise-equal t), and rays of the seeded batch
whose interval ends (tmax = t) or starts (tmin = t) exactly at the
reference's hit: quads accept t == tmax (quad.h:39), spheres and triangles
do not (sphere.h:37-40, mesh.h:63), and the BVH culls boxes it only touches
(aabb.h:50-51). For every ray the reference lists the candidates, the
primitives whose own hit() over the interval gives the closest t bitwise:
their count and smallest and largest canonical ID. The oracle and the GPU
must list the same candidates, return the same t and front face, and name a
candidate as the hit; the GPU's choice among tied candidates must be the
oracle's (the BVH walk order is the same)."""
import os

import numpy as np
import pytest

from conftest import scene_path

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "trace_ties.npz")
SCENES = ["cornell", "test_lights", "weekend_cover"]


def _golden(key):
    g = np.load(GOLDEN)
    return {k[len(key) + 1:]: g[k] for k in g.files if k.startswith(key + "_")}


def _check(out, g):
    assert np.array_equal(out["ties"], g["ties"]), "candidate counts differ"
    assert np.array_equal(out["tie_min"], g["tie_min"]) and np.array_equal(out["tie_max"], g["tie_max"])
    assert np.array_equal(out["t"].view(np.int64), g["t"].view(np.int64)), "t differs"
    hit = g["ties"] > 0
    assert np.array_equal(out["id"] >= 0, hit)
    assert np.all((out["id"][hit] >= g["tie_min"][hit]) & (out["id"][hit] <= g["tie_max"][hit]))
    one = g["ties"] == 1
    assert np.array_equal(out["id"][one], g["id"][one])
    assert np.array_equal(out["front"][one], g["front"][one])


def test_golden_batches_are_tie_rich():
    for key in SCENES:
        g = _golden(key)
        assert np.all(g["ties"] >= 0) and np.all(g["id"] != -2), "the reference left a hit unidentified"
        for kind in (1, 2):  # interval-bound rays all hit something first
            assert np.sum(g["kind"] == kind) >= 256
    assert sum(int(np.sum(_golden(k)["ties"] > 1)) for k in SCENES) >= 200
    # cornell's walls meet on axis-aligned edges: aimed rays there are ties
    c = _golden("cornell")
    assert np.sum((c["kind"] == 0) & (c["ties"] > 1)) >= 100


@pytest.mark.parametrize("key", SCENES)
def test_oracle_lists_the_reference_ties(key):
    import mypbr_amd as m
    from oracle_lib import OracleScene
    g = _golden(key)
    d = m.describe(scene_path(key + ".scene"))
    o = OracleScene(d.desc)
    out = o.trace_ties(g["rays"], g["tt"])
    out["front"] = o.trace(g["rays"], g["tt"])[2]
    _check(out, g)


@pytest.mark.gpu
@pytest.mark.parametrize("key", SCENES)
def test_gpu_lists_the_reference_ties(key):
    import mypbr_amd as m
    from oracle_lib import OracleScene
    g = _golden(key)
    d = m.describe(scene_path(key + ".scene"))
    s = m.Scene.from_desc(d.desc, d.camera)
    out = s.trace_ties(g["rays"], g["tt"])
    _check(out, g)
    orc = OracleScene(d.desc).trace_ties(g["rays"], g["tt"])
    assert np.array_equal(out["id"], orc["id"]), "the GPU took another candidate than the oracle's BVH walk"
    ids, t, front = s.trace_closest(g["rays"], g["tt"])
    assert np.array_equal(ids, out["id"]) and np.array_equal(t.view(np.int64), out["t"].view(np.int64))
    assert np.array_equal(front, out["front"])
    print(key, "rays", len(ids), "ties", int(np.sum(out["ties"] > 1)), flush=True)
