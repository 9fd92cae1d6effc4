"""The CPU oracle against the golden vectors the compiled reference wrote
(tests/golden/make_golden.py), and against brute force.

These pin the oracle: the GPU parity tests then compare the HIP path with
an oracle that is itself shown equal to the reference on the same inputs.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import mypbr_amd as m
from conftest import scene_path
from oracle_lib import OracleScene, oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TRACE_SCENES = ["cornell.scene", "weekend_cover.scene", "test_lights.scene"]


def gold(name):
    return np.load(os.path.join(GOLD, name))


@pytest.fixture(scope="module")
def digests():
    with open(os.path.join(GOLD, "trace_digests.json")) as f:
        return json.load(f)


def test_ggx_table_bit_exact():
    g = gold("ggx_table.npz")
    e = np.zeros(1024)
    ea = np.zeros(32)
    import oracle_lib as ol
    oracle().orc_ggx_table(ol._d(e), ol._d(ea))
    assert np.array_equal(e.reshape(32, 32), g["e"]) and np.array_equal(ea, g["e_avg"])
    pe, pea = m.ggx_albedo_table()  # the product's host-side table
    assert np.array_equal(pe, g["e"]) and np.array_equal(pea, g["e_avg"])


def test_ggx_table_closed_forms():
    """E <= 1, E(mu, alpha -> 0) -> 1 (a smooth mirror keeps everything), E
    falls with roughness at grazing views, E_avg = 2 int E mu dmu."""
    g = gold("ggx_table.npz")
    e, ea = g["e"], g["e_avg"]
    assert np.all(e <= 1) and np.all(e > 0)
    assert np.all(e[0, 4:] > 0.999)
    assert e[-1, 3] < e[8, 3] < e[0, 3]
    mu = np.linspace(0, 1, 32)
    for j in (0, 10, 31):
        integral = np.trapezoid(e[j] * mu, mu)
        assert abs(min(2 * integral, 1.0) - ea[j]) < 2e-3


@pytest.mark.parametrize("name", TRACE_SCENES)
def test_trace_subset_matches_reference(name):
    g = gold("trace_" + name.replace(".scene", ".npz"))
    d = m.describe(scene_path(name))
    o = OracleScene(d.desc)
    ids, t, f = o.trace(g["rays"], g["tt"])
    unique = g["ties"] <= 1
    assert np.array_equal(ids[unique], g["id"][unique])
    assert np.array_equal(t.view(np.int64), g["t"].view(np.int64))
    assert np.array_equal(f, g["front"])


@pytest.mark.parametrize("name", TRACE_SCENES)
def test_trace_1m_digest_matches_reference(name, digests):
    dg = digests[name]
    d = m.describe(scene_path(name))
    o = OracleScene(d.desc)
    rays, tt = o.ray_batch(d.camera, dg["seed"], dg["n"])
    g = gold("trace_" + name.replace(".scene", ".npz"))
    assert np.array_equal(rays[g["index"]], g["rays"]), "the seeded batch generator changed"
    ids, t, f = o.trace(rays, tt)
    assert dg["ties"] == 0 and dg["unidentified"] == 0
    assert hashlib.sha256(ids.astype("<i8").tobytes()).hexdigest() == dg["sha256_id"]
    assert hashlib.sha256(t.astype("<f8").tobytes()).hexdigest() == dg["sha256_t"]
    assert hashlib.sha256(f.astype("u1").tobytes()).hexdigest() == dg["sha256_front"]
    assert int((ids == -1).sum()) == dg["misses"]


@pytest.mark.parametrize("name", TRACE_SCENES)
def test_bvh_equals_brute_force(name):
    """Closest hit through the binned-SAH BVH == linear scan over all objects."""
    d = m.describe(scene_path(name))
    o = OracleScene(d.desc)
    rays, tt = o.ray_batch(d.camera, 11, 20000)
    ids, t, _ = o.trace(rays, tt)
    lid, lt = o.trace_linear(rays, tt)
    assert np.array_equal(ids, lid) and np.array_equal(t, lt)


def test_light_probabilities_match_reference():
    with open(os.path.join(GOLD, "light_probs.json")) as f:
        probs = json.load(f)
    for name, expect in probs.items():
        d = m.describe(scene_path(name))
        got = OracleScene(d.desc).light_probabilities(d.desc.n_lights)
        assert np.array_equal(got, np.array(expect)), name


def test_light_probabilities_closed_form():
    """pdf_test.cpp:178-215: p_k = 0.8 P_k / sum P + 0.2 / n, sum p = 1; a
    light of unknown (zero) power gets the mean of the others."""
    from scenebuild import DescBuilder
    b = DescBuilder()
    a = b.diffuse_light((1, 1, 1))
    c = b.diffuse_light((2, 2, 2))
    dark = b.diffuse_light((0, 0, 0))
    wall = b.lambertian((0.5, 0.5, 0.5))
    q1 = b.quad((0, 0, 0), (1, 0, 0), (0, 1, 0), a)        # power pi * 1 * 1
    q2 = b.quad((0, 0, 2), (2, 0, 0), (0, 1, 0), c)        # power pi * 2 * 2
    q3 = b.quad((0, 0, 4), (1, 0, 0), (0, 1, 0), dark)     # unknown: mean of the others
    b.quad((0, -1, 0), (5, 0, 0), (0, 0, 5), wall)
    for q in (q1, q2, q3):
        b.light(q)
    desc, keep = b.build()
    p = OracleScene(desc).light_probabilities(3)
    powers = np.array([np.pi, 4 * np.pi, 2.5 * np.pi])
    expect = 0.8 * powers / powers.sum() + 0.2 / 3
    assert np.allclose(p, expect, rtol=1e-12, atol=0)
    assert abs(p.sum() - 1) < 1e-12


def test_bsdf_matches_reference():
    g = gold("bsdf.npz")
    spec, q, front = g["spec"], g["q"], g["front"]
    from golden_materials import desc_for_specs
    desc, keep, mats = desc_for_specs(spec)
    o = OracleScene(desc)
    ev, pdf = o.bsdf_eval(mats, q, front)
    ok_e = np.abs(ev - g["eval"]) <= 1e-12 * np.maximum(np.abs(g["eval"]), 1e-300)
    ok_p = np.abs(pdf - g["pdf"]) <= 1e-12 * np.maximum(np.abs(g["pdf"]), 1e-300)
    assert ok_e.all(), f"{(~ok_e).sum()} eval values differ"
    assert ok_p.all(), f"{(~ok_p).sum()} pdf values differ"
    assert (g["pdf"] > 0).sum() > 1000  # the queries exercise the lobes


RENDER_SCENES = ["cornell.scene", "weekend_cover.scene", "test_lights.scene", "test_media.scene"]


def render_agrees(img, img_se, g, label):
    """validate.py's judging rules (validate.py:116-153), on low-res images:
    total within 1% or not significant at 4 sigma, no 4x4 block off by more
    than 3% and 4.5 sigma, and a relMSE bound (eps 0.01). Sigma comes from
    8 runs a side, so block z-scores are rescaled by their robust spread
    (median absolute deviation) when the per-pixel errors understate it:
    path-traced pixel means have heavy tails at these sample counts."""
    ref, ref_se = g["mean"], g["stderr"]
    se = np.sqrt(ref_se ** 2 + img_se ** 2)
    tot, tot_ref = img.sum(), ref.sum()
    tot_se = np.sqrt((se ** 2).sum())
    ratio = tot / tot_ref
    assert abs(ratio - 1) < 0.01 or abs(tot - tot_ref) < 4 * tot_se, f"{label}: total ratio {ratio}"
    h, w = ref.shape[:2]
    blocks = []
    for by in range(0, h - 3, 4):
        for bx in range(0, w - 3, 4):
            a = img[by:by + 4, bx:bx + 4].sum()
            b = ref[by:by + 4, bx:bx + 4].sum()
            s = np.sqrt((se[by:by + 4, bx:bx + 4] ** 2).sum())
            blocks.append((a, b, (a - b) / s))
    z = np.array([x[2] for x in blocks])
    scale = max(1.0, 1.4826 * np.median(np.abs(z)))
    bad = [(a, b, zz) for a, b, zz in blocks if abs(a - b) > 0.03 * max(b, 1e-6) and abs(zz) > 4.5 * scale]
    assert not bad, f"{label}: bad blocks {bad} (z scale {scale:.2f})"
    relmse = np.mean((img - ref) ** 2 / (ref ** 2 + 0.01))
    noise = np.mean(se ** 2 / (ref ** 2 + 0.01))
    assert relmse < 3 * noise + 1e-6, f"{label}: relMSE {relmse:.3e} vs noise {noise:.3e}"
    return ratio, relmse


@pytest.mark.parametrize("name", RENDER_SCENES)
def test_render_converges_to_reference(name):
    from mypbr_amd.abi import camera_copy
    g = gold("render_" + name.replace(".scene", ".npz"))
    d = m.describe(scene_path(name))
    spp = int(g["spp"]) // 8
    cam = camera_copy(d.camera, image_width=int(g["width"]), samples_per_pixel=spp, max_depth=int(g["max_depth"]))
    o = OracleScene(d.desc)
    runs = np.stack([o.render(cam, seed=100 + k)["radiance"] for k in range(8)])
    img = runs.mean(0)
    img_se = runs.std(0, ddof=1) / np.sqrt(8)
    render_agrees(img, img_se, g, name)
