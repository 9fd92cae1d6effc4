"""Photo studio pieces (studio.h; the `cyclorama` and `softbox` lines,
scene.h:935-1001): a backdrop mesh that curves into the floor with smooth
normals and arc-length uv, and quad lights aimed at a point.

CPU: the geometry against the settings' closed forms (profile, normals,
seams, uv, a softbox's corner, edges and facing, the straight-down case);
the lines' errors; the oracle against the compiled reference on its own
generators (tests/golden/studio.npz: hit t, front face and surface records,
converged means). GPU: a render at equal seed against the oracle and the
converged means against the reference's."""
import math
import os

import numpy as np
import pytest

import oracle_lib as ol
from conftest import ROOT, scene_path

GOLD = os.path.join(ROOT, "tests", "golden", "studio.npz")


def _scene(tmp_path, body):
    s = tmp_path / "s.scene"
    s.write_text("image_width 8\nmaterial w lambertian 0.5 0.5 0.5\nmaterial e emissive 4 4 4\n" + body)
    return str(s)


def test_cyclorama_profile(tmp_path):
    """The profile from the floor's front over the quarter circle (centre y =
    radius, z = -back) to the wall's top, its normals pointing at the
    centre, and uv: x across, arc length along."""
    import mypbr_amd as m
    from mypbr_amd.abi import PRIM_TRIANGLE
    d = m.describe(_scene(tmp_path, "cyclorama width 4 front 2 back 1 height 5 radius 2 segments 4 w\n"))
    desc = d.desc
    tri = [k for k in range(desc.n_prims) if desc.prim_kind[k] == PRIM_TRIANGLE]
    assert len(tri) == 2 * (1 + 4 + 1)  # floor strip, 4 curve strips, the wall
    pts = np.array([desc.prim_data[9 * k + j] for k in tri for j in range(9)]).reshape(-1, 3)
    r, back = 2.0, 1.0
    th = np.arange(5) * (math.pi / 2) / 4
    profile = [(0.0, 2.0)] + [(r * (1 - math.cos(t)), -back - r * math.sin(t)) for t in th] + [(5.0, -3.0)]
    for y, z in profile:
        assert np.any(np.all(np.isclose(pts, [2.0, y, z], atol=1e-12), axis=1)), (y, z)
    assert np.all(np.isin(np.round(pts[:, 0], 12), [-2.0, 2.0]))
    nrm = np.array([desc.prim_normals[9 * k + j] for k in tri for j in range(9)]).reshape(-1, 3)
    for p, n in zip(pts, nrm):
        if p[2] < -back and p[1] < r:  # on the curve: towards the centre
            c = np.array([p[0], r, -back])
            want = (c - p) / np.linalg.norm(c - p)
            assert np.allclose(n, want, atol=1e-12)
        assert abs(np.linalg.norm(n) - 1) < 1e-12 and n[0] == 0
    uv = np.array([desc.prim_uvs[6 * k + j] for k in tri for j in range(6)]).reshape(-1, 2)
    assert uv[:, 1].min() == 0 and uv[:, 1].max() == 1 and set(np.unique(uv[:, 0])) == {0.0, 1.0}


def test_softbox_faces_its_aim(tmp_path):
    import mypbr_amd as m
    from mypbr_amd.abi import PRIM_QUAD
    for line, aim, w, h in (("softbox center -4 5 4 aim 0 1 0 size 3 5 roll 10 e", (0, 1, 0), 3, 5),
                            ("softbox center 0 8 0 aim 0 0 0 size 4 2 e", (0, 0, 0), 4, 2)):
        d = m.describe(_scene(tmp_path, line + "\nsphere 0 1 0 1 w\n"))
        desc = d.desc
        q = [k for k in range(desc.n_prims) if desc.prim_kind[k] == PRIM_QUAD][0]
        Q = np.array([desc.prim_data[9 * q + j] for j in range(3)])
        u = np.array([desc.prim_data[9 * q + 3 + j] for j in range(3)])
        v = np.array([desc.prim_data[9 * q + 6 + j] for j in range(3)])
        centre = Q + u / 2 + v / 2
        c = np.array([float(x) for x in line.split("center")[1].split()[:3]])
        assert np.allclose(centre, c, atol=1e-12)
        assert abs(np.linalg.norm(u) - w) < 1e-12 and abs(np.linalg.norm(v) - h) < 1e-12
        n = np.cross(u, v)
        assert np.dot(n, np.array(aim) - c) > 0  # the front face looks at the aim point
        if "roll" not in line and aim == (0, 0, 0):
            assert np.allclose(u / np.linalg.norm(u), [1, 0, 0], atol=1e-12)  # straight down: width along x
        assert [desc.light_object[k] for k in range(desc.n_lights)]


@pytest.mark.parametrize("line,needle", [
    ("cyclorama front -5 back 4 w", "front must be in front of its back"),
    ("cyclorama width 0 w", "expected 'cyclorama"),
    ("cyclorama e", "cannot be a light"),
    ("softbox center 0 1 0 aim 0 1 0 size 1 1 e", "expected 'softbox"),
    ("softbox center 0 2 0 aim 0 0 0 size 1 1 w", "needs an emissive material"),
    ("softbox aim 0 0 0 size 1 1 e", "expected 'softbox"),
])
def test_studio_line_errors(tmp_path, line, needle):
    import mypbr_amd as m
    with pytest.raises(m.MptError) as e:
        m.describe(_scene(tmp_path, line + "\n"))
    assert needle in str(e.value), str(e.value)


def test_oracle_hits_and_surfaces_match_reference():
    import mypbr_amd as m
    g = np.load(GOLD)
    d = m.describe(scene_path("studio.scene"))
    o = ol.OracleScene(d.desc)
    ids, t, f = o.trace(g["rays"], g["tt"])
    hit = g["id"] != -1
    single = hit & (g["ties"] <= 1)
    assert single.sum() > 1000
    assert np.array_equal(ids != -1, hit)
    assert np.array_equal(t[single].view(np.int64), g["t"][single].view(np.int64))
    assert np.array_equal(f[single], g["front"][single])
    got, want = o.surface(g["rays"], g["tt"]), g["surface"]
    h = single & ~np.isnan(want[:, 0])
    assert np.array_equal(got[h].view(np.int64), want[h].view(np.int64))
    np.testing.assert_array_equal(o.light_probabilities(d.desc.n_lights), g["light_probabilities"])


def test_oracle_render_converges_to_reference():
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_oracle_golden import render_agrees
    g = np.load(GOLD)
    d = m.describe(scene_path("studio.scene"))
    cam = camera_copy(d.camera, image_width=int(g["width"]), samples_per_pixel=int(g["spp"]) // 8,
                      max_depth=int(g["max_depth"]))
    runs = np.stack([ol.OracleScene(d.desc).render(cam, seed=100 + k)["radiance"] for k in range(8)])
    render_agrees(runs.mean(0), runs.std(0, ddof=1) / np.sqrt(8), g, "studio.scene")


@pytest.mark.gpu
def test_gpu_render_matches_oracle():
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_gpu_render import _compare
    d = m.describe(scene_path("studio.scene"))
    cam = camera_copy(d.camera, image_width=64, samples_per_pixel=8, max_depth=8, collect_features=1)
    cpu = ol.OracleScene(d.desc).render(cam, seed=7)
    f = m.Scene.from_desc(d.desc, d.camera).render(cam, seed=7)
    _compare(f.radiance, cpu["radiance"], 0.995, 1e-3)
    _compare(f.albedo, cpu["albedo"], 0.995, 1e-3)


@pytest.mark.gpu
def test_gpu_render_converges_to_reference():
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_oracle_golden import render_agrees
    g = np.load(GOLD)
    d = m.describe(scene_path("studio.scene"))
    cam = camera_copy(d.camera, image_width=int(g["width"]), samples_per_pixel=int(g["spp"]) // 8,
                      max_depth=int(g["max_depth"]))
    s = m.Scene.from_desc(d.desc, d.camera)
    runs = np.stack([s.render(cam, seed=300 + k).radiance for k in range(8)])
    ratio, relmse = render_agrees(runs.mean(0), runs.std(0, ddof=1) / np.sqrt(8), g, "studio.scene")
    print(f"studio.scene: total ratio {ratio:.5f}, relMSE {relmse:.2e}", flush=True)
