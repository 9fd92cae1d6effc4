"""Alpha masks (mesh.h:24-32, 82-86; obj_loader.h:169-172): an MTL's map_d,
or without one a diffuse texture with a non-opaque alpha channel, cuts holes
in its triangles: at a hit, the nearest texel's coverage c (alpha / 255, or
the grey level of an image without alpha) keeps the hit when a uniform draw
is below c, so a ray passes through with probability 1 - c.

CPU: the OBJ/MTL loader's masks and flags; independent of the reference, the
closed form of a black masked quad before a white background (pixels of
coverage 0, 128/255 and 1 give 1, 127/255 and 0); the oracle's converged
render of scenes/alpha_leaves.scene against the reference's
(tests/golden/alpha_leaves.npz, the reference's own loaders). GPU: the
oracle at equal seed (both draw the mask test from the ray's stream keyed by
the triangle), and the converged means against the same golden. The random
draw inside the hit test keeps masked scenes out of the ray-query goldens."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as ol
from conftest import ROOT, scene_path

SCENE = "alpha_leaves.scene"
GOLD = os.path.join(ROOT, "tests", "golden", "alpha_leaves.npz")
sys.path.insert(0, os.path.join(ROOT, "scenes"))


def test_obj_masks_load():
    import mypbr_amd as m
    from mypbr_amd.abi import OBJ_MESH, TEXF_NO_ALPHA
    d = m.describe(scene_path(SCENE))
    alpha = [d.desc.object_alpha[o] for o in range(d.desc.n_objects)]
    kinds = [d.desc.object_kind[o] for o in range(d.desc.n_objects)]
    masked = [a for a in alpha if a >= 0]
    assert len(masked) == 2  # the leaves (map_d) and the fence (its diffuse texture's alpha)
    assert all(kinds[o] == OBJ_MESH for o in range(len(alpha)) if alpha[o] >= 0)
    for a in masked:
        assert not (d.desc.textures[a].flags & TEXF_NO_ALPHA)


def _masked_quad(tmp_path, alpha_row, grey=False):
    """A black quad, textured (0..1)^2 by a 3 x 1 mask, facing a camera
    under a white background."""
    from make_assets import png_write
    if grey:
        png_write(str(tmp_path / "mask.png"), [list(alpha_row)], 0, 8)
    else:
        png_write(str(tmp_path / "mask.png"), [[v for a in alpha_row for v in (255, a)]], 4, 8)
    (tmp_path / "q.mtl").write_text("newmtl black\nKd 0 0 0\nmap_d mask.png\n")
    (tmp_path / "q.obj").write_text("mtllib q.mtl\nv -1.5 -1 0\nv 1.5 -1 0\nv 1.5 1 0\nv -1.5 1 0\n"
                                    "vt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\nusemtl black\nf 1/1 2/2 3/3 4/4\n")
    (tmp_path / "q.scene").write_text("image_width 30\naspect_ratio 1.5\nsamples_per_pixel 256\nmax_depth 4\n"
                                      "background 1 1 1\ncamera lookfrom 0 0 5  lookat 0 0 0  vfov 20\n"
                                      f"mesh {tmp_path / 'q.obj'}\n")
    return str(tmp_path / "q.scene")


def test_grey_mask_reads_the_grey_level(tmp_path):
    import mypbr_amd as m
    from mypbr_amd.abi import TEXF_NO_ALPHA
    d = m.describe(_masked_quad(tmp_path, (0, 128, 255), grey=True))
    a = d.desc.object_alpha[0]
    assert a >= 0 and d.desc.textures[a].flags & TEXF_NO_ALPHA


@pytest.mark.parametrize("grey", [False, True])
def test_oracle_masked_quad_closed_form(tmp_path, grey):
    """Coverage 0 lets every ray through (exactly white), 255 none (exactly
    black), 128 a fraction 127/255 (binomial noise of 256 samples). A mask
    read from the wrong byte, with rows or u flipped, or a draw compared the
    wrong way round fails one of the three bands."""
    import mypbr_amd as m
    d = m.describe(_masked_quad(tmp_path, (0, 128, 255), grey=grey))
    img = ol.OracleScene(d.desc).render(d.camera, seed=3)["radiance"][..., 0]
    h, w = img.shape
    # the quad (x in [-1.5, 1.5], u = (x + 1.5) / 3) fills the view, whose half width at the quad is
    # 5 tan(10 deg) * 1.5; the mask's three texels are thirds of u. Columns whose whole footprint lies
    # inside one third, with a margin:
    half = 5 * np.tan(np.radians(10)) * w / h
    edges = (-half + 2 * half * np.arange(w + 1) / w + 1.5) / 3
    lo, hi = edges[:-1], edges[1:]
    xs = [(lo > a + 0.01) & (hi < b - 0.01) for a, b in ((0, 1 / 3), (1 / 3, 2 / 3), (2 / 3, 1))]
    assert all(x.sum() >= 3 for x in xs)
    band = [img[:, x] for x in xs]
    assert np.all(band[0] == 1.0)
    assert np.all(band[2] == 0.0)
    want = 127 / 255
    sigma = np.sqrt(want * (1 - want) / d.camera.samples_per_pixel)
    assert abs(band[1].mean() - want) < 4 * sigma / np.sqrt(band[1].size)
    assert np.all(np.abs(band[1] - want) < 5 * sigma)


def _runs(render, seed0):
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    g = np.load(GOLD)
    d = m.describe(scene_path(SCENE))
    cam = camera_copy(d.camera, image_width=int(g["width"]), samples_per_pixel=int(g["spp"]) // 8,
                      max_depth=int(g["max_depth"]))
    runs = np.stack([render(d, cam, seed0 + k) for k in range(8)])
    return runs.mean(0), runs.std(0, ddof=1) / np.sqrt(8), g


def test_oracle_render_converges_to_reference():
    from test_oracle_golden import render_agrees
    mean, se, g = _runs(lambda d, cam, k: ol.OracleScene(d.desc).render(cam, seed=k)["radiance"], 100)
    render_agrees(mean, se, g, SCENE)


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


@pytest.mark.gpu
def test_gpu_render_converges_to_reference():
    import mypbr_amd as m
    from test_oracle_golden import render_agrees
    mean, se, g = _runs(lambda d, cam, k: m.Scene.from_desc(d.desc, d.camera).render(cam, seed=k).radiance, 300)
    ratio, relmse = render_agrees(mean, se, g, SCENE)
    print(f"{SCENE}: total ratio {ratio:.5f}, relMSE {relmse:.2e}", flush=True)
