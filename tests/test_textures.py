"""Image textures (texture.h:58-208), tangent-space normal maps
(normal_from_map and normal_mapped, material.h:282-308, 631-674) and
principled's metallic / roughness maps (material.h:508-515), with the hit's
texture coordinates and dp/du, dp/dv (sphere.h:46-48, 99-120, quad.h:53-57,
mesh.h:71-93, 149-166, instance.h:38-39).

CPU: the PNG and JPEG decoders give stb_image's bytes (the golden holds
stbi_load's output for every texture file of scenes/textured.scene; where
the compiled reference is present, also for the reference's own textures and
for PNG and JPEG files of every color type and chroma layout written here);
the oracle's surface record (u, v, tangents, texel values, the normal after a
normal map, principled's metallic and alpha) equals the reference's bit for
bit at every hit of the golden ray subset; hits and converged means against
the reference's render of the scene. GPU: the HIP path against the golden
hits and converged means, and against the oracle at equal seed."""
import hashlib
import json
import os
import struct
import sys
import zlib

import numpy as np
import pytest

import oracle_lib as ol
from conftest import ROOT, scene_path

GOLD = os.path.join(ROOT, "tests", "golden", "textured.npz")
SCENE = "textured.scene"
MODELS = os.path.join(ROOT, "scenes", "models")
REF_TEXTURES = "/root/reference/textures"

sys.path.insert(0, os.path.join(ROOT, "scenes"))
from make_assets import png_write  # noqa: E402


def _golden():
    g = np.load(GOLD)
    return g, json.loads(str(g["digest"]))


def decoded(path, tmp_path, linear=True):
    """The RGBA bytes the scene loader keeps for an image texture file."""
    import mypbr_amd as m
    s = tmp_path / "t.scene"
    s.write_text(f"image_width 8\ntexture t image {path}{' linear' if linear else ''}\n"
                 "material a lambertian texture t\nsphere 0 0 -1 0.5 a\n")
    d = m.describe(str(s))
    t = d.desc.textures[0]
    n = t.width * t.height * 4
    px = np.ctypeslib.as_array(d.desc.texels, shape=(d.desc.n_texel_bytes,))[4 * t.offset: 4 * t.offset + n]
    return px.reshape(t.height, t.width, 4).copy()


@pytest.mark.parametrize("name", ["globe.jpg", "bumps_normal.png", "tiles_rough.png", "tiles_metal.png"])
def test_texture_files_decode_to_stb_bytes(tmp_path, name):
    g, _ = _golden()
    assert np.array_equal(decoded(os.path.join(MODELS, name), tmp_path), g["image/" + name])


def _png_cases():
    rng = np.random.default_rng(5)
    w, h = 13, 7  # odd sizes: partial bytes at low bit depths, odd chroma blocks
    cases = []
    for depth in (1, 2, 4, 8, 16):
        rows = rng.integers(0, 1 << depth, (h, w)).tolist()
        cases.append((f"grey{depth}", rows, 0, depth, None, None))
    for depth in (1, 2, 4, 8):
        n = 1 << depth
        pal = [tuple(int(v) for v in rng.integers(0, 256, 3)) for _ in range(n)]
        rows = rng.integers(0, n, (h, w)).tolist()
        cases.append((f"palette{depth}", rows, 3, depth, pal, None))
        cases.append((f"palette{depth}_trns", rows, 3, depth, pal, bytes(rng.integers(0, 256, n).tolist())))
    for depth in (8, 16):
        for ct, k in ((2, 3), (4, 2), (6, 4)):
            rows = rng.integers(0, 1 << depth, (h, w * k)).tolist()
            cases.append((f"ct{ct}_{depth}", rows, ct, depth, None, None))
    # colour keys: the pixels equal to the key become transparent
    grey = rng.integers(0, 4, (h, w)).tolist()
    cases.append(("grey2_key", grey, 0, 2, None, struct.pack(">H", 2)))
    grey8 = rng.integers(0, 3, (h, w)).tolist()
    cases.append(("grey8_key", grey8, 0, 8, None, struct.pack(">H", 1)))
    rgb16 = (rng.integers(0, 2, (h, w * 3)) * 40000).tolist()
    cases.append(("rgb16_key", rgb16, 2, 16, None, struct.pack(">HHH", 40000, 0, 40000)))
    return cases


def _write_png(path, rows, ct, depth, pal, trns):
    png_write(path, rows, ct, depth, palette=pal) if ct in (0, 2, 3) else _png_any(path, rows, ct, depth)
    if trns is not None:  # insert a tRNS chunk before the first IDAT
        data = open(path, "rb").read()
        i = data.index(b"IDAT") - 4
        chunk = struct.pack(">I", len(trns)) + b"tRNS" + trns + struct.pack(">I", zlib.crc32(b"tRNS" + trns))
        open(path, "wb").write(data[:i] + chunk + data[i:])


def _png_any(path, rows, ct, depth):
    """Filter-0 PNG for color types png_write does not take (grey + alpha, RGBA)."""
    body = b"".join(b"\0" + (b"".join(struct.pack(">H", v) for v in r) if depth == 16 else bytes(r)) for r in rows)
    k = {4: 2, 6: 4}[ct]

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d))
    hdr = struct.pack(">IIBBBBB", len(rows[0]) // k, len(rows), depth, ct, 0, 0, 0)
    open(path, "wb").write(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", hdr) + chunk(b"IDAT", zlib.compress(body)) +
                           chunk(b"IEND", b""))


@pytest.mark.skipif(not ol.reference_available(), reason="the compiled reference (oracle/_ref) is not built")
def test_png_color_types_match_stb(tmp_path):
    for name, rows, ct, depth, pal, trns in _png_cases():
        p = str(tmp_path / f"{name}.png")
        _write_png(p, rows, ct, depth, pal, trns)
        want, _ = ol.reference_image_u8(p)
        assert np.array_equal(decoded(p, tmp_path), want), name


@pytest.mark.skipif(not ol.reference_available(), reason="the compiled reference (oracle/_ref) is not built")
def test_jpeg_layouts_match_stb(tmp_path):
    """Baseline JPEGs from stb_image_write: grey and color, full-resolution
    chroma (quality 95) and 2x2-subsampled chroma (80, 30), odd sizes; and the
    reference's own textures where its sources are present."""
    rng = np.random.default_rng(9)
    img = np.clip(rng.normal(128, 60, (29, 45, 3)) + np.linspace(-60, 60, 45)[None, :, None], 0, 255).astype(np.uint8)
    paths = []
    for q in (95, 80, 30):
        for comp in (1, 3):
            p = str(tmp_path / f"q{q}c{comp}.jpg")
            ol.reference_write_jpg(p, img[:, :, 0] if comp == 1 else img, q)
            paths.append(p)
    if os.path.isdir(REF_TEXTURES):
        paths += [os.path.join(REF_TEXTURES, f) for f in sorted(os.listdir(REF_TEXTURES))
                  if f.endswith((".jpg", ".png"))]
    for p in paths:
        want, _ = ol.reference_image_u8(p)
        assert np.array_equal(decoded(p, tmp_path), want), p


def test_oracle_surface_matches_reference():
    """u, v, dp/du, dp/dv, texel values (JPEG sRGB, 16-bit grey, palette,
    Radiance floats), the normal after each normal map (quad, spheres, a
    rotated instance, a smooth mirror), principled's metallic and alpha from
    their maps: bitwise equal to the reference at every hit of the subset."""
    import mypbr_amd as m
    g, _ = _golden()
    d = m.describe(scene_path(SCENE))
    o = ol.OracleScene(d.desc)
    got = o.surface(g["rays"], g["tt"])
    want = g["surface"]
    assert np.array_equal(np.isnan(got), np.isnan(want))
    hit = ~np.isnan(want[:, 0])
    assert hit.sum() > 1000
    assert np.array_equal(got[hit].view(np.int64), want[hit].view(np.int64))
    assert np.unique(np.round(want[hit, 14], 6)).size > 2  # metallic varies: read from the palette map


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


@pytest.mark.parametrize("body,ok", [
    ("texture t image models/globe.jpg\nmaterial a lambertian texture t normal_map t\n", True),
    ("texture t image models/none.png\n", False),                                     # missing file
    ("texture t image models/globe.jpg srgb\n", False),                                # bad option
    ("texture t image models/blob.obj\n", False),                                      # not an image
    ("texture t image models/globe.jpg\nmaterial a dielectric 1.5 normal_map t\n", False),
    ("texture t image models/globe.jpg\nmaterial a lambertian 1 1 1 normal_map u\n", False),  # unknown texture
    ("texture t image models/globe.jpg\nmaterial a lambertian 1 1 1 bump t\n", False),
    ("texture t image models/globe.jpg\nmaterial a principled metallic texture t roughness 2\n", False),
    ("texture t image models/globe.jpg\nmaterial a principled metallic texture t roughness texture t\n", True),
])
def test_texture_lines(tmp_path, body, ok):
    import mypbr_amd as m
    s = tmp_path / "s.scene"
    s.write_text(f"image_width 8\n{body}sphere 0 0 -1 0.5\n")
    os.symlink(MODELS, tmp_path / "models")
    if ok:
        m.describe(str(s))
    else:
        with pytest.raises(m.MptError):
            m.describe(str(s))


@pytest.mark.parametrize("data", [
    b"\x89PNG\r\n\x1a\n",                                                  # no chunks
    b"\x89PNG\r\n\x1a\n" + struct.pack(">I", 13) + b"IHDR" + struct.pack(">IIBBBBB", 4, 4, 8, 2, 0, 0, 1) + b"\0" * 4,
    b"\xff\xd8\xff\xc2\x00\x0b\x08\x00\x08\x00\x08\x01\x01\x11\x00\xff\xd9",  # progressive JPEG
    b"\xff\xd8\xff\xd9",                                                   # no image data
])
def test_malformed_images_are_refused(tmp_path, data):
    import mypbr_amd as m
    p = tmp_path / "bad.img"
    p.write_bytes(data)
    with pytest.raises(m.MptError):
        decoded(str(p), tmp_path)


@pytest.mark.parametrize("mtl,alpha,opaque", [
    ("map_Kd rgba.png", 255, True),      # an alpha channel that is opaque everywhere is no mask
    ("map_Kd rgba.png", 0, False),       # a diffuse texture's alpha channel is a mask (obj_loader.h:171-172)
    ("map_d grey.png", 0, False),        # map_d's grey level is a mask
    ("map_d grey.png", 255, True),
])
def test_obj_alpha_masks_follow_the_loader_rule(tmp_path, mtl, alpha, opaque):
    """The MTL gives a mesh an alpha mask exactly when the reference's
    loader does (obj_loader.h:169-172); masks themselves: tests/test_alpha.py."""
    import mypbr_amd as m
    _png_any(str(tmp_path / "rgba.png"), [[200, 100, 50, 255, 10, 20, 30, alpha]] * 2, 6, 8)
    png_write(str(tmp_path / "grey.png"), [[255, alpha]] * 2, 0, 8)
    (tmp_path / "m.mtl").write_text(f"newmtl a\nKd 1 1 1\n{mtl}\n")
    (tmp_path / "m.obj").write_text("mtllib m.mtl\nusemtl a\nv 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\n"
                                    "f 1/1 2/2 3/3\n")
    (tmp_path / "s.scene").write_text("image_width 8\nmesh m.obj\n")
    d = m.describe(str(tmp_path / "s.scene"))
    masked = bool(d.desc.object_alpha) and d.desc.object_alpha[0] >= 0
    assert masked == (not opaque)


def test_normal_maps_off_drops_every_map():
    import mypbr_amd as m
    text = open(scene_path(SCENE)).read()
    d = m.describe(scene_path(SCENE))
    assert any(d.desc.materials[i].normal_map >= 0 for i in range(d.desc.n_materials))
    p = scene_path("textured_flat.scene")
    try:
        open(p, "w").write("normal_maps off\n" + text)
        d2 = m.describe(p)
        assert all(d2.desc.materials[i].normal_map < 0 for i in range(d2.desc.n_materials))
    finally:
        os.remove(p)


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
    """Equal seed: almost every pixel equal to 1e-12, guides included (the
    albedo guide reads the textures at first hits, through normal-mapped
    mirrors)."""
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_gpu_render import _compare
    d = m.describe(scene_path(SCENE))
    cam = camera_copy(d.camera, image_width=64, samples_per_pixel=8, max_depth=12, collect_features=1)
    cpu = ol.OracleScene(d.desc).render(cam, seed=7)
    f = m.Scene.from_desc(d.desc, d.camera).render(cam, seed=7)
    # the radiance bar is lower here: bilinear texel weights and the normal
    # maps' renormalisation send about 2% of this image's paths apart in the
    # last bits (97.7% of pixels equal on MI355X); the guides are exact
    _compare(f.radiance, cpu["radiance"], 0.97, 1e-3)
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
