"""glTF 2.0 models (gltf_loader.h; the `mesh file.gltf|.glb [replace a b]...
[material] [transforms]` line, scene.h:1129-1197): the reference's material
mapping and node flattening, read by this repository's own glTF reader.

CPU: what the loader builds from scenes/models/props.gltf and props_box.glb
(the texture wrappers of each material: scaled base colour under
KHR_texture_transform, metallic and roughness as channels, a normal map's
scaled tilt; thin glass, a clear coat, an emitter light, a glowing screen,
masks; the default material; replace and override), the line's errors; the
oracle against the compiled reference on its own cgltf-based loader
(tests/golden/gltf_props.npz: the hit t and front face of a 2^16-ray batch's
subset and every hit's surface record; converged means of gltf_props.scene
and gltf_glb.scene). Part order differs from the reference's (its parts sit
in a map keyed by pointers), so primitive IDs are not compared. GPU: renders
at equal seed against the oracle and converged means against the
reference's."""
import math
import os

import numpy as np
import pytest

import oracle_lib as ol
from conftest import ROOT, scene_path

GOLD = os.path.join(ROOT, "tests", "golden")


def _desc(name):
    import mypbr_amd as m
    return m.describe(scene_path(name))


def test_props_materials_and_textures():
    from mypbr_amd.abi import (MAT_DIELECTRIC, MAT_DIFFUSE_LIGHT, MAT_PRINCIPLED, TEX_CHANNEL, TEX_IMAGE,
                               TEX_NORMAL_SCALE, TEX_SCALED, TEX_UV_TRANSFORM)
    keep = _desc("gltf_props.scene")  # owns the arrays the description points into
    d = keep.desc
    mats = [d.materials[k] for k in range(d.n_materials)]
    tex = [d.textures[k] for k in range(d.n_textures)]
    types = [m.type for m in mats]
    assert types.count(MAT_DIFFUSE_LIGHT) == 1 and d.n_lights == 2  # the lamp and the sky
    lamp = next(m for m in mats if m.type == MAT_DIFFUSE_LIGHT)
    assert list(lamp.emit) == [4.0, 3.200000047683716 * 1.0, 2.0]  # emissiveFactor x strength, as floats
    glass = next(m for m in mats if m.type == MAT_DIELECTRIC)
    assert glass.thin == 1 and abs(glass.ior - float(np.float32(1.45))) == 0
    assert sum(m.coat for m in mats) == 1 and sum(m.glow for m in mats) == 1
    glow = next(m for m in mats if m.glow)
    assert tex[glow.glow_texture].type == TEX_SCALED and tex[tex[glow.glow_texture].inner].type == TEX_IMAGE
    # the default material: white, metallic 1, roughness 1
    default = [m for m in mats if m.type == MAT_PRINCIPLED and m.metallic_texture == -1 and m.metallic == 1 and
               m.roughness == 1 and tex[m.texture].type == 0 and list(tex[m.texture].even) == [1.0, 1.0, 1.0]]
    assert default
    # the painted base: scaled over a uv transform over the clamped checker
    uvt = [t for t in tex if t.type == TEX_UV_TRANSFORM]
    assert len(uvt) == 1
    u = uvt[0]
    rot = float(np.float32(0.4))
    assert list(u.uv) == [0.25, float(np.float32(-0.1)), 2.0, 1.5, math.cos(rot), math.sin(rot)]
    assert tex[u.inner].type == TEX_IMAGE and not (tex[u.inner].flags & 2)  # the sampler clamps
    assert any(t.type == TEX_SCALED and t.inner >= 0 and tex[t.inner].type == TEX_UV_TRANSFORM for t in tex)
    # metallic from blue, roughness from green of one linear image
    ch = sorted((t.channel, round(t.scale, 6)) for t in tex if t.type == TEX_CHANNEL)
    assert (2, 1.0) in ch and (1, 0.9) in ch
    ns = [t for t in tex if t.type == TEX_NORMAL_SCALE]
    assert len(ns) == 1 and ns[0].scale == 0.5
    # the leaf's mask: the base texture's alpha
    masks = [d.object_alpha[o] for o in range(d.n_objects) if d.object_alpha[o] >= 0]
    assert len(masks) == 1 and tex[masks[0]].type == TEX_IMAGE


def test_glb_replace_and_override():
    from mypbr_amd.abi import MAT_LAMBERTIAN
    keep = _desc("gltf_glb.scene")
    d = keep.desc
    mats = [d.materials[k] for k in range(d.n_materials)]
    # the constant coverage (0.46 as a byte) masks the ghost box in both placements
    masks = [d.object_alpha[o] for o in range(d.n_objects) if d.object_alpha[o] >= 0]
    ghost = [d.textures[k] for k in masks if d.textures[k].width == 1]
    assert len(ghost) == 2
    assert sum(m.type == MAT_LAMBERTIAN for m in mats) == 1  # the clay, standing in for every props material


@pytest.mark.parametrize("line,needle", [
    ("mesh models/props.gltf replace nosuch clay", "has no material named 'nosuch'"),
    ("mesh models/props.gltf replace painted clay clay", "can't be used together"),
    ("mesh models/blob.obj replace a clay", "'replace' is for glTF files"),
    ("mesh models/nosuch.gltf", "cannot read glTF file"),
    ("mesh models/props.gltf replace painted", "expected 'replace"),
])
def test_mesh_line_errors(tmp_path, line, needle):
    import mypbr_amd as m
    s = tmp_path / "s.scene"
    s.write_text(f"image_width 8\nmaterial clay lambertian 0.5 0.5 0.5\n"
                 f"{line.replace('models/', ROOT + '/scenes/models/')}\n")
    with pytest.raises(m.MptError) as e:
        m.describe(str(s))
    assert needle in str(e.value), str(e.value)


def test_broken_files(tmp_path):
    import mypbr_amd as m
    cases = {"bad.gltf": b"{\"asset\": {\"version\": \"2.0\"}, \"buffers\": [{\"uri\": \"missing.bin\", \"byteLength\": 8}]}",
             "trunc.glb": b"glTF\x02\x00\x00\x00\x40\x00\x00\x00\x30\x00\x00\x00JSON{}",
             "empty.gltf": b"{\"asset\": {\"version\": \"2.0\"}}", "junk.gltf": b"{\"asset\": [1, 2"}
    for name, data in cases.items():
        (tmp_path / name).write_bytes(data)
        s = tmp_path / "s.scene"
        s.write_text(f"image_width 8\nmesh {tmp_path / name}\n")
        with pytest.raises(m.MptError):
            m.describe(str(s))


def test_oracle_hits_and_surfaces_match_reference():
    """t and front face bitwise at every hit of the subset, and the surface
    record: u, v, dp/du, dp/dv (the flattened node matrices' float
    arithmetic, uvs flipped), base colour (texture transform, scale,
    clamping, normalized u16 texture coordinates), the normal after the
    scaled normal map, metallic and alpha from the map's channels."""
    g = np.load(os.path.join(GOLD, "gltf_props.npz"))
    d = _desc("gltf_props.scene")
    o = ol.OracleScene(d.desc)
    ids, t, f = o.trace(g["rays"], g["tt"])
    hit = g["id"] != -1
    assert hit.sum() > 1000
    assert np.array_equal(ids != -1, hit)
    single = hit & (g["ties"] <= 1)
    assert np.array_equal(t[single].view(np.int64), g["t"][single].view(np.int64))
    assert np.array_equal(f[single], g["front"][single])
    got, want = o.surface(g["rays"], g["tt"]), g["surface"]
    h = single & ~np.isnan(want[:, 0])
    assert np.array_equal(np.isnan(got[single]), np.isnan(want[single]))
    assert np.array_equal(got[h].view(np.int64), want[h].view(np.int64))
    assert len(np.unique(np.round(want[h, 8:11], 9), axis=0)) >= 6
    assert np.unique(np.round(want[h, 14], 6)).size >= 2  # metallic from the map's blue channel


@pytest.mark.parametrize("name,seed0", [("gltf_props.scene", 100), ("gltf_glb.scene", 200)])
def test_oracle_render_converges_to_reference(name, seed0):
    from mypbr_amd.abi import camera_copy
    from test_oracle_golden import render_agrees
    g = np.load(os.path.join(GOLD, name.replace(".scene", ".npz")))
    d = _desc(name)
    cam = camera_copy(d.camera, image_width=int(g["width"]), samples_per_pixel=int(g["spp"]) // 8,
                      max_depth=int(g["max_depth"]))
    runs = np.stack([ol.OracleScene(d.desc).render(cam, seed=seed0 + k)["radiance"] for k in range(8)])
    render_agrees(runs.mean(0), runs.std(0, ddof=1) / np.sqrt(8), g, name)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["gltf_props.scene", "gltf_glb.scene"])
def test_gpu_render_matches_oracle(name):
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_gpu_render import _compare
    d = _desc(name)
    cam = camera_copy(d.camera, image_width=64, samples_per_pixel=8, max_depth=12, collect_features=1)
    cpu = ol.OracleScene(d.desc).render(cam, seed=7)
    f = m.Scene.from_desc(d.desc, d.camera).render(cam, seed=7)
    _compare(f.radiance, cpu["radiance"], 0.995, 1e-3)
    _compare(f.albedo, cpu["albedo"], 0.995, 1e-3)
    _compare(f.normal, cpu["normal"], 0.995, 1e-3)


@pytest.mark.gpu
def test_gpu_hits_match_reference():
    import mypbr_amd as m
    g = np.load(os.path.join(GOLD, "gltf_props.npz"))
    d = _desc("gltf_props.scene")
    ids, t, f = m.Scene.from_desc(d.desc, d.camera).trace_closest(g["rays"], g["tt"])
    single = (g["id"] != -1) & (g["ties"] <= 1)
    assert np.array_equal(ids != -1, g["id"] != -1)
    assert np.array_equal(t[single].view(np.int64), g["t"][single].view(np.int64))
    assert np.array_equal(f[single], g["front"][single])


@pytest.mark.gpu
@pytest.mark.parametrize("name,seed0", [("gltf_props.scene", 300), ("gltf_glb.scene", 400)])
def test_gpu_render_converges_to_reference(name, seed0):
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_oracle_golden import render_agrees
    g = np.load(os.path.join(GOLD, name.replace(".scene", ".npz")))
    d = _desc(name)
    cam = camera_copy(d.camera, image_width=int(g["width"]), samples_per_pixel=int(g["spp"]) // 8,
                      max_depth=int(g["max_depth"]))
    s = m.Scene.from_desc(d.desc, d.camera)
    runs = np.stack([s.render(cam, seed=seed0 + k).radiance for k in range(8)])
    ratio, relmse = render_agrees(runs.mean(0), runs.std(0, ddof=1) / np.sqrt(8), g, name)
    print(f"{name}: total ratio {ratio:.5f}, relMSE {relmse:.2e}", flush=True)
