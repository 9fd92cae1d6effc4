"""Writes the golden vectors in tests/golden from the compiled reference.

Runs only where oracle/_ref/libmpt_ref.so was built (the reference renderer
compiled from its own sources by oracle/Makefile). Every expected value
here is computed by the reference itself; inputs (ray batches, material
queries) are seeded and regenerated identically by the tests:

  ggx_table.npz         the 32x32 E(mu, alpha) table and E_avg (microfacet.h:265-338)
  trace_<scene>.npz     4096 rays of the seeded batch with the reference's
                        closest-hit ID, t, front_face and tie count (bvh.h:45-86)
  trace_ties.npz        rays aimed at where primitives meet, and rays whose interval
                        ends or starts at the hit t, with the reference's closest hit
                        and its tie candidates (count, smallest and largest ID)
  trace_digests.json    SHA-256 of the reference's IDs and t bits over the full
                        1M-ray batches, tie and miss counts
  light_probs.json      light_list::probability per light (light_list.h:45-127)
  bsdf_coat.npz         the same for principled with a clear coat (set_clearcoat,
                        set_clearcoat_color, material.h:326-552)
  bsdf.npz              material::eval and the scatter pdf's value at fixed
                        directions (material.h, microfacet.h, pdf.h, medium.h)
  render_<scene>.npz    converged low-resolution means of camera::render and
                        their per-pixel standard error over 8 runs
  denoise.npz           denoiser::run (denoiser.h:45-193) on the synthetic frames
                        of tests/denoise_frames.py, for two settings
  metals.npz            scenes/metals.scene (measured metals, thin glass) as below
  ground.npz            scenes/ground.scene (the ground plane beside the BVH) as textured.npz
  sun.npz               scenes/sun.scene (sun_light, sun.h) as below, plus the reference's
                        light_list::probability per light
  height_fog.npz        scenes/height_fog.scene (height_fog, medium.h:98-155) as sun.npz
  alpha_leaves.npz      scenes/alpha_leaves.scene (alpha masks, mesh.h:82-86, through the
                        reference's load_obj/load_mtl) as sun.npz
  clearcoat.npz         scenes/clearcoat.scene (principled's clear coat) as sun.npz
  sky_atmosphere.npz    scenes/sky_atmosphere.scene (the sky bake, atmosphere.h) as sun.npz,
                        plus the reference's environment_light::radiance of its baked sky
                        along 4096 seeded directions
  textured.npz          scenes/textured.scene (image textures, normal maps, texture
                        maps) as below, plus the reference's surface record at each
                        hit of the subset and stbi_load's bytes of its texture files
  obj_room.npz,         scenes/obj_room.scene (Wavefront OBJ/MTL models) and
  hdr_env.npz           scenes/hdr_env.scene (a run-length encoded .hdr environment)
                        loaded by the reference's own loaders: a 4096-ray subset of
                        the seeded batch with the reference's hit IDs, t and front
                        face, digests over the 1M rays, and converged means
  display.npz           postprocess (postprocess.h:108-153) and the bytes write_png
                        writes (display_settings::apply + to_byte, tonemap.h,
                        color3.h) on the images and settings of tests/display_cases.py

Usage: python tests/golden/make_golden.py [step ...]  (steps: ggx trace ties bsdf render denoise display obj textures materials sun fog alpha coat; all by default)
(all steps take about a minute)
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import mypbr_amd as m  # noqa: E402  (host-side scene description only)
import oracle_lib as ol  # noqa: E402

TRACE_SCENES = ["cornell.scene", "weekend_cover.scene", "test_lights.scene"]
N_BATCH = 1 << 20
N_SUBSET = 4096
BATCH_SEED = 1
RENDERS = {  # scene: (width, spp per run, runs, max_depth)
    "cornell.scene": (32, 512, 8, 16),
    "weekend_cover.scene": (48, 256, 8, 50),
    "test_lights.scene": (40, 256, 8, 24),
    "test_media.scene": (32, 512, 8, 32),
}

# material specs for the BSDF vectors: type r g b roughness ior metallic g sigma_a(3) unused
BSDF_MATERIALS = [
    [0, 0.7, 0.5, 0.3, 0, 0, 0, 0, 0, 0, 0, 0],
    [1, 0.9, 0.6, 0.3, 0.3, 0, 0, 0, 0, 0, 0, 0],
    [1, 0.95, 0.95, 0.95, 0.8, 0, 0, 0, 0, 0, 0, 0],
    [2, 1, 1, 1, 0.2, 1.5, 0, 0, 0, 0, 0, 0],
    [2, 1, 1, 1, 0.6, 1.33, 0, 0, 0, 0, 0, 0],
    [3, 0.6, 0.2, 0.1, 0.4, 1.5, 0.0, 0, 0, 0, 0, 0],
    [3, 0.8, 0.7, 0.3, 0.7, 1.5, 1.0, 0, 0, 0, 0, 0],
    [3, 0.5, 0.5, 0.9, 0.05, 1.45, 0.5, 0, 0, 0, 0, 0],
    [5, 0.9, 0.8, 0.7, 0, 0, 0, 0.3, 0, 0, 0, 0],
    [5, 0.6, 0.6, 0.6, 0, 0, 0, 0.0, 0, 0, 0, 0],
    # measured metals (find_conductor): gold and aluminium, principled over chromium
    [1, 0, 0, 0, 0.3, 0, 0, 0, 0, 0, 0, 4],
    [1, 0, 0, 0, 0.7, 0, 0, 0, 0, 0, 0, 1],
    [3, 0.5, 0.4, 0.3, 0.35, 1.5, 0.8, 0, 0, 0, 0, 6],
]
N_PER_MATERIAL = 256


def scene_path(name):
    return os.path.join(ROOT, "scenes", name)


# principled with a clear coat: the 12 columns above, then coat weight, coat roughness, tint (3)
COAT_MATERIALS = [
    [3, 0.6, 0.2, 0.1, 0.4, 1.5, 0.0, 0, 0, 0, 0, 0, 1.0, 0.0, 1, 1, 1],
    [3, 0.8, 0.7, 0.3, 0.5, 1.5, 1.0, 0, 0, 0, 0, 0, 0.7, 0.1, 0.6, 0.3, 0.2],
    [3, 0.5, 0.4, 0.3, 0.3, 1.5, 1.0, 0, 0, 0, 0, 6, 1.0, 0.05, 0.6, 0.35, 0.28],
    [3, 0.2, 0.5, 0.8, 0.7, 1.45, 0.3, 0, 0, 0, 0, 0, 0.3, 0.3, 1, 1, 1],
]


def bsdf_queries(seed=5, materials=None):
    """Seeded (normal, ray direction, wi, front) queries; the ray direction
    points against the normal as after set_face_normal (hittable.h:23-26)."""
    materials = BSDF_MATERIALS if materials is None else materials
    rng = np.random.default_rng(seed)
    n = len(materials) * N_PER_MATERIAL

    def unit(k):
        v = rng.normal(size=(k, 3))
        return v / np.linalg.norm(v, axis=1, keepdims=True)

    nrm = unit(n)
    rd = unit(n)
    flip = np.sum(rd * nrm, axis=1) > 0
    rd[flip] *= -1
    rd *= rng.uniform(0.5, 3.0, size=(n, 1))  # camera and instance rays are not unit length
    wi = unit(n)
    # a quarter of the queries ask for the mirror direction (sharp lobes)
    k = n // 4
    ud = rd[:k] / np.linalg.norm(rd[:k], axis=1, keepdims=True)
    wi[:k] = ud - 2 * np.sum(ud * nrm[:k], axis=1, keepdims=True) * nrm[:k]
    front = rng.integers(0, 2, size=n).astype(np.uint8)
    spec = np.repeat(np.array(materials, dtype=np.float64), N_PER_MATERIAL, axis=0)
    q = np.concatenate([nrm, rd, wi], axis=1)
    return spec, np.ascontiguousarray(q), front


def golden_denoise():
    import denoise_frames as dfr
    arrays = {}
    for name, make in dfr.CASES.items():
        f = make()
        for k in ("radiance", "albedo", "normal", "variance"):
            arrays[f"{name}/{k}"] = f[k]
        for si, st in enumerate(dfr.SETTINGS):
            arrays[f"{name}/out{si}"] = ol.reference_denoise(f, st)
    arrays["settings"] = np.array(dfr.SETTINGS)
    np.savez_compressed(os.path.join(HERE, "denoise.npz"), **arrays)
    print("denoise.npz", len(dfr.CASES), "frames", flush=True)


def golden_display():
    import display_cases as dc
    arrays = {}
    for name, make in dc.CASES.items():
        img = make()
        arrays[f"{name}/radiance"] = img
        for si, st in enumerate(dc.SETTINGS):
            arrays[f"{name}/post{si}"] = ol.reference_postprocess(img, st)
            arrays[f"{name}/bytes{si}"] = ol.reference_display_bytes(img, st)
        arrays[f"{name}/bytes_nopost"] = ol.reference_display_bytes(img, dc.SETTINGS[2], post=False)
    arrays["settings"] = np.array(dc.SETTINGS)
    np.savez_compressed(os.path.join(HERE, "display.npz"), **arrays)
    print("display.npz", len(dc.CASES), "images", flush=True)


ASSET_SCENES = {  # scene: (width, spp per run, runs, max_depth) of the converged render
    "obj_room.scene": (40, 256, 8, 16),
    "hdr_env.scene": (48, 256, 8, 16),
}
TEXTURE_SCENES = {
    "textured.scene": (48, 256, 8, 12),
    "ground.scene": (48, 256, 8, 12),
}
MATERIAL_SCENES = {
    "metals.scene": (48, 256, 8, 16),
}
SUN_SCENES = {
    "sun.scene": (48, 256, 8, 12),
    "sky_atmosphere.scene": (48, 256, 8, 12),
    "sun_disk.scene": (48, 256, 8, 12),
}


FOG_SCENES = {
    "height_fog.scene": (48, 256, 8, 12),
}
ALPHA_SCENES = {
    "alpha_leaves.scene": (48, 256, 8, 12),
    "clearcoat.scene": (48, 256, 8, 12),
}


VOXEL_SCENES = {"voxel_island.scene": (64, 128, 8, 12), "voxel_quads.scene": (64, 128, 8, 12)}


def golden_voxel():
    """The voxel scenes through the reference's own .vox loader, palette
    and voxel_grid walk (or its faces, voxel_quads.scene): the hit subset
    with t and front face, the surface records (u, v, tangents, base colour:
    palette or outline material) and the converged means; and MagicaVoxel's
    default palette as the reference holds it (vox_default_palette), read
    back through a version-150 file without an RGBA chunk."""
    for name, render in VOXEL_SCENES.items():
        # the reference's tie count tests every face of the quads scene per ray:
        # the subset is drawn from a batch of 2^15
        golden_asset_scene(name, render, n_batch=1 << 15)
        path = os.path.join(HERE, name.replace(".scene", ".npz"))
        g = dict(np.load(path))
        r = ol.RefScene(scene_path(name))
        g["surface"] = r.surface(g["rays"], g["tt"])
        g["light_probabilities"] = r.light_probabilities()
        np.savez_compressed(path, **g)
        print(name, "surface", flush=True)
    # the default palette: straight down onto the row of voxels 1-255 of
    # plain150.vox (cell x spans world x in [x - 127, x - 126], y in [0, 1])
    r = ol.RefScene(scene_path("vox_palette.scene"))
    xs = np.arange(255) - 127 + 0.5
    rays = np.zeros((255, 6))
    rays[:, 0], rays[:, 1], rays[:, 2], rays[:, 4] = xs, 5.0, -0.5, -1.0
    tt = np.tile([0.001, np.inf], (255, 1))
    np.savez_compressed(os.path.join(HERE, "vox_palette.npz"), rays=rays, tt=tt, surface=r.surface(rays, tt),
                        trace_t=r.trace(rays, tt)["t"])
    print("vox_palette", flush=True)


def golden_gltf():
    """The glTF scenes through the reference's own load_gltf (cgltf) and
    material mapping: for gltf_props.scene the hit subset and each hit's
    surface record (u, v, tangents, base colour through the texture
    transform and scale, the normal after a scaled normal map, principled's
    metallic and alpha from the channels of the metallic-roughness map), and
    converged means; for gltf_glb.scene (a constant coverage, whose hit test
    draws a random number) the converged means only."""
    golden_asset_scene("gltf_props.scene", (64, 128, 8, 12), n_batch=1 << 16)
    path = os.path.join(HERE, "gltf_props.npz")
    g = dict(np.load(path))
    r = ol.RefScene(scene_path("gltf_props.scene"))
    g["surface"] = r.surface(g["rays"], g["tt"])
    g["light_probabilities"] = r.light_probabilities()
    np.savez_compressed(path, **g)
    R = ol.reference()
    d = m.describe(scene_path("gltf_glb.scene"))
    width, spp, runs, depth = 64, 128, 8, 8
    h = max(1, int(width / d.camera.aspect_ratio))
    r = ol.RefScene(scene_path("gltf_glb.scene"))
    imgs = []
    for _ in range(runs):
        img = np.zeros((h, width, 3))
        R.ref_render(r.h, width, spp, depth, os.cpu_count() or 8, ol._d(img))
        imgs.append(img)
    imgs = np.stack(imgs)
    np.savez_compressed(os.path.join(HERE, "gltf_glb.npz"), mean=imgs.mean(0),
                        stderr=imgs.std(0, ddof=1) / np.sqrt(runs), width=width, height=h, spp=spp * runs,
                        max_depth=depth)
    print("gltf", flush=True)


def golden_studio():
    """studio.scene through the reference's make_cyclorama / make_softbox:
    the hit subset, its surface records (the cyclorama's interpolated
    normals and arc-length uv) and converged means."""
    golden_asset_scene("studio.scene", (64, 128, 8, 8), n_batch=1 << 16)
    path = os.path.join(HERE, "studio.npz")
    g = dict(np.load(path))
    r = ol.RefScene(scene_path("studio.scene"))
    g["surface"] = r.surface(g["rays"], g["tt"])
    g["light_probabilities"] = r.light_probabilities()
    np.savez_compressed(path, **g)
    print("studio", flush=True)


def golden_alpha():
    for name, render in ALPHA_SCENES.items():
        golden_asset_scene(name, render)


def golden_fog():
    for name, render in FOG_SCENES.items():
        golden_asset_scene(name, render)


def golden_sun():
    for name, render in SUN_SCENES.items():
        golden_asset_scene(name, render)
        path = os.path.join(HERE, name.replace(".scene", ".npz"))
        g = dict(np.load(path))
        r = ol.RefScene(scene_path(name))
        g["light_probabilities"] = r.light_probabilities()
        if name == "sky_atmosphere.scene":
            dirs = np.random.default_rng(11).normal(size=(4096, 3))
            g["sky_dirs"] = dirs
            g["sky_radiance"] = r.env_radiance(dirs)
        np.savez_compressed(path, **g)
        print(name, "light probabilities", g["light_probabilities"], flush=True)


def golden_materials():
    for name, render in MATERIAL_SCENES.items():
        golden_asset_scene(name, render)


def golden_obj():
    for name, render in ASSET_SCENES.items():
        golden_asset_scene(name, render)


def golden_textures():
    """textured.scene as golden_asset_scene writes it, plus what each hit of
    the 4096-ray subset hands its material (the reference's rec.u, rec.v,
    dp/du, dp/dv, base texture value, shading normal after normal maps,
    principled metallic and alpha), and stbi_load's bytes of every texture
    file the scene names."""
    for name, render in TEXTURE_SCENES.items():
        golden_asset_scene(name, render)
        path = os.path.join(HERE, name.replace(".scene", ".npz"))
        g = dict(np.load(path))
        r = ol.RefScene(scene_path(name))
        g["surface"] = r.surface(g["rays"], g["tt"])
        for f in ("globe.jpg", "bumps_normal.png", "tiles_rough.png", "tiles_metal.png"):
            px, ch = ol.reference_image_u8(os.path.join(ROOT, "scenes", "models", f))
            g["image/" + f] = px
            g["channels/" + f] = np.array(ch)
        np.savez_compressed(path, **g)
        print(name, "surface and images", flush=True)


def golden_asset_scene(name, render, n_batch=None):
    """A scene of the repository's own assets through the reference's own
    loaders (load_obj/load_mtl, stb_image's .hdr decoder): the seeded
    1M-ray batch's subset with the reference's hit IDs, t and front face,
    digests over the 1M rays, and converged low-resolution means."""
    R = ol.reference()
    d = m.describe(scene_path(name))
    o = ol.OracleScene(d.desc)
    n_batch = n_batch or N_BATCH
    rays, tt = o.ray_batch(d.camera, BATCH_SEED, n_batch)
    r = ol.RefScene(scene_path(name))
    out = r.trace(rays, tt, threads=os.cpu_count() or 8)
    sel = np.linspace(0, n_batch - 1, N_SUBSET).astype(np.int64)
    digest = {"seed": BATCH_SEED, "n": n_batch,
              "sha256_id": hashlib.sha256(out["id"].astype("<i8").tobytes()).hexdigest(),
              "sha256_t": hashlib.sha256(out["t"].astype("<f8").tobytes()).hexdigest(),
              "sha256_front": hashlib.sha256(out["front"].astype("u1").tobytes()).hexdigest(),
              "ties": int((out["ties"] > 1).sum()), "misses": int((out["id"] == -1).sum()),
              "unidentified": int((out["id"] == -2).sum())}
    width, spp, runs, depth = render
    h = max(1, int(width / d.camera.aspect_ratio))
    imgs = []
    for _ in range(runs):
        img = np.zeros((h, width, 3))
        R.ref_render(r.h, width, spp, depth, os.cpu_count() or 8, ol._d(img))
        imgs.append(img)
    imgs = np.stack(imgs)
    np.savez_compressed(os.path.join(HERE, name.replace(".scene", ".npz")), rays=rays[sel], tt=tt[sel], index=sel,
                        id=out["id"][sel], t=out["t"][sel], front=out["front"][sel], ties=out["ties"][sel],
                        digest=json.dumps(digest), mean=imgs.mean(0), stderr=imgs.std(0, ddof=1) / np.sqrt(runs),
                        width=width, height=h, spp=spp * runs, max_depth=depth)
    print(name, digest, flush=True)


def main():
    if not ol.reference_available():
        sys.exit("oracle/_ref/libmpt_ref.so is not built (make -C oracle ref, where the reference is present)")
    R = ol.reference()
    steps = set(sys.argv[1:]) or {"ggx", "trace", "ties", "bsdf", "render", "denoise", "display", "obj",
                                  "textures", "materials", "sun", "fog", "alpha", "coat", "voxel", "gltf", "studio"}
    if "denoise" in steps:
        golden_denoise()
    if "display" in steps:
        golden_display()
    if "obj" in steps:
        golden_obj()
    if "textures" in steps:
        golden_textures()
    if "materials" in steps:
        golden_materials()
    if "sun" in steps:
        golden_sun()
    if "fog" in steps:
        golden_fog()
    if "alpha" in steps:
        golden_alpha()
    if "voxel" in steps:
        golden_voxel()
    if "gltf" in steps:
        golden_gltf()
    if "studio" in steps:
        golden_studio()
    if "ggx" in steps:
        golden_ggx(R)
    if "trace" in steps:
        golden_trace()
    if "ties" in steps:
        golden_ties()
    if "bsdf" in steps:
        golden_bsdf(R)
    if "coat" in steps:
        golden_coat(R)
    if "render" in steps:
        golden_render(R)


def golden_ggx(R):
    e = np.zeros(1024)
    ea = np.zeros(32)
    R.ref_ggx_table(ol._d(e), ol._d(ea))
    np.savez_compressed(os.path.join(HERE, "ggx_table.npz"), e=e.reshape(32, 32), e_avg=ea)
    print("ggx_table.npz", flush=True)


def golden_trace():
    digests = {}
    probs = {}
    for name in TRACE_SCENES:
        d = m.describe(scene_path(name))
        o = ol.OracleScene(d.desc)  # generates the seeded input batch only
        rays, tt = o.ray_batch(d.camera, BATCH_SEED, N_BATCH)
        r = ol.RefScene(scene_path(name))
        out = r.trace(rays, tt, threads=os.cpu_count() or 8)
        digests[name] = {
            "seed": BATCH_SEED, "n": N_BATCH,
            "sha256_id": hashlib.sha256(out["id"].astype("<i8").tobytes()).hexdigest(),
            "sha256_t": hashlib.sha256(out["t"].astype("<f8").tobytes()).hexdigest(),
            "sha256_front": hashlib.sha256(out["front"].astype("u1").tobytes()).hexdigest(),
            "ties": int((out["ties"] > 1).sum()), "misses": int((out["id"] == -1).sum()),
            "unidentified": int((out["id"] == -2).sum()),
        }
        sel = np.linspace(0, N_BATCH - 1, N_SUBSET).astype(np.int64)
        np.savez_compressed(os.path.join(HERE, "trace_" + name.replace(".scene", ".npz")),
                            rays=rays[sel], tt=tt[sel], index=sel, id=out["id"][sel], t=out["t"][sel],
                            front=out["front"][sel], ties=out["ties"][sel])
        probs[name] = r.light_probabilities().tolist()
        print(name, digests[name], flush=True)
    with open(os.path.join(HERE, "trace_digests.json"), "w") as f:
        json.dump(digests, f, indent=1)
    with open(os.path.join(HERE, "light_probs.json"), "w") as f:
        json.dump(probs, f, indent=1)


TIE_SCENES = ["cornell.scene", "test_lights.scene", "weekend_cover.scene"]
N_TIE_AIMED = 3072
N_TIE_BOUND = 1024


def _desc_arrays(desc):
    """Primitive and object arrays of a scene description as numpy."""
    n = desc.n_prims
    kind = np.ctypeslib.as_array(desc.prim_kind, (n,)).copy()
    data = np.ctypeslib.as_array(desc.prim_data, (n * 9,)).reshape(n, 9).copy()
    no = desc.n_objects
    okind = np.ctypeslib.as_array(desc.object_kind, (no,)).copy()
    first = np.ctypeslib.as_array(desc.object_first, (no,)).copy()
    count = np.ctypeslib.as_array(desc.object_count, (no,)).copy()
    oxf = np.ctypeslib.as_array(desc.object_xform, (no,)).copy()
    xfs = [(np.array(desc.xforms[k].m[:]).reshape(3, 3), np.array(desc.xforms[k].t[:])) for k in range(desc.n_xforms)]
    return kind, data, okind, first, count, oxf, xfs


def tie_targets(desc):
    """Points where primitives meet or end: corners and dyadic edge points of
    quads, the lowest point of spheres (where they rest on a floor), vertices
    and edge midpoints of triangles; an instanced object's points are mapped
    into the world by its transform."""
    kind, data, okind, first, count, oxf, xfs = _desc_arrays(desc)
    pts = []
    for o in range(len(okind)):
        local = []
        for p in range(first[o], first[o] + (1 if okind[o] == 0 else count[o])):
            q = data[p]
            if kind[p] == 1:  # quad: Q, u, v
                Q, u, v = q[0:3], q[3:6], q[6:9]
                for s in np.arange(17) / 16.0:
                    local += [Q + s * u, Q + s * v, Q + u + s * v, Q + v + s * u]
            elif kind[p] == 0:  # sphere: its lowest point
                local.append(q[0:3] - np.array([0.0, q[3], 0.0]))
            else:  # triangle
                a, b, c = q[0:3], q[3:6], q[6:9]
                local += [a, b, c, 0.5 * (a + b), 0.5 * (b + c), 0.5 * (a + c)]
        if oxf[o] >= 0:
            mm, t = xfs[oxf[o]]
            local = [mm @ x + t for x in local]
        pts += local
    return np.unique(np.array(pts), axis=0)


def tie_batch(name, R, rng):
    """Rays from dyadic origins (and the camera) aimed exactly at tie_targets,
    so that the planes of primitives meeting there give bitwise-equal t; and
    seeded-batch rays whose interval ends or starts exactly at the reference's
    hit t (inclusive quads, exclusive spheres and triangles)."""
    d = m.describe(scene_path(name))
    targets = tie_targets(d.desc)
    lo, hi = targets.min(0), targets.max(0)
    span = np.maximum(hi - lo, 1.0)
    idx = rng.integers(0, len(targets), N_TIE_AIMED)
    P = targets[idx]
    org = np.round((lo + (0.05 + 0.9 * rng.random((N_TIE_AIMED, 3))) * span) * 8) / 8  # dyadic interior points
    cam = np.array(d.camera.lookfrom[:])
    org[: N_TIE_AIMED // 4] = cam
    dirs = P - org
    keep = np.linalg.norm(dirs, axis=1) > 1e-9
    aimed = np.concatenate([org[keep], dirs[keep]], axis=1)
    aimed_tt = np.tile([0.001, np.inf], (aimed.shape[0], 1))
    o = ol.OracleScene(d.desc)
    rays, tt = o.ray_batch(d.camera, BATCH_SEED + 7, 4 * N_TIE_BOUND)
    first = R_trace(name, rays, tt)
    hit = np.nonzero(first["id"] >= 0)[0][: 2 * (N_TIE_BOUND // 2)]
    half = len(hit) // 2
    tmax_rays, tmax_tt = rays[hit[:half]], tt[hit[:half]].copy()
    tmax_tt[:, 1] = first["t"][hit[:half]]  # the interval ends at the hit
    tmin_rays, tmin_tt = rays[hit[half:]], tt[hit[half:]].copy()
    tmin_tt[:, 0] = first["t"][hit[half:]]  # the interval starts at the hit
    tmin_tt[:, 1] = np.inf
    all_rays = np.concatenate([aimed, tmax_rays, tmin_rays])
    all_tt = np.concatenate([aimed_tt, tmax_tt, tmin_tt])
    kind = np.concatenate([np.zeros(len(aimed), np.int8), np.ones(half, np.int8), np.full(len(hit) - half, 2, np.int8)])
    return all_rays, all_tt, kind


def R_trace(name, rays, tt):
    return ol.RefScene(scene_path(name)).trace(rays, tt, threads=os.cpu_count() or 8)


def golden_ties():
    """trace_ties.npz: per scene, rays aimed at where primitives meet plus
    rays whose interval bound is the hit t, with the reference's closest-hit
    t and front face, candidate count and smallest/largest candidate ID."""
    rng = np.random.default_rng(11)
    out = {}
    for name in TIE_SCENES:
        rays, tt, kind = tie_batch(name, None, rng)
        r = R_trace(name, rays, tt)
        key = name.replace(".scene", "")
        out.update({f"{key}_rays": rays, f"{key}_tt": tt, f"{key}_kind": kind, f"{key}_id": r["id"],
                    f"{key}_t": r["t"], f"{key}_front": r["front"], f"{key}_ties": r["ties"],
                    f"{key}_tie_min": r["tie_min"], f"{key}_tie_max": r["tie_max"]})
        print(name, "rays", len(rays), "ties", int((r["ties"] > 1).sum()), "misses", int((r["id"] == -1).sum()),
              flush=True)
    np.savez_compressed(os.path.join(HERE, "trace_ties.npz"), **out)


def golden_bsdf(R):
    spec, q, front = bsdf_queries()
    n = spec.shape[0]
    ev = np.zeros((n, 3))
    pdf = np.zeros(n)
    R.ref_bsdf_eval(ol._d(spec), ol._d(q), front.ctypes.data_as(ol.P(ol.C.c_uint8)), n, ol._d(ev), ol._d(pdf))
    np.savez_compressed(os.path.join(HERE, "bsdf.npz"), spec=spec, q=q, front=front, eval=ev, pdf=pdf)
    print("bsdf.npz", n, "queries", flush=True)


def golden_coat(R):
    """bsdf_coat.npz: principled::eval and its scatter pdf's value with a clear
    coat (material.h:326-552), through the reference's set_clearcoat."""
    spec, q, front = bsdf_queries(seed=9, materials=COAT_MATERIALS)
    n = spec.shape[0]
    ev = np.zeros((n, 3))
    pdf = np.zeros(n)
    f = R.ref_bsdf_eval_rows
    f.restype = ol.C.c_int
    f.argtypes = [ol.P(ol.C.c_double), ol.C.c_int, ol.P(ol.C.c_double), ol.P(ol.C.c_uint8), ol.C.c_int64,
                  ol.P(ol.C.c_double), ol.P(ol.C.c_double)]
    f(ol._d(spec), 17, ol._d(q), front.ctypes.data_as(ol.P(ol.C.c_uint8)), n, ol._d(ev), ol._d(pdf))
    np.savez_compressed(os.path.join(HERE, "bsdf_coat.npz"), spec=spec, q=q, front=front, eval=ev, pdf=pdf)
    print("bsdf_coat.npz", n, "queries", flush=True)


def golden_render(R):
    for name, (width, spp, runs, depth) in RENDERS.items():
        d = m.describe(scene_path(name))
        h = max(1, int(width / d.camera.aspect_ratio))
        r = ol.RefScene(scene_path(name))
        imgs = []
        for _ in range(runs):
            img = np.zeros((h, width, 3))
            R.ref_render(r.h, width, spp, depth, os.cpu_count() or 8, ol._d(img))
            imgs.append(img)
        imgs = np.stack(imgs)
        mean = imgs.mean(0)
        stderr = imgs.std(0, ddof=1) / np.sqrt(runs)
        np.savez_compressed(os.path.join(HERE, "render_" + name.replace(".scene", ".npz")),
                            mean=mean, stderr=stderr, width=width, height=h, spp=spp * runs, max_depth=depth)
        print(name, "render", mean.mean((0, 1)), flush=True)


if __name__ == "__main__":
    main()
