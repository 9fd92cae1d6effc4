"""Writes the golden vectors in tests/golden from the compiled reference.

Runs only where oracle/_ref/libmpt_ref.so was built (the reference renderer
compiled from its own sources by oracle/Makefile). Every expected value
here is computed by the reference itself; inputs (ray batches, material
queries) are seeded and regenerated identically by the tests:

  ggx_table.npz         the 32x32 E(mu, alpha) table and E_avg (microfacet.h:265-338)
  trace_<scene>.npz     4096 rays of the seeded batch with the reference's
                        closest-hit ID, t, front_face and tie count (bvh.h:45-86)
  trace_digests.json    SHA-256 of the reference's IDs and t bits over the full
                        1M-ray batches, tie and miss counts
  light_probs.json      light_list::probability per light (light_list.h:45-127)
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

Usage: python tests/golden/make_golden.py [step ...]  (steps: ggx trace bsdf render denoise display obj textures materials; all by default)
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


def bsdf_queries(seed=5):
    """Seeded (normal, ray direction, wi, front) queries; the ray direction
    points against the normal as after set_face_normal (hittable.h:23-26)."""
    rng = np.random.default_rng(seed)
    n = len(BSDF_MATERIALS) * N_PER_MATERIAL

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
    spec = np.repeat(np.array(BSDF_MATERIALS, dtype=np.float64), N_PER_MATERIAL, axis=0)
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
}


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


def golden_asset_scene(name, render):
    """A scene of the repository's own assets through the reference's own
    loaders (load_obj/load_mtl, stb_image's .hdr decoder): the seeded
    1M-ray batch's subset with the reference's hit IDs, t and front face,
    digests over the 1M rays, and converged low-resolution means."""
    R = ol.reference()
    d = m.describe(scene_path(name))
    o = ol.OracleScene(d.desc)
    rays, tt = o.ray_batch(d.camera, BATCH_SEED, N_BATCH)
    r = ol.RefScene(scene_path(name))
    out = r.trace(rays, tt, threads=os.cpu_count() or 8)
    sel = np.linspace(0, N_BATCH - 1, N_SUBSET).astype(np.int64)
    digest = {"seed": BATCH_SEED, "n": N_BATCH,
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
    steps = set(sys.argv[1:]) or {"ggx", "trace", "bsdf", "render", "denoise", "display", "obj", "textures",
                                  "materials", "sun"}
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
    if "ggx" in steps:
        golden_ggx(R)
    if "trace" in steps:
        golden_trace()
    if "bsdf" in steps:
        golden_bsdf(R)
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


def golden_bsdf(R):
    spec, q, front = bsdf_queries()
    n = spec.shape[0]
    ev = np.zeros((n, 3))
    pdf = np.zeros(n)
    R.ref_bsdf_eval(ol._d(spec), ol._d(q), front.ctypes.data_as(ol.P(ol.C.c_uint8)), n, ol._d(ev), ol._d(pdf))
    np.savez_compressed(os.path.join(HERE, "bsdf.npz"), spec=spec, q=q, front=front, eval=ev, pdf=pdf)
    print("bsdf.npz", n, "queries", flush=True)


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
