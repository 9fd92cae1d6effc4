"""White-furnace, energy and Beer-Lambert cases of the reference's
tests/furnace_test.cpp, as programmatic scenes for either renderer.

Each case is (name, builder, check) where builder returns (DescBuilder,
Camera) and check(mean, worst) -> (ok, message). The scenes, camera and
tolerances are those of furnace_test.cpp:31-330 (32x32 image, 64 spp,
256 with an environment, max depth 100, white background).
"""
import math

import numpy as np

from scenebuild import DescBuilder, camera

FURNACE_CENTER = (0, 0, -1)


def furnace_camera(spp=64):
    # camera defaults of camera.h:21-56, furnace_test.cpp:35-41
    return camera(32, 1.0, spp, 100, (0, 0, 0), (0, 0, -1), 90, background=(1, 1, 1))


def sphere_case(make_material, env=False):
    def build():
        b = DescBuilder()
        b.light_strategy = 2  # a plain hittable_list of lights (furnace_test.cpp:33)
        m = make_material(b)
        b.sphere(FURNACE_CENTER, 0.5, m)
        cam = furnace_camera(256 if env else 64)
        if env:
            b.environment(np.ones((8, 16, 3)))
        return b, cam
    return build


def medium_case(g):
    def build():
        b = DescBuilder()
        b.light_strategy = 2
        ph = b.phase((1, 1, 1), g)
        b.medium_sphere(FURNACE_CENTER, 0.5, 2.0, ph)
        return b, furnace_camera(64)
    return build


def slab_case(scaled):
    # furnace_test.cpp:157-188
    d = 0.8
    sigma_a = (0.2, 1.0, 3.0)

    def build():
        b = DescBuilder()
        b.light_strategy = 2
        glass = b.dielectric(1.0, 0.0, sigma_a)
        if scaled:
            x = b.xform(np.diag([2.0, 2.0, 2.0]), [0, 0, 0])
            b.box((-50, -50, -1 - d / 2), (50, 50, -1), glass, xform=x)
        else:
            b.box((-100, -100, -2 - d), (100, 100, -2), glass)
        cam = camera(8, 1.0, 4, 10, (0, 0, 0), (0, 0, -1), 0.1, background=(1, 1, 1))
        return b, cam
    expect = np.exp(-np.array(sigma_a) * d)
    return build, expect


def furnace_check(tol):
    def check(img):
        mean, worst = img.mean(), np.abs(img - 1).max()
        ok = abs(mean - 1) < tol and worst < 4 * tol
        return ok, f"mean {mean:.8f}, worst pixel off by {worst:.3e} (tolerance {tol})"
    return check


def bounded_check(img):
    mean = img.mean()
    return mean <= 1.005 and mean > 0.5, f"mean {mean:.6f} (bounded)"


FURNACE = [
    ("white lambertian", sphere_case(lambda b: b.lambertian((1, 1, 1))), furnace_check(1e-6)),
    ("perfect mirror", sphere_case(lambda b: b.metal((1, 1, 1), 0.0)), furnace_check(1e-6)),
    ("clear glass", sphere_case(lambda b: b.dielectric(1.5)), furnace_check(0.02)),
    ("clear glass, explicit zero absorption", sphere_case(lambda b: b.dielectric(1.5, 0.0, (0, 0, 0))),
     furnace_check(0.02)),
    ("white lambertian, environment map", sphere_case(lambda b: b.lambertian((1, 1, 1)), env=True),
     furnace_check(0.02)),
    ("white medium, g 0.0", medium_case(0.0), furnace_check(0.03)),
    ("white medium, g 0.7", medium_case(0.7), furnace_check(0.03)),
]

BOUNDED = [("absorbing glass", sphere_case(lambda b: b.dielectric(1.5, 0.0, (0.5, 0.2, 0.1))))]
for _r in (0.1, 0.5, 1.0):
    BOUNDED += [
        (f"GGX metal, roughness {_r}", sphere_case(lambda b, r=_r: b.metal((1, 1, 1), r))),
        (f"rough glass, roughness {_r}", sphere_case(lambda b, r=_r: b.dielectric(1.5, r))),
        (f"rough absorbing glass, roughness {_r}", sphere_case(lambda b, r=_r: b.dielectric(1.5, r, (0.5, 0.2, 0.1)))),
        (f"principled dielectric, roughness {_r}", sphere_case(lambda b, r=_r: b.principled((1, 1, 1), 0.0, r))),
        (f"principled metallic, roughness {_r}", sphere_case(lambda b, r=_r: b.principled((1, 1, 1), 1.0, r))),
    ]

# a clear coat never adds energy (material.h:334-341): it takes F_c into its own lobe and passes
# (1 - w F_c) twice, times the tint's transmittance, to the layers below
for _r in (0.05, 0.4):
    BOUNDED += [
        (f"principled metallic under a clear coat, coat roughness {_r}",
         sphere_case(lambda b, r=_r: b.principled((1, 1, 1), 1.0, 0.3, coat=1.0, coat_roughness=r))),
        (f"principled dielectric under a tinted clear coat, coat roughness {_r}",
         sphere_case(lambda b, r=_r: b.principled((1, 1, 1), 0.0, 0.6, coat=0.8, coat_roughness=r,
                                                   coat_tint=(0.6, 0.3, 0.2)))),
    ]

# furnace_test.cpp:237-248: white rough conductors keep all the energy
ENERGY = []
for _r in (0.5, 1.0):
    ENERGY += [
        (f"compensated GGX metal, roughness {_r}", lambda b, r=_r: b.metal((1, 1, 1), r)),
        (f"compensated principled metallic, roughness {_r}", lambda b, r=_r: b.principled((1, 1, 1), 1.0, r)),
    ]

SLABS = [("absorbing slab", *slab_case(False)), ("absorbing slab, scaled instance", *slab_case(True))]


def run_furnace(render, build):
    b, cam = build()
    return render(b, cam)


def object_energy(render, make_material):
    """furnace_test.cpp:101-106: sphere pixels' mean over the covered area."""
    img = run_furnace(render, sphere_case(make_material))
    black = run_furnace(render, sphere_case(lambda b: b.lambertian((0, 0, 0))))
    return (img.mean() - black.mean()) / (1 - black.mean())


def slab_worst(render, build, expect):
    img = run_furnace(render, build)
    return float(np.abs(img - expect[None, None, :]).max())


__all__ = ["FURNACE", "BOUNDED", "ENERGY", "SLABS", "run_furnace", "object_energy", "slab_worst",
           "bounded_check", "math"]
