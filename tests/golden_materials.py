"""Scene descriptions holding the materials of the golden BSDF vectors.

A spec row is: type (0 lambertian, 1 metal, 2 dielectric, 3 principled,
5 phase), r, g, b, roughness, ior, metallic, g, sigma_a (3), measured metal
(0 none, 1.. Al Ti Cu Au Ag Cr) - the parameters make_golden.py gave the
reference's material constructors. Rows of 17 add principled's clear coat:
weight (0 none), roughness, tint (3).
"""
METALS = ["Al", "Ti", "Cu", "Au", "Ag", "Cr"]
import numpy as np

from scenebuild import DescBuilder


def desc_for_specs(spec):
    b = DescBuilder()
    index = {}
    mats = np.zeros(spec.shape[0], dtype=np.int32)
    for i, row in enumerate(spec):
        key = tuple(row.tolist())
        if key not in index:
            t = int(row[0])
            rgb = tuple(row[1:4])
            if t == 0:
                mi = b.lambertian(rgb)
            elif t == 1:
                mi = b.metal(rgb, row[4], conductor=METALS[int(row[11]) - 1] if row[11] > 0 else None)
            elif t == 2:
                mi = b.dielectric(row[5], row[4], tuple(row[8:11]))
            elif t == 3:
                coat = {}
                if len(row) >= 17 and row[12] > 0:
                    coat = dict(coat=row[12], coat_roughness=row[13], coat_tint=tuple(row[14:17]))
                mi = b.principled(rgb, row[6], row[4], row[5],
                                  conductor=METALS[int(row[11]) - 1] if row[11] > 0 else None, **coat)
            else:
                mi = b.phase(rgb, row[7])
            index[key] = mi
        mats[i] = index[key]
    b.sphere((0, 0, -5), 1.0, 0)
    desc, keep = b.build()
    return desc, keep, mats
