"""Seeded radiance images and display settings for the display pipeline
tests (postprocess.h, tonemap.h, color3.h, image_io.h).

Images: a sky-and-sun frame (a smooth gradient with a few pixels far above
the bloom threshold, as the sun and lamps are in a render), a heavy-tailed
random frame of odd size, one-pixel-wide strips (the pyramid stops at once)
and tiny frames. Settings cover each effect alone and together, every
tonemapper and signed exposures; radius 2.0 reaches the 10-level cap.
"""
import numpy as np

# (exposure, tonemap, bloom, threshold, radius, vignette); tonemap 0 none, 1 reinhard, 2 aces
SETTINGS = [
    (0.0, 0, 0.0, 1.0, 0.04, 0.0),
    (0.0, 0, 0.3, 1.0, 0.04, 0.0),
    (1.5, 2, 0.5, 0.5, 0.5, 0.3),
    (-1.25, 1, 0.0, 1.0, 0.04, 0.6),
    (0.0, 0, 1.0, 0.0, 2.0, 0.0),
    (0.7, 2, 0.25, 1.0, 0.04, 0.25),
]


def sky_sun(w=64, h=36, seed=3):
    rng = np.random.default_rng(seed)
    y = (np.arange(h) + 0.5) / h
    x = (np.arange(w) + 0.5) / w
    img = np.zeros((h, w, 3))
    img[..., 0] = 0.3 + 0.4 * y[:, None] + 0.05 * x[None, :]
    img[..., 1] = 0.4 + 0.3 * y[:, None]
    img[..., 2] = 0.9 - 0.5 * y[:, None]
    img *= 1 + 0.1 * rng.standard_normal((h, w, 1))
    for _ in range(5):
        i, j = rng.integers(0, w), rng.integers(0, h)
        img[j, i] = rng.uniform(5, 60, 3)
    return img


def heavy(w=37, h=23, seed=4):
    rng = np.random.default_rng(seed)
    return rng.lognormal(-1.0, 1.5, (h, w, 3))


def strip_h(seed=5):
    return np.random.default_rng(seed).lognormal(0.0, 1.0, (1, 40, 3))


def strip_v(seed=6):
    return np.random.default_rng(seed).lognormal(0.0, 1.0, (40, 1, 3))


def tiny(seed=7):
    return np.random.default_rng(seed).lognormal(0.5, 1.0, (2, 2, 3))


CASES = {"sky_sun": sky_sun, "heavy": heavy, "strip_h": strip_h, "strip_v": strip_v, "tiny": tiny}
