"""Synthetic frames for the denoiser tests.

The first four are the known-answer cases of the reference's own denoiser
test (tests/denoise_test.cpp:50-125): a flat grey wall under noise, an
albedo edge, a normal edge, and a converged wall with one firefly and one
real one-pixel highlight. The noise comes from numpy here (the reference
draws it from std::mt19937); the checks are the same. `mixed` is a frame
with every guide varying (random albedo, including black, random unit
normals, random variance) for bit-level comparisons.
"""
import numpy as np

W = H = 64


def make(noise, seed, pixel):
    """pixel(i, j) -> (albedo, normal, radiance); Gaussian relative noise with
    the matching variance, as make_frame (denoise_test.cpp:22-39)."""
    rng = np.random.default_rng(seed)
    rad = np.zeros((H, W, 3))
    alb = np.zeros((H, W, 3))
    nrm = np.zeros((H, W, 3))
    var = np.zeros((H, W))
    truth = np.zeros((H, W, 3))
    for j in range(H):
        for i in range(W):
            a, n, r = pixel(i, j)
            r = np.asarray(r, float)
            lum = 0.2126 * r[0] + 0.7152 * r[1] + 0.0722 * r[2]
            sigma = noise * lum
            truth[j, i] = r
            rad[j, i] = r * (1 + noise * rng.standard_normal())
            alb[j, i] = a
            nrm[j, i] = n
            var[j, i] = sigma * sigma
    return dict(radiance=rad, albedo=alb, normal=nrm, variance=var, truth=truth)


def flat_wall():
    return make(0.3, 1, lambda i, j: ((0.5, 0.5, 0.5), (0, 0, 1), (0.5, 0.5, 0.5)))


def albedo_edge():
    def px(i, j):
        v = 0.2 if i < W // 2 else 0.8
        return (v, v, v), (0, 0, 1), (v, v, v)
    return make(0.3, 2, px)


def normal_edge():
    def px(i, j):
        n = (0, 0, 1) if i < W // 2 else (1, 0, 0)
        r = (1, 1, 1) if i < W // 2 else (0.1, 0.1, 0.1)
        return (0.7, 0.7, 0.7), n, r
    return make(0.3, 3, px)


FIREFLY = (20, 20)
HIGHLIGHT = (40, 40)


def firefly():
    f = make(0.0, 4, lambda i, j: ((0.5, 0.5, 0.5), (0, 0, 1), (0.5, 0.5, 0.5)))
    f["radiance"][FIREFLY] = 50
    f["variance"][FIREFLY] = 50 * 50
    f["radiance"][HIGHLIGHT] = 50
    f["variance"][HIGHLIGHT] = 5 * 5
    return f


def mixed(w=48, h=40, seed=9):
    rng = np.random.default_rng(seed)
    alb = rng.uniform(0, 1, (h, w, 3))
    alb[rng.uniform(size=(h, w)) < 0.1] = 0.0
    nrm = rng.standard_normal((h, w, 3))
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    nrm[rng.uniform(size=(h, w)) < 0.05] = 0.0
    rad = alb * rng.uniform(0, 2, (h, w, 1)) + rng.exponential(0.05, (h, w, 3))
    rad[rng.uniform(size=(h, w)) < 0.01] *= 40.0
    var = rng.exponential(0.02, (h, w))
    return dict(radiance=rad, albedo=alb, normal=nrm, variance=var)


CASES = {"flat_wall": flat_wall, "albedo_edge": albedo_edge, "normal_edge": normal_edge, "firefly": firefly,
         "mixed": mixed}
# (iterations, sigma_luminance, sigma_normal, sigma_albedo): the defaults and a non-default set
SETTINGS = [(5, 4.0, 0.1, 0.1), (3, 2.0, 0.25, 0.3)]


def rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum(-1).mean()))


def lum(c):
    return 0.2126 * c[..., 0] + 0.7152 * c[..., 1] + 0.0722 * c[..., 2]
